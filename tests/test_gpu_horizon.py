"""GPU: the apparent-horizon finder (horizon_finder.f90, shooting method) on
the device against the oracle, on the same states.

The device walks the reference's sequence of shots and decisions, so the
number of integrations, the axis position where the search stops, the
found flag and the accepted steps of the last integration agree exactly;
the area and the shape of the last integration agree to 1e-10 relative
(the interpolated fields differ from the oracle's by rounding only).
States: Schwarzschild on one and two levels (initial data and after
coarse steps of 1+log slicing), equatorial symmetry, horizons centred off
the origin, Schwarzschild through a non-trivial coordinate change,
Painleve-Gullstrand (curvature-only horizon) and flat space (none).
"""
import time

import numpy as np
import pytest

from test_oracle_horizon import (disguised_schwarzschild, oracle_with, painleve_gullstrand,
                                 schwarzschild_fields)

pytestmark = pytest.mark.gpu

EXACT = ("found", "shots", "errorflag", "ntheta", "zinit", "zcenter")
TOL = 1e-10


def device_from_oracle(oaxis, orc, p):
    st = oaxis.Spacetime(p)
    for lev in range(st.nlevels):
        for name in oaxis.state_fields(p):
            st.upload(name, orc.get(name, lev), lev)
    return st


def compare(dev, ref, label):
    for key in EXACT:
        assert dev[key] == ref[key], (label, key, dev[key], ref[key])
    scale = max(abs(ref["area"]), 1.0)
    assert abs(dev["area"] - ref["area"]) <= TOL * scale, (label, dev["area"], ref["area"])
    assert abs(dev["mass"] - ref["mass"]) <= TOL * max(abs(ref["mass"]), 1.0), label
    assert abs(dev["dradlast"] - ref["dradlast"]) <= 1e-8 * max(abs(ref["dradlast"]), 1.0), label
    n = ref["ntheta"] + 1
    assert dev["theta"].shape == (n,) and dev["rad"].shape == (n,)
    assert np.abs(dev["theta"] - ref["theta"]).max() <= TOL, label
    assert np.abs(dev["rad"] - ref["rad"]).max() <= TOL * max(np.abs(ref["rad"]).max(), 1.0), label


def synthetic(oaxis, oracle, h, fields, k=0, extra="", eqsym=False):
    orc = oracle_with(oaxis, oracle, h, fields, extra=extra, eqsym=eqsym)
    p = orc._params
    st = device_from_oracle(oaxis, orc, p)
    st.initial_aux()
    t0 = time.perf_counter()
    dev = st.ahfind(k, shape=True)
    t1 = time.perf_counter()
    ref = orc.ahfind(k, shape=True)
    t2 = time.perf_counter()
    st.close()
    orc.close()
    print("horizon k=%d: found %d after %d shots, mass %.10f; device %.3f s, oracle %.3f s"
          % (k, dev["found"], dev["shots"], dev["mass"], t1 - t0, t2 - t1))
    return dev, ref


def test_schwarzschild_matches_oracle(oaxis_mod, oracle_mod):
    dev, ref = synthetic(oaxis_mod, oracle_mod, 0.05, schwarzschild_fields(2.0))
    assert dev["found"] == 1 and abs(dev["mass"] - 2.0) < 2e-4
    compare(dev, ref, "schwarzschild")


def test_equatorial_symmetry_matches_oracle(oaxis_mod, oracle_mod):
    dev, ref = synthetic(oaxis_mod, oracle_mod, 0.05, schwarzschild_fields(2.0), eqsym=True)
    assert dev["found"] == 1
    compare(dev, ref, "eqsym")


@pytest.mark.parametrize("k,z0", [(1, 1.2), (2, -1.2)])
def test_off_centre_horizon_matches_oracle(oaxis_mod, oracle_mod, k, z0):
    extra = "N_ah = 3\nz_ah1 = 1.2\nz_ah2 = -1.2\n"
    dev, ref = synthetic(oaxis_mod, oracle_mod, 0.05, schwarzschild_fields(1.0, z0), k=k, extra=extra)
    assert dev["found"] == 1 and dev["zcenter"] == z0
    compare(dev, ref, "off-centre %d" % k)


def test_disguised_schwarzschild_matches_oracle(oaxis_mod, oracle_mod):
    fields, zstar = disguised_schwarzschild(2)
    dev, ref = synthetic(oaxis_mod, oracle_mod, 0.05, fields, extra="ahzinit = 1.3\n")
    assert dev["found"] == 1 and abs(dev["zinit"] - zstar) < 1e-4
    compare(dev, ref, "disguised")


def test_painleve_gullstrand_matches_oracle(oaxis_mod, oracle_mod):
    dev, ref = synthetic(oaxis_mod, oracle_mod, 0.1, painleve_gullstrand(0.5))
    assert dev["found"] == 1 and abs(dev["mass"] - 0.5) < 1e-4
    compare(dev, ref, "painleve-gullstrand")


def test_flat_space_no_horizon(oaxis_mod, oracle_mod):
    dev, ref = synthetic(oaxis_mod, oracle_mod, 0.1, {})
    assert dev["found"] == 0 and dev["area"] == 0.0
    compare(dev, ref, "flat")


EVOLVED = """
order = four
dtfac = 0.25
boundtype = radiative
geodiss = 0.1
idata = schwarzschild
BH1mass = 2.0
ilapse = psiminus2
slicing = 1+log
gauge_f = 2.0
integrator = rk4
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 32
Nztotal = 64
Nl0 = 1
"""


@pytest.mark.parametrize("steps", [0, 3])
def test_evolved_two_levels_match_oracle(oaxis_mod, oracle_mod, steps):
    p = oaxis_mod.params_from_text(EVOLVED)
    orc = oracle_mod.Oracle(p)
    orc.initial_data()
    st = device_from_oracle(oaxis_mod, orc, p)
    orc.initial_aux()
    st.initial_aux()
    for _ in range(steps):
        orc.coarse_step()
        st.coarse_step()
    dev = st.ahfind(0, shape=True)
    ref = orc.ahfind(0, shape=True)
    print("evolved %d steps: found %d, %d shots, mass %.10f" % (steps, dev["found"], dev["shots"], dev["mass"]))
    assert dev["found"] == 1
    compare(dev, ref, "evolved %d" % steps)
    # a second search on the same state repeats the first
    again = st.ahfind(0, shape=True)
    for key in EXACT + ("area",):
        assert again[key] == dev[key]
    st.close()
    orc.close()


def _schwarzschild_text(M, extra=""):
    return EVOLVED.replace("BH1mass = 2.0", "BH1mass = %r" % M) + extra


def test_batch_search_equals_alone(oaxis_mod):
    """oa_batch_ahfind runs the searches of several members in lockstep; each
    result equals that member's search alone (separate handles, same state)."""
    oaxis = oaxis_mod
    texts = [_schwarzschild_text(2.0), _schwarzschild_text(1.6, "Nl0 = 0\n"),
             _schwarzschild_text(2.0, "N_ah = 2\nz_ah1 = 0.0\n"), CASES_FLAT]
    ks = [0, 0, 1, 0]
    alone, members = [], []
    for text, k in zip(texts, ks):
        p = oaxis.params_from_text(text)
        with oaxis.Spacetime(p) as st:
            st.initial_data()
            st.initial_aux()
            st.evolve(2)
            alone.append(st.ahfind(k, shape=True))
        st = oaxis.Spacetime(p)
        st.initial_data()
        st.initial_aux()
        members.append(st)
    batch = oaxis.Batch(members)
    batch.evolve(2)
    t0 = time.perf_counter()
    got = batch.ahfind(list(zip(members, ks)))
    t1 = time.perf_counter()
    print("batch of %d searches: %.3f s" % (len(got), t1 - t0))
    for q, (g, a) in enumerate(zip(got, alone)):
        for key in EXACT + ("area", "mass", "dradlast"):
            assert g[key] == a[key], (q, key, g[key], a[key])
    assert [g["found"] for g in got] == [1, 1, 1, 0]
    batch.close()
    for st in members:
        st.close()


CASES_FLAT = """
order = four
dtfac = 0.25
boundtype = radiative
geodiss = 0.1
idata = minkowski
slicing = 1+log
gauge_f = 2.0
integrator = rk4
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 32
Nztotal = 64
"""


@pytest.mark.parametrize("method", ["shoot", "spectral"])
def test_serving_retires_on_horizon(oaxis_mod, method):
    """With ahafter > 0 the search at t = 0 is skipped, so the puncture's
    horizon counts as formed at the first search after ahafter: the search
    runs beside the next coarse step and the request retires after it with
    reason "horizon"; with ahafter = 0 the horizon exists from t = 0 and the
    request runs to its last step.  Either finding method."""
    oaxis = oaxis_mod
    from oaxis.serve import Request, Server
    m = "ahmethod = %s\n" % method
    late = _schwarzschild_text(2.0, "ahfind = .true.\nahfind_every = 2\nahafter = 0.1\n" + m)
    early = _schwarzschild_text(2.0, "ahfind = .true.\nahfind_every = 2\n" + m)
    reqs = [Request(0, late, 12, 0.0, "late"), Request(1, early, 6, 0.0, "early")]
    dt0 = oaxis.params_from_text(late).dtfac * 0.1
    records, summary = Server(1e12, check_every=0).run(reqs)
    by = {r.rid: r for r in records}
    first = next(s for s in range(2, 13, 2) if s * dt0 >= 0.1)
    assert by[0].reason == "horizon" and by[0].steps == first + 1, (by[0].reason, by[0].steps, first)
    assert by[0].horizon["found"] == 1 and abs(by[0].horizon["mass"] - 2.0) < 1e-3
    assert by[1].reason == "final_time" and by[1].steps == 6


COLLAPSE = """
mattertype = scalar
scalarpotential = none
idata = scalarpulse
scalarpulse_type = shell
scalar_a0 = 0.6
scalar_r0 = 0.0
scalar_z0 = 0.0
scalar_sr0 = 1.0
scalar_sz0 = 1.0
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
drivercsi = 0.75
integrator = rk4
order = four
eqsym = .true.
Nrtotal = 32
Nztotal = 32
Nl0 = 1
dr0 = 0.125
dz0 = 0.125
dtfac = 0.25
geodiss = 0.1
scalardiss = 0.1
boundtype = radiative
"""


def test_scalar_collapse_forms_a_horizon(oaxis_mod, oracle_mod):
    """A strong scalar pulse (Gamma-driver, 2 levels) collapses: searching
    every 2 coarse steps, device and oracle agree on every search and find
    the horizon at the same step; the serving loop, whose search runs beside
    the next coarse step, retires the request after that step with reason
    "horizon" and the search's result."""
    oaxis = oaxis_mod
    p = oaxis.params_from_text(COLLAPSE)
    orc = oracle_mod.Oracle(p)
    orc.initial_data()
    st = device_from_oracle(oaxis, orc, p)
    orc.initial_aux()
    st.initial_aux()
    formed = None
    for step in range(1, 49):
        orc.coarse_step()
        st.coarse_step()
        if step % 2:
            continue
        dev = st.ahfind(0, shape=True)
        ref = orc.ahfind(0, shape=True)
        compare(dev, ref, "collapse step %d" % step)
        if dev["found"]:
            formed = (step, dev["mass"])
            break
    st.close()
    orc.close()
    print("collapse: horizon at coarse step %s" % (formed,))
    # the elliptic solve makes psi about 2 at the centre: the horizon forms
    # at about step 36 with mass about 1.8
    assert formed is not None and 4 <= formed[0] <= 46 and 0.1 < formed[1] < 3.0
    from oaxis.serve import Request, Server
    records, _ = Server(1e12, check_every=0).run(
        [Request(0, COLLAPSE + "ahfind = .true.\nahfind_every = 2\n", 60, 0.0, "collapse")])
    assert records[0].reason == "horizon" and records[0].steps == formed[0] + 1
    assert abs(records[0].horizon["mass"] - formed[1]) < 1e-9


def test_critical_collapse_bisection(oaxis_mod):
    """Two rounds of amplitude bisection of the collapsing pulse family
    through the serving loop: outcomes are "no horizon" below and "horizon"
    above a threshold, and the bracket shrinks by members - 1 per round."""
    from oaxis.bisection import critical_bisection
    # with psi from the elliptic solve, pulses above about 0.65 are trapped
    # from the start (a horizon at t = 0 does not count as formed)
    res = critical_bisection(lambda a: COLLAPSE.replace("scalar_a0 = 0.6", "scalar_a0 = %r" % a),
                             0.2, 0.6, members=7, rounds=2, steps=48, ahfind_every=2)
    r1, r2 = res["rounds"]
    print("bisection: %s -> [%.6f, %.6f], %.2f s" % ([(round(a, 3), h) for a, h in zip(r1["amplitudes"], r1["horizon"])],
                                                    res["lo"], res["hi"], res["wall_s"]))
    assert r1["horizon"] == sorted(r1["horizon"])                 # no horizon, then horizon
    assert not r1["horizon"][0] and r1["horizon"][-1]
    k = r1["horizon"].index(True)
    assert r2["amplitudes"][0] == r1["amplitudes"][k - 1] and r2["amplitudes"][-1] == r1["amplitudes"][k]
    assert abs((res["hi"] - res["lo"]) - 0.4 / 36) < 1e-12
    assert res["lo"] < 0.6 and res["hi"] > 0.3


@pytest.mark.parametrize("k", [1, 2])
def test_moving_centres_match_oracle(oaxis_mod, oracle_mod, k):
    """ahmove on Brill-Lindquist punctures at z = +-3: the centre walk (Dz_phi
    interpolated on the device) and the search agree with the oracle."""
    from test_oracle_horizon import BASE, BINARY
    h = 0.05
    n = int(round(4.8 / h))
    p = oaxis_mod.params_from_text(BASE % (".false.", h, h, n, 2 * n) + BINARY)
    orc = oracle_mod.Oracle(p)
    orc.initial_data()
    st = device_from_oracle(oaxis_mod, orc, p)
    orc.initial_aux()
    st.initial_aux()
    dev = st.ahfind(k, shape=True)
    ref = orc.ahfind(k, shape=True)
    st.close()
    orc.close()
    assert dev["found"] == 1 and abs(abs(dev["zcenter"]) - 3.0) < h
    compare(dev, ref, "ahmove k=%d" % k)


def test_begun_search_matches_a_direct_search(oaxis_mod):
    """oa_ahfind_begin / oa_ahfind_end with a coarse step enqueued between
    them give what oa_ahfind gives on the state at the begin: the search
    reads only what the begin formed, and its shots run beside the step.
    A spacetime with a begun search cannot begin another before it ends."""
    oaxis = oaxis_mod
    text = _schwarzschild_text(2.0, "")
    p = oaxis.params_from_text(text)
    with oaxis.Spacetime(p) as a, oaxis.Spacetime(p) as b:
        for st in (a, b):
            st.initial_data()
            st.initial_aux()
            st.evolve(2)
        direct = a.ahfind(0)
        search = b.ahfind_begin(0)
        with pytest.raises(oaxis.OaxisError):
            b.ahfind_begin(0)
        b.coarse_step()
        begun = search.end()
        for key in ("found", "shots", "ntheta", "zinit", "area", "mass"):
            assert begun[key] == direct[key], (key, begun[key], direct[key])
        # the step ran: the next search sees the evolved state
        a.coarse_step()
        assert b.ahfind(0)["zinit"] == a.ahfind(0)["zinit"]

