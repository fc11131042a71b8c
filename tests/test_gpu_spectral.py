"""GPU: the spectral apparent-horizon finder (ahmethod = spectral / both;
horizon_finder.f90 ahspectral :1113-1972) on the device against the oracle,
on the same states.

The device runs the reference's fast flow whole on the device, one
workgroup per search; the interpolated fields differ from the oracle's by
rounding only, so the found flag, the error flag, the shooting search that
precedes the flow (both, ahshootfirst) agree exactly, the iteration counts
within one (the flow's step control compares sums of coefficients against
thresholds), and the surface, area and mass to 1e-9 relative.  The flow is
run to ahepsilon_spectral = 1e-10 so that one iteration more or less moves
the surface by far less than that.  States: Schwarzschild (and a second
search starting from the horizon found), equatorial symmetry, a horizon off
the origin started from an ellipse, Schwarzschild through a coordinate
change with ahmethod = both, ahshootfirst, 25 modes (the tables read from
device memory), flat space, an evolved state on two levels, a batch of
searches and a begun search.
"""
import time

import numpy as np
import pytest

from test_gpu_horizon import EVOLVED, device_from_oracle
from test_oracle_horizon import disguised_schwarzschild, oracle_with, schwarzschild_fields

pytestmark = pytest.mark.gpu

EXACT = ("found", "errorflag", "method", "ntheta", "shots", "zcenter", "zinit")
TOL = 1e-9
TIGHT = "ahepsilon_spectral = 1e-10\n"


def compare(dev, ref, label):
    for key in EXACT:
        assert dev[key] == ref[key], (label, key, dev[key], ref[key])
    assert abs(dev["iterations"] - ref["iterations"]) <= 1, (label, dev["iterations"], ref["iterations"])
    scale = max(abs(ref["area"]), 1.0)
    assert abs(dev["area"] - ref["area"]) <= TOL * scale, (label, dev["area"], ref["area"])
    assert abs(dev["mass"] - ref["mass"]) <= TOL * max(abs(ref["mass"]), 1.0), label
    n = ref["ntheta"] + 1
    assert dev["theta"].shape == (n,) and dev["rad"].shape == (n,)
    assert np.abs(dev["theta"] - ref["theta"]).max() <= 1e-15, label
    err = np.abs(dev["rad"] - ref["rad"]).max()
    assert err <= TOL * max(np.abs(ref["rad"]).max(), 1.0), (label, err)


def pair(oaxis, oracle, h, fields, method="spectral", k=0, extra="", eqsym=False):
    orc = oracle_with(oaxis, oracle, h, fields, extra="ahmethod = %s\n" % method + extra, eqsym=eqsym)
    st = device_from_oracle(oaxis, orc, orc._params)
    st.initial_aux()
    return st, orc


def both_find(st, orc, k, label):
    t0 = time.perf_counter()
    dev = st.ahfind(k, shape=True)
    t1 = time.perf_counter()
    ref = orc.ahfind(k, shape=True)
    t2 = time.perf_counter()
    print("%s: found %d after %d iterations (oracle %d), %d shots, mass %.10f; device %.3f s, oracle %.3f s"
          % (label, dev["found"], dev["iterations"], ref["iterations"], dev["shots"], dev["mass"], t1 - t0, t2 - t1),
          flush=True)
    compare(dev, ref, label)
    return dev, ref


def test_schwarzschild_and_second_search(oaxis_mod, oracle_mod):
    st, orc = pair(oaxis_mod, oracle_mod, 0.1, schwarzschild_fields(2.0), extra=TIGHT)
    dev, _ = both_find(st, orc, 0, "schwarzschild")
    assert dev["found"] == 1 and abs(dev["mass"] - 2.0) < 2e-4 and dev["iterations"] > 50
    again, _ = both_find(st, orc, 0, "second search")
    assert again["iterations"] <= 2
    st.close()
    orc.close()


def test_equatorial_symmetry(oaxis_mod, oracle_mod):
    st, orc = pair(oaxis_mod, oracle_mod, 0.1, schwarzschild_fields(2.0), extra=TIGHT, eqsym=True)
    dev, _ = both_find(st, orc, 0, "eqsym")
    assert dev["found"] == 1 and abs(dev["theta"][-1] - np.pi / 2) < 1e-14
    st.close()
    orc.close()


def test_off_centre_from_an_ellipse(oaxis_mod, oracle_mod):
    extra = TIGHT + "N_ah = 3\nz_ah1 = 1.2\nz_ah2 = -1.2\naheccentric = 0.3\n"
    st, orc = pair(oaxis_mod, oracle_mod, 0.05, schwarzschild_fields(1.0, 1.2), k=1, extra=extra)
    dev, _ = both_find(st, orc, 1, "off centre, ellipse")
    assert dev["found"] == 1 and dev["zcenter"] == 1.2
    st.close()
    orc.close()


def test_both_methods_on_a_disguised_horizon(oaxis_mod, oracle_mod):
    fields, zstar = disguised_schwarzschild(2)
    st, orc = pair(oaxis_mod, oracle_mod, 0.1, fields, method="both", extra=TIGHT)
    dev, _ = both_find(st, orc, 0, "both, disguised")
    assert dev["found"] == 1 and dev["shots"] > 10 and abs(dev["rad"][0] - zstar) < 1e-4
    st.close()
    orc.close()


def test_shoot_first(oaxis_mod, oracle_mod):
    st, orc = pair(oaxis_mod, oracle_mod, 0.1, schwarzschild_fields(2.0), extra=TIGHT + "ahshootfirst = .true.\n")
    first, _ = both_find(st, orc, 0, "shoot first")
    second, _ = both_find(st, orc, 0, "after the first call")
    assert first["shots"] > 10 and second["shots"] == 0 and second["found"] == 1
    st.close()
    orc.close()


def test_many_modes(oaxis_mod, oracle_mod):
    """ahNmodes = 25 (Ntheta = 250): the cos / sin tables (104 KB) no longer
    fit the kernel's LDS and are read from device memory, and the surface
    takes two chunks of the source evaluation."""
    st, orc = pair(oaxis_mod, oracle_mod, 0.1, schwarzschild_fields(2.0), extra=TIGHT + "ahNmodes = 25\n")
    dev, _ = both_find(st, orc, 0, "25 modes")
    assert dev["found"] == 1 and dev["ntheta"] == 250
    st.close()
    orc.close()


def test_flat_space(oaxis_mod, oracle_mod):
    st, orc = pair(oaxis_mod, oracle_mod, 0.1, {})
    dev, _ = both_find(st, orc, 0, "flat")
    assert dev["found"] == 0 and dev["errorflag"] == 1 and np.all(dev["rad"] == 0.0)
    st.close()
    orc.close()


def test_evolved_two_levels(oaxis_mod, oracle_mod):
    p = oaxis_mod.params_from_text(EVOLVED + "ahmethod = spectral\n" + TIGHT)
    orc = oracle_mod.Oracle(p)
    orc.initial_data()
    st = device_from_oracle(oaxis_mod, orc, p)
    orc.initial_aux()
    st.initial_aux()
    for _ in range(3):
        orc.coarse_step()
        st.coarse_step()
    dev, _ = both_find(st, orc, 0, "evolved 3 steps")
    assert dev["found"] == 1
    st.close()
    orc.close()


def _text(M, extra=""):
    return EVOLVED.replace("BH1mass = 2.0", "BH1mass = %r" % M) + extra


def test_batch_and_begun_searches_equal_alone(oaxis_mod):
    """oa_batch_ahfind runs the flows of several members in one launch (one
    workgroup each) beside a shooting member; each result equals that
    member's search alone.  A begun search with a coarse step enqueued
    before its end gives what a direct search gives."""
    oaxis = oaxis_mod
    texts = [_text(2.0, "ahmethod = spectral\n"), _text(1.6, "Nl0 = 0\nahmethod = both\n"),
             _text(2.0, "N_ah = 2\nz_ah1 = 0.0\nahmethod = spectral\n"), _text(2.0)]
    ks = [0, 0, 1, 0]
    keys = EXACT + ("iterations", "residual", "area", "mass")
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
    print("batch of %d searches: %.3f s" % (len(got), time.perf_counter() - t0), flush=True)
    for q, (g, a) in enumerate(zip(got, alone)):
        for key in keys:
            assert g[key] == a[key], (q, key, g[key], a[key])
    assert [g["found"] for g in got] == [1, 1, 1, 1] and [g["method"] for g in got] == [1, 2, 1, 0]
    batch.close()
    for st in members:
        st.close()

    p = oaxis.params_from_text(texts[0])
    with oaxis.Spacetime(p) as a, oaxis.Spacetime(p) as b:
        for st in (a, b):
            st.initial_data()
            st.initial_aux()
            st.evolve(2)
        direct = a.ahfind(0, shape=True)
        search = b.ahfind_begin(0)
        b.coarse_step()
        begun = search.end()
        for key in keys:
            assert begun[key] == direct[key], (key, begun[key], direct[key])
        shape = b.ahfind(0, shape=True)          # the next search starts from ack
        assert shape["found"] == 1 and shape["iterations"] < direct["iterations"]
