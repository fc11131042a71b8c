"""GPU: device 0D norms and the continuous-batching serving loop.

The serving test admits requests at different times under a memory budget
that holds only some of them at once, so members join and leave a running
batch; every request's final state must equal the same request evolved
alone (bit for bit: the same kernels run on the same work items)."""
import numpy as np
import pytest

from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


def test_norms0d_matches_downloaded_arrays(oaxis_mod):
    oaxis = oaxis_mod
    p = oaxis.params_from_text(CASES["scalar_pulse_gd2_levels"][0])
    with oaxis.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        st.evolve(2)
        g = st.level_info(0)["ghost"]
        for lev in range(st.nlevels):
            for name in ("alpha", "chi", "trK", "scalar_phi", "beta_r"):
                a = st.download(name, lev)
                # interior range of save0Ddata: i in [1, Nr-g], j in [1, Nz-g]
                sub = a[g:a.shape[0] - g, g:a.shape[1] - g]
                n = st.norms0d(name, lev)
                assert n["max"] == sub.max() and n["min"] == sub.min()
                assert n["nm1"] == np.abs(sub).max()
                ref = np.sqrt(np.sum(sub * sub) / sub.size)
                assert abs(n["nm2"] - ref) <= 1e-13 * ref, (name, lev, n["nm2"], ref)


def _requests(oaxis):
    from oaxis import workloads
    from oaxis.serve import Request
    pulse = workloads.scalar_pulse(seed=3, index=0) + "\nNrtotal = 32\nNztotal = 64\nNl0 = 1\ndr0 = 0.375\ndz0 = 0.375\n"
    pulse2 = workloads.scalar_pulse(seed=3, index=1) + "\nNrtotal = 48\nNztotal = 96\nNl0 = 2\ndr0 = 0.25\ndz0 = 0.25\n"
    schw = workloads.schwarzschild(1.3, 32, 64, 1)
    return [
        Request(0, pulse, 5, 0.0, "scalar_pulse"),
        Request(1, CASES["brill_gd2_odd_sizes_icn"][0], 3, 0.0, "brill"),
        Request(2, schw, 6, 0.0, "schwarzschild"),
        Request(3, pulse2, 4, 0.02, "scalar_pulse"),
        Request(4, CASES["schwarzschild_3levels_icn_eqsym"][0], 2, 0.05, "schwarzschild"),
        Request(5, CASES["gauge_pulse_gammadriver2"][0], 7, 0.05, "gauge"),
        # the deepest requests of the serving mix: 6 levels
        Request(6, workloads.schwarzschild(0.8, 24, 48, 5, extent=12.8), 3, 0.03, "schwarzschild"),
    ]


def _final_state(oaxis, st):
    return {(lev, name): st.download(name, lev)
            for lev in range(st.nlevels) for name in oaxis.state_fields(st.params)}


def test_serving_results_equal_alone(oaxis_mod):
    oaxis = oaxis_mod
    from oaxis.serve import Server
    reqs = _requests(oaxis)
    alone = {}
    for r in reqs:
        with oaxis.Spacetime(oaxis.params_from_text(r.text)) as st:
            st.initial_data()
            st.initial_aux()
            st.evolve(r.steps)
            alone[r.rid] = _final_state(oaxis, st)
    fps = [oaxis.footprint(oaxis.params_from_text(r.text)) for r in reqs]
    budget = sum(sorted(fps)[-3:])           # at most about three members at once
    got = {}
    peak = []

    def keep(req, st, rec):
        got[req.rid] = _final_state(oaxis, st)

    srv = Server(budget, check_every=0, on_retire=keep)
    orig_admit = srv._admit

    def admit(req, now, batch, active, used, prepared=None):
        peak.append(used)            # the footprints admitted so far, this request's included
        return orig_admit(req, now, batch, active, used, prepared)

    srv._admit = admit
    print("serving test: %d requests, budget %.3g B" % (len(reqs), budget), flush=True)
    records, summary = srv.run(reqs)
    assert summary["requests"] == len(reqs)
    assert max(peak) <= budget
    for rec in records:
        assert rec.reason == "final_time" and rec.steps == reqs[rec.rid].steps
        assert rec.latency >= 0.0 and rec.admitted >= rec.arrival
    assert summary["p99_s"] >= summary["p50_s"] > 0.0
    for r in reqs:
        a, b = got[r.rid], alone[r.rid]
        bad = [k for k in a if not np.array_equal(a[k], b[k])]
        assert not bad, "request %d (%s) differs in the mix: %s" % (r.rid, r.kind, bad[:6])


DISPERSING_PULSE = """
order = four
dtfac = 0.25
boundtype = radiative
integrator = rk4
eqsym = .false.
dr0 = 0.25
dz0 = 0.25
Nrtotal = 32
Nztotal = 64
Nl0 = 0
mattertype = scalar
idata = scalarpulse
scalarpulse_type = shell
scalar_a0 = 0.001
scalar_r0 = 0.0
scalar_z0 = 0.0
scalar_sr0 = 1.0
scalar_sz0 = 1.0
"""


def test_plain_serving_matches_batched_outcomes(oaxis_mod):
    """run_plain serves the same requests one spacetime at a time: every
    request ends with the same reason, step count and point updates as in
    the continuous batch."""
    from oaxis.serve import Server
    reqs = _requests(oaxis_mod)
    srv = Server(1e12, check_every=0)
    rec_b, sum_b = srv.run(reqs)
    rec_p, sum_p = srv.run_plain(reqs)
    print("serving: batched %.3g updates/s, plain %.3g updates/s" % (sum_b["updates_per_s"], sum_p["updates_per_s"]))
    for a, b in zip(rec_b, rec_p):
        assert (a.rid, a.reason, a.steps, a.point_updates) == (b.rid, b.reason, b.steps, b.point_updates)
    assert sum_b["point_updates"] == sum_p["point_updates"]


def test_serving_retires_a_dispersed_pulse(oaxis_mod):
    """A weak central scalar pulse spreads out and leaves through the
    radiative boundary: the request retires as "dispersed" at the first
    check (every check_every coarse steps) where max |scalar_phi| on level 0
    is below disperse_frac of its initial value (the oracle's evolution of
    this pulse crosses 0.1 at step 48, after a dip to 0.17 and a rebound to
    0.45).  The same pulse with a smaller step count retires at its final
    time first."""
    oaxis = oaxis_mod
    from oaxis.serve import Request, Server
    frac, every, cap = 0.1, 4, 400
    p = oaxis.params_from_text(DISPERSING_PULSE)
    # the step of the first check that sees the field dispersed, evolving alone
    with oaxis.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        phi0 = st.norms0d("scalar_phi", 0)["nm1"]
        expect = None
        for k in range(every, cap, every):
            st.evolve(every)
            if st.norms0d("scalar_phi", 0)["nm1"] < frac * phi0:
                expect = k
                break
    print("dispersing pulse: phi0 %.3e, dispersed at step %s" % (phi0, expect), flush=True)
    assert phi0 > 0.0 and expect == 48
    reqs = [Request(0, DISPERSING_PULSE, cap, 0.0, "scalar_pulse"),
            Request(1, DISPERSING_PULSE, expect - every // 2, 0.0, "scalar_pulse")]
    srv = Server(1e10, check_every=every, disperse_frac=frac)
    records, summary = srv.run(reqs)
    assert summary["requests"] == 2
    assert records[0].reason == "dispersed" and records[0].steps == expect, (records[0].reason, records[0].steps)
    assert records[1].reason == "final_time" and records[1].steps == expect - every // 2
