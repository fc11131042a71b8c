"""GPU: batched evolution equals evolution alone, bit for bit.

A batch advances all its members with one launch per kernel kind over their
(spacetime, level) work items.  Members here differ in grid size, level
count, integrator (RK4 and ICN-3 in the same launches), shift and matter, and
join the batch at different clocks.  Every evolved array of every level of
every member must equal the array of the same spacetime evolved alone with
oa_evolve (same kernels' arithmetic, so exact equality is required)."""
import numpy as np
import pytest

from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


def _pulse(oaxis, index, Nr, Nz, nl0):
    from oaxis import workloads
    text = workloads.scalar_pulse(seed=7, index=index)
    return oaxis.params_from_text(text + "\nNrtotal = %d\nNztotal = %d\nNl0 = %d\ndr0 = %r\ndz0 = %r\n"
                                  % (Nr, Nz, nl0, 12.0 / Nr, 12.0 / Nr))


def _members(oaxis):
    """(name, params) of a mixed set."""
    out = [
        ("pulse_32x64_3lev", _pulse(oaxis, 0, 32, 64, 2)),
        ("pulse_48x96_2lev", _pulse(oaxis, 1, 48, 96, 1)),
        ("schwarzschild_rk4", oaxis.params_from_text(CASES["schwarzschild_rk4_noeqsym"][0])),
        ("schwarzschild_3lev_icn_eqsym", oaxis.params_from_text(CASES["schwarzschild_3levels_icn_eqsym"][0])),
        ("gauge_pulse_gd2", oaxis.params_from_text(CASES["gauge_pulse_gammadriver2"][0])),
        ("brill_icn_odd", oaxis.params_from_text(CASES["brill_gd2_odd_sizes_icn"][0])),
    ]
    return out


def _fresh(oaxis, p):
    st = oaxis.Spacetime(p)
    st.initial_data()
    st.initial_aux()
    return st


def _state(oaxis, st):
    return {(lev, name): st.download(name, lev)
            for lev in range(st.nlevels) for name in oaxis.state_fields(st.params)}


def _assert_same(tag, a, b):
    bad = []
    for k in a:
        if not np.array_equal(a[k], b[k]):
            d = np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-300)
            bad.append((k, d))
    assert not bad, "%s: batched result differs from the spacetime alone: %s" % (tag, bad[:6])


def test_batch_equals_alone(oaxis_mod):
    oaxis = oaxis_mod
    members = _members(oaxis)
    nsteps = 3
    alone = {}
    for name, p in members:
        print("batch test: alone", name, flush=True)
        st = _fresh(oaxis, p)
        st.evolve(nsteps)
        alone[name] = _state(oaxis, st)
        st.close()
    sts = [(name, _fresh(oaxis, p)) for name, p in members]
    print("batch test: batched, %d members" % len(sts), flush=True)
    with oaxis.Batch([st for _, st in sts]) as b:
        b.evolve(nsteps)
        launches, items = b.launch_stats()
        assert launches > 0 and items > launches
        expect = sum(float(st.params.Nrtotal * st.params.Nztotal) * sum(2 ** l for l in range(st.nlevels))
                     for _, st in sts) * nsteps
        assert b.point_updates() == expect
        for name, st in sts:
            _assert_same(name, _state(oaxis, st), alone[name])
    for _, st in sts:
        st.close()


def test_batch_join_and_leave(oaxis_mod):
    """A member that joins late (clock 0 beside members at t > 0) and one
    that leaves and continues alone."""
    oaxis = oaxis_mod
    members = _members(oaxis)
    ref = {}
    for name, p in members[:3]:
        st = _fresh(oaxis, p)
        st.evolve(4)
        ref[name] = _state(oaxis, st)
        st.close()
    (na, pa), (nb, pb), (nc, pc) = members[:3]
    a, bb, c = _fresh(oaxis, pa), _fresh(oaxis, pb), _fresh(oaxis, pc)
    print("batch test: join/leave", flush=True)
    with oaxis.Batch([a, bb]) as b:
        b.evolve(2)
        # a spacetime stepped alone joins at its own clocks
        c.evolve(1)
        b.add(c)
        b.evolve(1)
        # b leaves and finishes alone
        b.remove(bb)
        with pytest.raises(oaxis.OaxisError):
            a.evolve(1)           # a member cannot step on its own
        bb.evolve(1)
        b.evolve(1)
        # c has now done 1 + 1 + 1 = 3 steps
        b.remove(c)
        c.evolve(1)
        assert len(b) == 1
    _assert_same(na, _state(oaxis, a), ref[na])
    _assert_same(nb, _state(oaxis, bb), ref[nb])
    _assert_same(nc, _state(oaxis, c), ref[nc])
    for st in (a, bb, c):
        st.close()


def test_graph_steps_equal_direct_launches(oaxis_mod):
    """Coarse steps replayed from captured HIP graphs (reused across steps,
    with fresh work items each step) equal steps launched kernel by kernel,
    for a spacetime alone and for a batch."""
    oaxis = oaxis_mod
    members = _members(oaxis)[:4]
    nsteps = 5
    direct = {}
    for name, p in members:
        st = _fresh(oaxis, p)
        st.set_graphs(False)
        st.evolve(nsteps)
        direct[name] = _state(oaxis, st)
        st.close()
    for name, p in members[:2]:
        st = _fresh(oaxis, p)
        st.set_graphs(True)
        for _ in range(nsteps):
            st.coarse_step()
        _assert_same("graph " + name, _state(oaxis, st), direct[name])
        st.close()
    sts = [(name, _fresh(oaxis, p)) for name, p in members]
    with oaxis.Batch([st for _, st in sts]) as b:
        b.set_graphs(True)
        b.evolve(2)
        b.set_graphs(False)
        b.evolve(1)
        b.set_graphs(True)
        b.evolve(2)
        for name, st in sts:
            _assert_same("batched graph " + name, _state(oaxis, st), direct[name])
    for _, st in sts:
        st.close()


def test_schedule_autotuner_keeps_results(oaxis_mod):
    """The schedule autotuner (graph or direct launches, XCD block numbering)
    settles within 16 steady-state steps and changes no result."""
    oaxis = oaxis_mod
    name, p = _members(oaxis)[0]
    nsteps = 20
    ref = _fresh(oaxis, p)
    ref.set_graphs(False)
    ref.evolve(nsteps)
    want = _state(oaxis, ref)
    ref.close()
    st = _fresh(oaxis, p)
    st.evolve(nsteps)
    sched = st.schedule()
    assert sched["tuned"], sched
    assert all(v > 0.0 for v in sched["step_ms"].values()), sched
    _assert_same("autotuned " + name, _state(oaxis, st), want)
    st.close()
