"""GPU: the piecewise drop-in path.  A maintainer who keeps the reference's
onestep recursion (onestep.f90:97-463) and replaces only its body calls
oa_saveold, then per internal iteration oa_substep, oa_finebound (level > 0)
and oa_auxiliary, then oa_advance_time, the recursion into level+1 twice,
and, when the fine level's step count is even, oa_restrict and
oa_auxiliary(level-1).  Driven that way from Python, every evolved array of
every level must equal oa_coarse_step bit for bit (same kernels on the same
work items), the level clocks and point-update counts must be equal, and the
result must agree with the oracle to 1e-10.  Cases: RK4 and ICN on 2 and 3
levels, with and without equatorial symmetry, the Gamma-driver and the
scalar field."""
import numpy as np
import pytest

from parity_util import TOL, compare_state
from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


def weights(integrator_rk4, niter, it, dt):
    """dtw and the RK4 weight of internal iteration it (onestep.f90:141-177)."""
    if integrator_rk4:
        return {1: (0.5 * dt, 1 / 6), 2: (0.5 * dt, 1 / 3), 3: (dt, 1 / 3), 4: (dt, 1 / 6)}[it]
    return (0.5 * dt if it < niter else dt), 0.0


def onestep(oaxis, st, p, level):
    rk4 = p.integrator == oaxis.ENUMS["OA_RK4"]
    niter = 4 if rk4 else p.icniter
    dt = st.level_info(level)["dt"]
    st.saveold(level)
    for it in range(1, niter + 1):
        dtw, w = weights(rk4, niter, it, dt)
        st.substep(level, it, niter, dtw, w)
        if level > 0:
            st.finebound(level, dtw)
        st.auxiliary(level)
    st.advance_time(level)
    if level < st.nlevels - 1:
        onestep(oaxis, st, p, level + 1)
        onestep(oaxis, st, p, level + 1)
    if level > 0 and st.time(level)[3] % 2 == 0:
        st.restrict(level)
        st.auxiliary(level - 1)


PIECEWISE_CASES = ["schwarzschild_2levels_rk4", "schwarzschild_3levels_icn_eqsym", "scalar_pulse_gd2_levels",
                   "brill_gd2_odd_sizes_icn"]


@pytest.mark.parametrize("case", PIECEWISE_CASES)
def test_piecewise_steps_equal_coarse_steps(oaxis_mod, oracle_mod, case):
    oaxis = oaxis_mod
    text, nsteps = CASES[case]
    p = oaxis.params_from_text(text)
    orc = oracle_mod.Oracle(p)
    orc.initial_data()
    sts = [oaxis.Spacetime(p), oaxis.Spacetime(p)]
    try:
        for st in sts:
            for lev in range(st.nlevels):
                for name in oaxis.state_fields(p):
                    st.upload(name, orc.get(name, lev), lev)
            st.initial_aux()
        orc.initial_aux()
        whole, piece = sts
        for _ in range(nsteps):
            whole.coarse_step()
            onestep(oaxis, piece, p, 0)
            orc.coarse_step()
        print("case %s: %d levels, %d coarse steps" % (case, whole.nlevels, nsteps), flush=True)
        for lev in range(whole.nlevels):
            assert whole.time(lev) == piece.time(lev) == orc.time(lev), lev
            for name in oaxis.state_fields(p):
                a, b = whole.download(name, lev), piece.download(name, lev)
                assert np.array_equal(a, b), (lev, name, np.abs(a - b).max())
        assert whole.point_updates() == piece.point_updates()
        errs = compare_state(oaxis, piece, orc, p)
        worst = max(errs.items(), key=lambda kv: kv[1])
        print("  piecewise vs oracle: worst %s" % (worst,), flush=True)
        assert worst[1] <= TOL
    finally:
        for st in sts:
            st.close()
        orc.close()
