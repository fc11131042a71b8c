"""GPU: checkpoint and restart (checkpointsave / checkpointload).

A spacetime saved between coarse steps and loaded into a new handle with the
same parameters continues bit for bit as the original: every evolved array
of every level and the level clocks agree exactly after further steps.
Cases: two levels with RK4; the Gamma-driver with the scalar field on
levels (its pre-initialisation dtbeta planes matter right after the initial
auxiliary pass); a Brill wave with ICN.  A snapshot is refused by a handle
with other parameters.
"""
import numpy as np
import pytest

from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


def _final(oaxis, st):
    out = {(lev, name): st.download(name, lev)
           for lev in range(st.nlevels) for name in oaxis.state_fields(st.params)}
    for lev in range(st.nlevels):
        out[(lev, "clock")] = np.array(st.time(lev)[:3])
    return out


@pytest.mark.parametrize("case,before,after", [
    ("schwarzschild_2levels_rk4", 2, 2),
    ("scalar_pulse_gd2_levels", 0, 2),
    ("scalar_pulse_gd2_levels", 1, 2),
    ("brill_gd2_odd_sizes_icn", 1, 1),
])
def test_restart_continues_bit_for_bit(oaxis_mod, tmp_path, case, before, after):
    oaxis = oaxis_mod
    p = oaxis.params_from_text(CASES[case][0])
    with oaxis.Spacetime(p) as a:
        a.initial_data()
        a.initial_aux()
        if before:
            a.evolve(before)
        path = str(tmp_path / "ckpt.bin")
        oaxis.save_checkpoint(a, path)
        a.evolve(after)
        want = _final(oaxis, a)
    with oaxis.load_checkpoint(path) as b:
        b.evolve(after)
        got = _final(oaxis, b)
    bad = [k for k in want if not np.array_equal(want[k], got[k])]
    assert not bad, "restart differs: %s" % bad[:6]


def test_snapshot_of_other_parameters_is_refused(oaxis_mod):
    oaxis = oaxis_mod
    p1 = oaxis.params_from_text(CASES["schwarzschild_2levels_rk4"][0])
    p2 = oaxis.params_from_text(CASES["schwarzschild_rk4_noeqsym"][0])
    with oaxis.Spacetime(p1) as a, oaxis.Spacetime(p2) as b:
        a.initial_data()
        a.initial_aux()
        data = a.checkpoint()
        with pytest.raises(oaxis.OaxisError):
            b.restore(data)
