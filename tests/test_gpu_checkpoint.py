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


def test_output_writers_on_device_state(oaxis_mod, tmp_path):
    """save0D/save1D/save2D on a device spacetime: the .2D values are the
    downloaded array (15 digits) and the .tl norms the device reduction."""
    import os
    from oaxis import output
    oaxis = oaxis_mod
    p = oaxis.params_from_text(CASES["schwarzschild_2levels_rk4"][0])
    with oaxis.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        st.evolve(1)
        out = str(tmp_path)
        for lev in range(st.nlevels):
            output.save0D(st, out, ["alpha", "trK"], lev, new=True)
            output.save1D(st, out, ["alpha"], lev, new=True)
            output.save2D(st, out, ["alpha"], lev, new=True)
        a = st.download("alpha", 1)
        rows = open(os.path.join(out, "alpha_b0l1.2D")).read().split("\n")
        vals = [float(l[46:]) for l in rows[1:1 + a.shape[0]]]
        assert np.allclose(vals, a[:, 0] + 1e-30, rtol=1e-14, atol=0)
        nm2 = float(open(os.path.join(out, "trK_b0l0_nm2.tl")).read().split("\n")[1][20:])
        assert abs(nm2 - st.norms0d("trK", 0)["nm2"]) <= 1e-8 * abs(nm2)
