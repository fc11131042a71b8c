"""GPU: the constraints (constraints.f90) on the device.

  * parity with the oracle after several coarse steps, on every level: the
    L2 norms (save0Ddata's nm2) of ham, mom_r, mom_z, CDelta_r, CDelta_z
    agree to 1e-8 relative, and the arrays to 1e-8 of the max norm of their
    leading term (RSCAL for ham, DIV_KT for mom, Delta for CDelta);
  * 4th-order convergence of the Hamiltonian constraint over three
    resolutions, for Schwarzschild initial data and after an evolution to
    a fixed time (away from the puncture and the outer boundary).
"""
import numpy as np
import pytest

from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu

NAMES = ["ham", "mom_r", "mom_z", "CDelta_r", "CDelta_z"]
# Each constraint is a small difference of O(1) terms (the cancellation is
# what it measures), so pointwise differences are compared against the
# scale of its leading term; the L2 norms are compared directly.
TERMS = {"ham": ("RSCAL", "KT2"), "mom_r": ("DIV_KT_r",), "mom_z": ("DIV_KT_z",),
         "CDelta_r": ("Delta_r",), "CDelta_z": ("Delta_z",)}


def nm2(a, g=3):
    sub = a[g:a.shape[0] - g, g:a.shape[1] - g]
    return np.sqrt(np.sum(sub * sub) / sub.size)


@pytest.mark.parametrize("case", ["schwarzschild_2levels_rk4", "scalar_pulse_gd2_levels",
                                  "gauge_pulse_gammadriver2", "brill_gd2_odd_sizes_icn"])
def test_constraints_match_oracle(oaxis_mod, oracle_mod, case):
    oaxis, pyoracle = oaxis_mod, oracle_mod
    text, nsteps = CASES[case]
    p = oaxis.params_from_text(text)
    orc = pyoracle.Oracle(p)
    orc.initial_data()
    st = oaxis.Spacetime(p)
    for lev in range(st.nlevels):
        for name in oaxis.state_fields(p):
            st.upload(name, orc.get(name, lev), lev)
    orc.initial_aux()
    st.initial_aux()
    for _ in range(nsteps):
        orc.coarse_step()
        st.coarse_step()
    worst = []
    for lev in range(st.nlevels):
        orc.constraints(lev)
        st.constraints(lev)
        for name in NAMES:
            a, b = st.download(name, lev), orc.get(name, lev)
            scale = max([np.abs(b).max()] + [np.abs(orc.get(t, lev)).max() for t in TERMS[name]])
            if scale == 0.0:
                assert np.abs(a).max() < 1e-12, (case, lev, name)
                continue
            err = np.abs(a - b).max() / scale
            n_a, n_b = st.norms0d(name, lev)["nm2"], nm2(b)
            nerr = abs(n_a - n_b) / n_b if n_b > 0 else abs(n_a)
            worst.append((max(err, nerr), lev, name))
            assert err <= 1e-8 and nerr <= 1e-8, (case, lev, name, err, nerr)
    print("constraints %s: worst %s" % (case, max(worst)))
    st.close()
    orc.close()


def _schwarzschild(oaxis, n, extra=""):
    text = ("idata = schwarzschild\nBH1mass = 1.0\nilapse = psiminus2\nslicing = 1+log\ngauge_f = 2.0\n"
            "order = four\nintegrator = rk4\ndtfac = 0.25\neqsym = .false.\nboundtype = radiative\n"
            "geodiss = 0.1\ndr0 = %r\ndz0 = %r\nNrtotal = %d\nNztotal = %d\n" % (3.2 / n, 3.2 / n, n, 2 * n))
    return oaxis.params_from_text(text + extra)


def _constraint_errors(oaxis, n, tfinal, names=("ham",)):
    p = _schwarzschild(oaxis, n)
    with oaxis.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        dt = st.level_info(0)["dt"]
        steps = int(round(tfinal / dt))
        if steps:
            st.evolve(steps)
        st.constraints(0)
        arrays = {name: st.download(name, 0) for name in names}
        ham = arrays[names[0]]
        info = st.level_info(0)
        g, dr = info["ghost"], info["dr"]
        i = np.arange(ham.shape[0]) + 1 - g
        j = np.arange(ham.shape[1]) + 1 - g
        r = (i - 0.5) * dr
        zoff = 0.5 * (info["Nz"] - g + 1)
        z = (j - zoff) * dr
        rr = np.sqrt(r[:, None] ** 2 + z[None, :] ** 2)
        mask = (rr > 1.2) & (r[:, None] > 0) & (r[:, None] < 2.4) & (np.abs(z[None, :]) < 2.4)
        return {name: np.abs(a[mask]).max() for name, a in arrays.items()}


def _ham_error(oaxis, n, tfinal):
    return _constraint_errors(oaxis, n, tfinal)["ham"]


def test_momentum_converges_at_fourth_order(oaxis_mod):
    """The momentum constraints of the evolved puncture (zero at t = 0, where
    K = 0) are truncation error after 0.2 of evolution: they fall by >= 10
    per halving of the grid spacing over 32/64/128."""
    errs = [_constraint_errors(oaxis_mod, n, 0.2, ("mom_r", "mom_z")) for n in (32, 64, 128)]
    print("mom convergence: %s" % errs)
    for name in ("mom_r", "mom_z"):
        e = [x[name] for x in errs]
        assert e[0] > 1e-8
        assert e[0] / e[1] >= 10.0 and e[1] / e[2] >= 10.0, (name, e)


@pytest.mark.parametrize("tfinal", [0.0, 0.2])
def test_hamiltonian_converges_at_fourth_order(oaxis_mod, tfinal):
    errs = [_ham_error(oaxis_mod, n, tfinal) for n in (32, 64, 128)]
    ratios = [errs[0] / errs[1], errs[1] / errs[2]]
    print("ham convergence t=%.1f: %s ratios %s" % (tfinal, errs, ratios))
    assert all(rt >= 10.0 for rt in ratios), (errs, ratios)
