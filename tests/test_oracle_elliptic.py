"""Pins of the oracle's WaveElliptic solve of the Hamiltonian constraint
(oracle/or_elliptic.c; src/elliptic/wave_elliptic.f90, called by
idata_brillwave.f90:128-135 and idata_scalarpulse.f90:171-178):

  * zero amplitude: the equation is lap u = 0 with the guess u = 1, which is
    its solution: u stays exactly 1 and the loop stops after 100 steps;
  * the coefficient S1 of the Brill equation equals (q_rr + q_zz)/4 of the
    Holz q function, derived here with sympy; for the scalar pulse S1 tends
    to pi |grad phi|^2 of the analytic pulse at 4th order;
  * the solved u satisfies the discrete equation lap u + S1 u = 0, evaluated
    with centred 4th-order stencils written here, to the solver's tolerance;
  * the Hamiltonian constraint of the resulting initial data (constraints
    of the BSSN variables, a path independent of the solver) converges to
    zero at 4th order over three resolutions, for a Brill wave and for a
    scalar pulse;
  * ICN and RK4 relaxation reach the same discrete solution;
  * with refinement levels, the stationary exit ends on the same solution
    as the reference's rule run to WE_maxiter.
"""
import numpy as np
import pytest
import sympy as sp

BRILL = """
order = four
idata = BrillWave
brill_a0 = %g
brill_sr0 = 1.0
brill_sz0 = 1.0
eqsym = .false.
dr0 = %g
dz0 = %g
Nrtotal = %d
Nztotal = %d
Nl0 = %d
"""

PULSE = """
order = four
mattertype = scalar
idata = scalarpulse
scalarpulse_type = shell
scalar_a0 = 0.05
scalar_r0 = 2.0
scalar_z0 = 2.0
scalar_sr0 = 0.7
scalar_sz0 = 0.7
eqsym = .false.
dr0 = %g
dz0 = %g
Nrtotal = %d
Nztotal = %d
Nl0 = 0
"""


def solve(oaxis_mod, oracle_mod, text, stationary=True):
    oracle_mod.ell_stationary(stationary)
    try:
        o = oracle_mod.Oracle(oaxis_mod.params_from_text(text))
        o.initial_data()
    finally:
        oracle_mod.ell_stationary(True)
    return o


def test_zero_amplitude_keeps_the_guess(oaxis_mod, oracle_mod):
    o = solve(oaxis_mod, oracle_mod, BRILL % (0.0, 0.2, 0.2, 24, 48, 0))
    st = o.ell_stats()
    assert st["steps0"] == 100 and st["steps1"] == 0, st
    assert np.all(o.view("ell_u", 0) == 1.0)
    assert np.all(o.get("chi", 0) == 1.0)
    o.close()
    # with a finer level the interpolation weights of the fine boundary sum
    # to one only up to rounding
    o = solve(oaxis_mod, oracle_mod, BRILL % (0.0, 0.2, 0.2, 24, 48, 1))
    st = o.ell_stats()
    assert st["steps0"] == 100 and st["steps1"] == 100, st
    for lev in range(2):
        assert np.abs(o.view("ell_u", lev) - 1.0).max() <= 1e-14
    o.close()


def brill_s1(a0, sr, sz):
    r, z = sp.symbols("r z", real=True)
    q = a0 * r ** 2 * sp.exp(-(r / sr) ** 2 - (z / sz) ** 2)
    return sp.lambdify((r, z), (sp.diff(q, r, 2) + sp.diff(q, z, 2)) / 4, "numpy")


def test_brill_coefficient_is_the_laplacian_of_q(oaxis_mod, oracle_mod):
    o = solve(oaxis_mod, oracle_mod, BRILL % (2.5, 0.2, 0.2, 24, 48, 0))
    r, z = o.coords()
    want = brill_s1(2.5, 1.0, 1.0)(r, z)
    got = o.view("ell_S1", 0)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert np.all(o.view("ell_S0", 0) == 0.0) and np.all(o.view("ell_S5", 0) == 0.0)
    o.close()


def centred_residual(u, s1, r, dr, dz, g, Nr, Nz):
    """lap u + S1 u at i in [1, Nr-3], j in [3-g, Nz-3] (array index i+g-1)."""
    i0, i1 = g, Nr + g - 3            # array rows of i = 1 .. Nr-3
    j0, j1 = 2, Nz + g - 3            # array columns of j = 3-g .. Nz-3
    c = u[i0:i1, j0:j1]
    up1, um1 = u[i0 + 1:i1 + 1, j0:j1], u[i0 - 1:i1 - 1, j0:j1]
    up2, um2 = u[i0 + 2:i1 + 2, j0:j1], u[i0 - 2:i1 - 2, j0:j1]
    zp1, zm1 = u[i0:i1, j0 + 1:j1 + 1], u[i0:i1, j0 - 1:j1 - 1]
    zp2, zm2 = u[i0:i1, j0 + 2:j1 + 2], u[i0:i1, j0 - 2:j1 - 2]
    urr = (-up2 + 16 * up1 - 30 * c + 16 * um1 - um2) / (12 * dr * dr)
    uzz = (-zp2 + 16 * zp1 - 30 * c + 16 * zm1 - zm2) / (12 * dz * dz)
    ur = (-up2 + 8 * up1 - 8 * um1 + um2) / (12 * dr)
    return urr + uzz + ur / r[i0:i1, j0:j1] + s1[i0:i1, j0:j1] * c


def test_solution_satisfies_the_discrete_equation(oaxis_mod, oracle_mod):
    h, n = 0.1, 48
    o = solve(oaxis_mod, oracle_mod, BRILL % (2.0, h, h, n, 2 * n, 0))
    d = o.dims(0)
    r, z = o.coords()
    u = o.view("ell_u", 0)
    res = centred_residual(u, brill_s1(2.0, 1.0, 1.0)(r, z), r, h, h, d["ghost"], d["Nr"], d["Nz"])
    st = o.ell_stats()
    print("steps", st, "max residual", np.abs(res).max(), "rms", np.sqrt(np.mean(res ** 2)))
    assert st["residual"] <= 1e-8
    assert np.abs(res).max() <= 1e-6
    # and the solution is far from the guess: the equation was really solved
    assert np.abs(u - 1.0).max() > 0.05
    o.close()


def hamiltonian_norm(o, cut):
    o.initial_aux()
    o.constraints(0)
    r, z = o.coords()
    ham = o.get("ham", 0)
    m = (r > 0) & (r <= cut) & (np.abs(z) <= cut)
    return np.sqrt(np.mean(ham[m] ** 2))


@pytest.mark.parametrize("kind", ["brill", "scalar"])
def test_hamiltonian_constraint_converges(oaxis_mod, oracle_mod, kind):
    # a tight solver tolerance keeps the solver's residual (ham ~ -8 psi^-5
    # times it) below the truncation error at all three resolutions
    norms = []
    for h, n in ((0.2, 20), (0.1, 40), (0.05, 80)):
        if kind == "brill":
            text = BRILL % (2.0, h, h, n, 2 * n, 0)
        else:
            text = (PULSE % (h, h, n, 2 * n)).replace("scalar_a0 = 0.05", "scalar_a0 = 0.2")
        o = solve(oaxis_mod, oracle_mod, text + "WE_epsilon = 1.d-11\n")
        norms.append(hamiltonian_norm(o, 3.0))
        o.close()
    print(kind, norms)
    for c, f in zip(norms, norms[1:]):
        assert c / f >= 10.0, norms


def test_scalar_coefficient_converges_to_grad_phi(oaxis_mod, oracle_mod):
    r_, z_ = sp.symbols("r z", real=True)
    R = sp.sqrt(r_ ** 2 + z_ ** 2)
    # shell with r0 = z0 and sr0 = sz0: A [exp(-(R-r0)^2/s^2) + exp(-(R+r0)^2/s^2)]
    phi = 0.05 * (sp.exp(-(R - 2.0) ** 2 / 0.49) + sp.exp(-(R + 2.0) ** 2 / 0.49))
    g2 = sp.lambdify((r_, z_), sp.pi * (sp.diff(phi, r_) ** 2 + sp.diff(phi, z_) ** 2), "numpy")
    errs = []
    for h, n in ((0.1, 40), (0.05, 80)):
        o = solve(oaxis_mod, oracle_mod, PULSE % (h, h, n, 2 * n))
        r, z = o.coords()
        m = (r > 0) & (r < n * h - 4 * h) & (np.abs(z) < n * h - 4 * h)
        errs.append(np.abs(o.view("ell_S1", 0) - g2(r, z))[m].max())
        o.close()
    assert errs[0] / errs[1] >= 10.0, errs


def test_icn_and_rk4_reach_the_same_solution(oaxis_mod, oracle_mod):
    base = BRILL % (2.0, 0.2, 0.2, 24, 48, 0)
    o1 = solve(oaxis_mod, oracle_mod, base)
    o2 = solve(oaxis_mod, oracle_mod, base + "WE_method = rk4\n")
    u1, u2 = o1.view("ell_u", 0), o2.view("ell_u", 0)
    assert o1.ell_stats()["steps0"] != o2.ell_stats()["steps0"]
    assert np.abs(u1 - u2).max() <= 1e-6
    o1.close()
    o2.close()


def test_stationary_exit_matches_the_reference_rule(oaxis_mod, oracle_mod):
    text = BRILL % (2.0, 0.4, 0.4, 32, 64, 2) + "WE_maxiter = 3000\n"
    full = solve(oaxis_mod, oracle_mod, text, stationary=False)
    fast = solve(oaxis_mod, oracle_mod, text, stationary=True)
    sf, ss = full.ell_stats(), fast.ell_stats()
    print("reference rule", sf, "stationary exit", ss)
    assert sf["steps0"] == ss["steps0"] and sf["steps1"] == 3000 and ss["steps1"] < 3000
    for lev in range(3):
        a, b = full.view("ell_u", lev), fast.view("ell_u", lev)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), lev
    full.close()
    fast.close()
