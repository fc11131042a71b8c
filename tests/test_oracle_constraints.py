"""Pins of the oracle's constraints (src/geometry/constraints.f90) against
closed forms, at 4th-order convergence:

  * Minkowski: every constraint is exactly zero;
  * flat space in curvilinear disguise (non-trivial conformal metric and
    conformal factor, K = 0, Delta set to the metric's own DelDef): ham and
    the Delta constraints -> 0, which pins the Ricci scalar through a path
    independent of the evolution equations;
  * flat metric with K_ij = eps * Hess(f) (the extrinsic curvature of a
    slightly tilted slice of Minkowski, to first order): the momentum
    constraint -> 0, and ham = K^2 - K_ij K^ij exactly (pointwise algebra);
  * Schwarzschild in isotropic coordinates: ham -> 0 away from the puncture.

The residual must fall by >= 10 (ideal 16) when the grid spacing halves.
"""
import numpy as np
import sympy as sp

from test_oracle_rhs_pins import BASE, CUT, SIZE, R, Z, disguised_flat_fields, lamb

CONSTRAINTS = ["ham", "mom_r", "mom_z", "CDelta_r", "CDelta_z"]


def run(oaxis_mod, oracle_mod, h, extra, fields, size=SIZE):
    n = int(round(size / h))
    p = oaxis_mod.params_from_text(BASE % (h, h, n, 2 * n) + extra)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    r, z = o.coords()
    for name, fn in fields.items():
        o.set(name, fn(r, z))
    o.run_auxiliary(0)
    o.constraints(0)
    return o, r, z


def converge(oaxis_mod, oracle_mod, extra, fields, expected, mask_fn, hs=(0.1, 0.05)):
    res = []
    for h in hs:
        o, r, z = run(oaxis_mod, oracle_mod, h, extra, fields)
        m = mask_fn(r, z)
        res.append({name: np.abs(o.get(name, 0) - (fn(r, z) if fn else 0.0))[m].max()
                    for name, fn in expected.items()})
        o.close()
    return res


def inner(r, z):
    return (r > 0) & (r <= CUT) & (np.abs(z) <= CUT)


def assert_ratio(res, names, floor=1e-12):
    c, f = res
    for name in names:
        if c[name] < floor:
            assert f[name] < 10 * floor, (name, c[name], f[name])
            continue
        ratio = c[name] / max(f[name], 1e-300)
        assert ratio >= 10.0, "%s: %.3e -> %.3e (ratio %.1f)" % (name, c[name], f[name], ratio)


def test_minkowski_constraints_vanish(oaxis_mod, oracle_mod):
    o, r, z = run(oaxis_mod, oracle_mod, 0.1, "mattertype = scalar\nshift = Gammadriver2\n", {})
    for name in CONSTRAINTS:
        assert np.abs(o.get(name, 0)).max() < 1e-14, name
    o.close()


def test_disguised_flat_space_satisfies_constraints(oaxis_mod, oracle_mod):
    fields = disguised_flat_fields()

    def with_deldef(h):
        # Delta = DelDef of the disguised metric, as the initial auxiliary sets it
        o, r, z = run(oaxis_mod, oracle_mod, h, "", fields)
        o.set("Delta_r", o.get("DelDef_r", 0))
        o.set("Delta_z", o.get("DelDef_z", 0))
        o.run_auxiliary(0)
        o.constraints(0)
        m = inner(r, z)
        out = {n: np.abs(o.get(n, 0))[m].max() for n in CONSTRAINTS}
        assert np.abs(o.get("Dr_phi", 0)).max() > 0.05     # not trivially flat
        assert np.abs(o.get("C", 0)).max() > 0.05
        o.close()
        return out

    res = [with_deldef(0.1), with_deldef(0.05)]
    assert_ratio(res, ["ham"])
    for name in ("mom_r", "mom_z", "CDelta_r", "CDelta_z"):
        assert res[1][name] < 1e-12, (name, res[1][name])


def test_tilted_slice_momentum_and_hamiltonian(oaxis_mod, oracle_mod):
    eps = sp.Rational(1, 20)
    f = sp.exp(-(R ** 2 + (Z - sp.Rational(1, 4)) ** 2))
    frr, fzz, frz, fr = sp.diff(f, R, 2), sp.diff(f, Z, 2), sp.diff(f, R, Z), sp.diff(f, R)
    K = eps * (frr + fzz + fr / R)                      # flat Laplacian
    Krr, Kzz, Krz, Kpp = eps * frr, eps * fzz, eps * frz, eps * fr / R    # Kpp = K_phiphi / r^2
    fields = {
        "trK": lamb(K),
        "KTA": lamb(Krr - K / 3),
        "KTB": lamb(Kzz - K / 3),
        "KTC": lamb(sp.cancel(Krz / R)),
    }
    ham = K ** 2 - (Krr ** 2 + Kzz ** 2 + 2 * Krz ** 2 + Kpp ** 2)
    expected = {"ham": lamb(ham), "mom_r": None, "mom_z": None}
    res = converge(oaxis_mod, oracle_mod, "", fields, expected, inner, hs=(0.1, 0.05, 0.025))
    # ham involves no derivative here: K^2 - K_ij K^ij is pointwise algebra
    assert res[0]["ham"] < 1e-13 and res[1]["ham"] < 1e-13
    assert res[0]["mom_r"] > 1e-7                            # the discretisation is visible
    # three resolutions: 4th order over both halvings
    assert_ratio(res[:2], ["mom_r", "mom_z"])
    assert_ratio(res[1:], ["mom_r", "mom_z"])


def test_schwarzschild_hamiltonian_converges(oaxis_mod, oracle_mod):
    extra = "idata = schwarzschild\nBH1mass = 1.0\nilapse = isotropic\n"

    def away(r, z):
        rr = np.sqrt(r * r + z * z)
        return (rr > 1.2) & (r <= CUT) & (np.abs(z) <= CUT)

    res = converge(oaxis_mod, oracle_mod, extra, {}, {"ham": None, "mom_r": None, "CDelta_r": None}, away)
    assert_ratio(res, ["ham"])
    assert res[1]["mom_r"] < 1e-13 and res[1]["CDelta_r"] < 1e-12


def test_hamiltonian_boundary_copies(oaxis_mod, oracle_mod):
    """constraints.f90:68-93: ham copied from the neighbouring row at the
    outer boundaries and at 1-ghost, then axis parity images."""
    o, r, z = run(oaxis_mod, oracle_mod, 0.1, "idata = schwarzschild\nBH1mass = 1.0\n", {})
    h = o.get("ham", 0)
    g = 3
    assert np.array_equal(h[-1, :], h[-2, :])            # i = Nr
    assert np.array_equal(h[:, -1], h[:, -2])            # j = Nz
    assert np.array_equal(h[1:, 0], h[1:, 1])            # j = 1-ghost
    for i in range(1, g + 1):                            # axis images
        assert np.array_equal(h[g - i, :], h[g - 1 + i, :])
    o.close()


def test_conformal_transverse_traceless_momentum(oaxis_mod, oracle_mod):
    """K = 0, conformal factor psi = e^phi, and KT_ij = psi^-6 * eps Hess(f)
    with f harmonic (f = z^3 - 3 z rho^2 / 2): Hess(f) is trace-free and
    divergence-free in flat space, so the physical traceless curvature
    psi^-10 Hess(f)^ij satisfies the momentum constraint for any psi (York's
    conformal scaling).  This pins the 6 KT^ij d_j phi coupling."""
    eps = sp.Rational(1, 10)
    phi = sp.Rational(1, 10) * sp.exp(-(R ** 2 + (Z + sp.Rational(1, 5)) ** 2))
    f = Z ** 3 - sp.Rational(3, 2) * Z * R ** 2
    w = sp.exp(-6 * phi) * eps
    fields = {
        "phi": lamb(phi),
        "KTA": lamb(w * sp.diff(f, R, 2)),
        "KTB": lamb(w * sp.diff(f, Z, 2)),
        "KTC": lamb(sp.cancel(w * sp.diff(f, R, Z) / R)),
    }
    res = converge(oaxis_mod, oracle_mod, "", fields, {"mom_r": None, "mom_z": None}, inner)
    assert res[0]["mom_r"] > 1e-8 and res[0]["mom_z"] > 1e-8
    assert_ratio(res, ["mom_r", "mom_z"])
