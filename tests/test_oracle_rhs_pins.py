"""Pins of the oracle's right-hand sides against closed forms of covariant
BSSN (lagrangian flavour, eta = 2, hdet = 1) derived here with sympy, at
4th-order convergence.

Each case is a configuration whose time derivatives at t = 0 are known in
closed form, or known to vanish:

  * Minkowski: every source is exactly zero;
  * flat space in curvilinear disguise (a smooth, axis-regular remapping of
    rho and z): the conformal metric and the conformal factor are both
    non-trivial, their Ricci parts cancel, so every source -> 0 (this pins
    the conformal Ricci tensor, the phi-Ricci terms and DelDef);
  * flat metric with a non-trivial lapse: trK and KT* sources are minus the
    flat Hessian of alpha and its trace-free part;
  * flat metric with a non-trivial shift and the Gamma-driver: the metric
    sources are the Lie derivative of the flat metric, the Delta sources the
    flat vector Laplacian of beta plus grad(div beta)/3, the driver
    dtbeta sources drivercsi times those;
  * real scalar field on flat space: wave equation, potential, and the
    stress-energy terms of trK, KT* and Delta;
  * static Schwarzschild with the isotropic lapse: every source -> 0.

The residual against the closed form must fall by >= 10 (ideal 16) when
the grid spacing halves, over a region that excludes the outer closures.
"""
import numpy as np
import pytest
import sympy as sp

R, Z = sp.symbols("r z", real=True)
PI = float(np.pi)

BASE = """
order = four
eqsym = .false.
boundtype = none
geodiss = 0.0
scalardiss = 0.0
integrator = rk4
chimethod = .true.
dtfac = 0.25
slicing = 1+log
gauge_f = 2.0
dr0 = %r
dz0 = %r
Nrtotal = %d
Nztotal = %d
"""

SIZE = 3.0          # rho in (0, 3], z in [-3, 3]
CUT = 2.6           # residuals are measured for rho <= CUT, |z| <= CUT


def lamb(expr):
    f = sp.lambdify((R, Z), expr, "numpy")
    return lambda r, z: f(r, z) + 0.0 * r


def build(oaxis_mod, oracle_mod, h, extra, fields):
    n = int(round(SIZE / h))
    text = BASE % (h, h, n, 2 * n) + extra
    p = oaxis_mod.params_from_text(text)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    r, z = o.coords()
    for name, fn in fields.items():
        o.set(name, fn(r, z))
    o.run_auxiliary(0)
    o.run_sources(0)
    return o, r, z


def residuals(oaxis_mod, oracle_mod, extra, fields, expected, hs=(0.1, 0.05)):
    """max |source - expected| over the measured region for each grid."""
    out = []
    for h in hs:
        o, r, z = build(oaxis_mod, oracle_mod, h, extra, fields)
        mask = (r > 0) & (r <= CUT) & (np.abs(z) <= CUT)
        res = {}
        for name, fn in expected.items():
            want = fn(r, z) if fn is not None else 0.0 * r
            got = o.view(name, 0, source=True)
            res[name] = np.abs(got - want)[mask].max()
        out.append((res, o))
    return out


def assert_converges(runs, names, floor=1e-12):
    (c, oc), (f, of) = runs
    for name in names:
        if c[name] < floor:
            assert f[name] < 10 * floor, (name, c[name], f[name])
            continue
        ratio = c[name] / max(f[name], 1e-300)
        assert ratio >= 10.0, "%s: residual %.3e -> %.3e (ratio %.1f)" % (name, c[name], f[name], ratio)
    oc.close()
    of.close()


GEOM = ["alpha", "chi", "A", "B", "H", "C", "trK", "KTA", "KTB", "KTC", "Delta_r", "Delta_z"]


def test_minkowski_sources_vanish(oaxis_mod, oracle_mod):
    o, r, z = build(oaxis_mod, oracle_mod, 0.1, "shift = Gammadriver2\nmattertype = scalar\n", {})
    for name in GEOM + ["beta_r", "beta_z", "dtbeta_r", "dtbeta_z",
                        "scalar_phi", "scalar_xi_r", "scalar_xi_z", "scalar_pi"]:
        assert np.all(o.view(name, 0, source=True) == 0.0), name
    o.close()


def disguised_flat():
    """Flat space in coordinates (r, z) related to cylindrical (rho, zeta) by
    rho = r (1 + e a), zeta = z + e b with a, b even in r."""
    e = sp.Rational(1, 5)
    G = sp.exp(-(R ** 2 + Z ** 2))
    rho = R * (1 + e * G * (1 + Z / 2))
    zeta = Z + e * sp.Rational(3, 10) * sp.exp(-(R ** 2 + (Z - sp.Rational(1, 2)) ** 2))
    J = sp.Matrix([[sp.diff(rho, R), sp.diff(rho, Z)], [sp.diff(zeta, R), sp.diff(zeta, Z)]])
    g = J.T * J                                   # metric of the (r, z) block
    gpp = rho ** 2                                # g_phiphi
    detJ = J.det()
    psi4 = (detJ ** 2 * (rho / R) ** 2) ** sp.Rational(1, 3)
    A = g[0, 0] / psi4
    B = g[1, 1] / psi4
    C = sp.cancel(g[0, 1] / R) / psi4
    H = (rho / R) ** 2 / psi4
    phi = sp.log(psi4) / 4
    return {"A": A, "B": B, "C": C, "H": H, "phi": phi, "psi4": psi4, "rho": rho, "zeta": zeta,
            "J": J, "g": g}


def disguised_flat_fields():
    d = disguised_flat()
    return {k: lamb(d[k]) for k in ("A", "B", "C", "H", "phi")}


def test_disguised_flat_space_is_static(oaxis_mod, oracle_mod):
    fields = disguised_flat_fields()
    runs = residuals(oaxis_mod, oracle_mod, "", fields, {n: None for n in GEOM})
    # the configuration is not trivially flat: the conformal metric, the
    # conformal factor and DelDef are all far from their flat values
    o = runs[1][1]
    assert np.abs(o.get("A", 0) - 1.0).max() > 0.1
    assert np.abs(o.get("C", 0)).max() > 0.05
    assert np.abs(o.get("Dr_phi", 0)).max() > 0.05
    assert np.abs(o.get("Delta_z", 0)).max() > 0.05
    assert_converges(runs, GEOM)


def flat_laplacian(f):
    return sp.diff(f, R, 2) + sp.diff(f, R) / R + sp.diff(f, Z, 2)


def test_lapse_hessian(oaxis_mod, oracle_mod):
    a = 1 + sp.Rational(3, 10) * sp.exp(-(R ** 2 + (Z - sp.Rational(1, 2)) ** 2))
    lap = flat_laplacian(a)
    expected = {
        "trK": lamb(-lap),
        "KTA": lamb(-(sp.diff(a, R, 2) - lap / 3)),
        "KTB": lamb(-(sp.diff(a, Z, 2) - lap / 3)),
        "KTC": lamb(-sp.diff(a, R, Z) / R),
        "alpha": None, "chi": None, "A": None, "B": None, "H": None, "C": None,
        "Delta_r": None, "Delta_z": None,
    }
    runs = residuals(oaxis_mod, oracle_mod, "", {"alpha": lamb(a)}, expected)
    assert_converges(runs, expected)


def christoffel_flat_cyl():
    """Christoffel symbols of the flat metric diag(1, 1, r^2) in (r, z, p)."""
    P = sp.symbols("p")
    x = [R, Z, P]
    g = sp.diag(1, 1, R ** 2)
    gi = g.inv()
    Gam = [[[sum(gi[i, m] * (sp.diff(g[m, j], x[k]) + sp.diff(g[m, k], x[j]) - sp.diff(g[j, k], x[m]))
                 for m in range(3)) / 2 for k in range(3)] for j in range(3)] for i in range(3)]
    return x, g, gi, Gam


def vector_terms(br, bz):
    """flat covariant vector Laplacian of beta and grad(div beta)."""
    x, g, gi, Gam = christoffel_flat_cyl()
    b = [br, bz, 0]
    Db = [[sp.diff(b[i], x[j]) + sum(Gam[i][j][k] * b[k] for k in range(3)) for j in range(3)] for i in range(3)]
    DDb = [[[sp.diff(Db[i][j], x[k]) + sum(Gam[i][k][m] * Db[m][j] for m in range(3))
             - sum(Gam[m][k][j] * Db[i][m] for m in range(3)) for k in range(3)] for j in range(3)]
           for i in range(3)]
    vlap = [sp.simplify(sum(gi[j, k] * DDb[i][j][k] for j in range(3) for k in range(3))) for i in range(2)]
    div = sum(Db[k][k] for k in range(3))
    grad_div = [sum(gi[i, j] * sp.diff(div, x[j]) for j in range(3)) for i in range(2)]
    return vlap, div, grad_div


def test_shift_lie_derivative_and_gamma_driver(oaxis_mod, oracle_mod):
    br = sp.Rational(1, 5) * R * sp.exp(-(R ** 2 + Z ** 2))
    bz = sp.Rational(3, 20) * sp.exp(-(R ** 2 + (Z - sp.Rational(3, 10)) ** 2))
    vlap, div, grad_div = vector_terms(br, bz)
    csi = 0.75
    sDr = vlap[0] + grad_div[0] / 3
    sDz = vlap[1] + grad_div[1] / 3
    expected = {
        "chi": lamb(-div / 3),
        "A": lamb(2 * sp.diff(br, R) - 2 * div / 3),
        "B": lamb(2 * sp.diff(bz, Z) - 2 * div / 3),
        "H": lamb(2 * br / R - 2 * div / 3),
        "C": lamb((sp.diff(bz, R) + sp.diff(br, Z)) / R),
        "Delta_r": lamb(sDr),
        "Delta_z": lamb(sDz),
        "dtbeta_r": lamb(csi * sDr),
        "dtbeta_z": lamb(csi * sDz),
        "beta_r": None, "beta_z": None,
        "alpha": None, "trK": None, "KTA": None, "KTB": None, "KTC": None,
    }
    extra = "shift = Gammadriver2\ndrivercsi = 0.75\ndrivereta = 0.0\ndriveradv = .false.\n"
    runs = residuals(oaxis_mod, oracle_mod, extra, {"beta_r": lamb(br), "beta_z": lamb(bz)}, expected)
    assert_converges(runs, expected)


@pytest.mark.parametrize("method", ["first", "second"])
@pytest.mark.parametrize("potential", ["none", "phi2"])
def test_scalar_field_on_flat_space(oaxis_mod, oracle_mod, method, potential):
    m = sp.Rational(7, 10)
    f = sp.Rational(1, 10) * sp.exp(-(R ** 2 + (Z - sp.Rational(2, 5)) ** 2) / sp.Rational(4, 5))
    pi = sp.Rational(1, 20) * sp.exp(-(R ** 2 + Z ** 2))
    xr, xz = sp.diff(f, R), sp.diff(f, Z)
    V = m ** 2 * f ** 2 / 2 if potential == "phi2" else 0
    VP = m ** 2 * f if potential == "phi2" else 0
    q2 = xr ** 2 + xz ** 2
    p = sp.pi
    expected = {
        "scalar_phi": lamb(pi),
        "scalar_xi_r": lamb(sp.diff(pi, R)),
        "scalar_xi_z": lamb(sp.diff(pi, Z)),
        "scalar_pi": lamb(flat_laplacian(f) - VP),
        "trK": lamb(8 * p * (pi ** 2 - V)),
        "KTA": lamb(-8 * p * (xr ** 2 - q2 / 3)),
        "KTB": lamb(-8 * p * (xz ** 2 - q2 / 3)),
        "KTC": lamb(-8 * p * xr * xz / R),
        "Delta_r": lamb(16 * p * pi * xr),
        "Delta_z": lamb(16 * p * pi * xz),
        "alpha": None, "chi": None, "A": None, "B": None, "H": None, "C": None,
    }
    extra = "mattertype = scalar\nscalarmethod = %s\nscalarpotential = %s\nscalar_mass = 0.7\n" % (
        method, potential)
    fields = {"scalar_phi": lamb(f), "scalar_pi": lamb(pi), "scalar_xi_r": lamb(xr), "scalar_xi_z": lamb(xz)}
    runs = residuals(oaxis_mod, oracle_mod, extra, fields, expected)
    assert_converges(runs, expected)


def test_static_schwarzschild(oaxis_mod, oracle_mod):
    """Isotropic Schwarzschild with the static lapse (1-M/2R)/(1+M/2R) and
    zero shift: every source vanishes; measured outside R = 1.5 (M = 1)."""
    M = 1.0
    out = []
    for h in (0.1, 0.05):
        o, r, z = build(oaxis_mod, oracle_mod, h,
                        "idata = schwarzschild\nBH1mass = 1.0\nilapse = isotropic\n", {})
        rr = np.sqrt(r * r + z * z)
        want_alpha = (1 - M / (2 * rr)) / (1 + M / (2 * rr))
        assert np.abs(o.get("alpha", 0) - want_alpha).max() < 1e-14
        mask = (rr >= 1.5) & (r <= CUT) & (np.abs(z) <= CUT)
        out.append(({n: np.abs(o.view(n, 0, source=True))[mask].max() for n in GEOM}, o))
    # the configuration is not trivial: the pieces of the trK and KT sources do not vanish
    o = out[1][1]
    assert np.abs(o.get("RIC_A", 0)[np.sqrt(r * r + z * z) >= 1.5]).max() > 1e-3
    assert_converges(out, GEOM)


@pytest.mark.parametrize("slicing", ["1+log", "harmonic"])
def test_extrinsic_curvature_algebra(oaxis_mod, oracle_mod, slicing):
    """Flat metric, constant lapse 1.3, no shift, non-uniform trK and a
    traceless, axis-regular KT (KT_H = -(KTA + KTB)):
    d_t gamma~ = -2 alpha A~, d_t chi = alpha chi K / 3,
    d_t K = alpha (A~_ij A~^ij + K^2 / 3),
    d_t A~_ij = alpha (K A~_ij - 2 A~_ik A~^k_j),
    d_t Delta^i = -(4/3) alpha d^i K (eta = 2 cancels the divergence of A~),
    d_t alpha = -f alpha K (1+log) or -f alpha^2 K (harmonic)."""
    al = sp.Rational(13, 10)
    gf = sp.Rational(3, 2)
    G1 = sp.exp(-(R ** 2 + (Z - sp.Rational(1, 5)) ** 2))
    G2 = sp.exp(-(R ** 2 + Z ** 2) / 2)
    K = sp.Rational(1, 5) * sp.exp(-((R - 0) ** 2 + (Z + sp.Rational(3, 10)) ** 2))
    a = sp.Rational(1, 10) * G1
    b = -2 * a + R ** 2 * sp.Rational(1, 20) * G2         # KTA - KT_H = O(r^2)
    c = sp.Rational(1, 15) * G2 * (1 + Z)
    h = -(a + b)
    AA = a ** 2 + b ** 2 + 2 * R ** 2 * c ** 2 + h ** 2
    dalpha = -gf * al * K if slicing == "1+log" else -gf * al ** 2 * K
    expected = {
        "alpha": lamb(dalpha),
        "chi": lamb(al * K / 3),
        "A": lamb(-2 * al * a), "B": lamb(-2 * al * b), "H": lamb(-2 * al * h), "C": lamb(-2 * al * c),
        "trK": lamb(al * (AA + K ** 2 / 3)),
        "KTA": lamb(al * (K * a - 2 * (a ** 2 + R ** 2 * c ** 2))),
        "KTB": lamb(al * (K * b - 2 * (b ** 2 + R ** 2 * c ** 2))),
        "KTC": lamb(al * (K * c - 2 * c * (a + b))),
        "Delta_r": lamb(-sp.Rational(4, 3) * al * sp.diff(K, R)),
        "Delta_z": lamb(-sp.Rational(4, 3) * al * sp.diff(K, Z)),
    }
    fields = {"alpha": lamb(al + 0 * R), "trK": lamb(K), "KTA": lamb(a), "KTB": lamb(b), "KTC": lamb(c)}
    extra = "slicing = %s\ngauge_f = 1.5\n" % slicing
    runs = residuals(oaxis_mod, oracle_mod, extra, fields, expected)
    assert_converges(runs, expected)


def test_scalar_field_on_disguised_flat_space(oaxis_mod, oracle_mod):
    """The real scalar field on the disguised flat metric (alpha = 1, K = 0).
    The field is defined in the physical cylindrical frame (rho, zeta), where
    its Laplacian and its stress-energy invariants are elementary; in the
    grid coordinates its gradient is xi = J^T grad F, raised indices use
    J^-1, and the Ricci parts cancel as in the vacuum case.  This pins the
    stress-energy terms of trK, KT* and Delta with a non-diagonal metric."""
    d = disguised_flat()
    A, B, C, psi4, J, g = d["A"], d["B"], d["C"], d["psi4"], d["J"], d["g"]
    u, v = sp.symbols("u v", real=True)             # physical rho, zeta
    F = sp.Rational(1, 10) * sp.exp(-(u ** 2 + (v - sp.Rational(2, 5)) ** 2))
    Pi = sp.Rational(1, 20) * sp.exp(-(u ** 2 + v ** 2))
    m = sp.Rational(7, 10)
    lapF = sp.diff(F, u, 2) + sp.diff(F, u) / u + sp.diff(F, v, 2)
    on = {u: d["rho"], v: d["zeta"]}
    at = lambda e: e.subs(on)
    gradF = sp.Matrix([at(sp.diff(F, u)), at(sp.diff(F, v))])
    gradPi = sp.Matrix([at(sp.diff(Pi, u)), at(sp.diff(Pi, v))])
    xi = J.T * gradF                                # grid-frame gradient of f
    f_, pi_ = at(F), at(Pi)
    q2 = gradF.dot(gradF)
    V = m ** 2 * f_ ** 2 / 2
    rhoE = (pi_ ** 2 + q2) / 2 + V
    trS = -q2 / 2 + 3 * pi_ ** 2 / 2 - 3 * V
    w = (q2 - pi_ ** 2) / 2 + V
    S_rr = xi[0] ** 2 - g[0, 0] * w
    S_zz = xi[1] ** 2 - g[1, 1] * w
    S_rz = xi[0] * xi[1] - g[0, 1] * w
    up = J.inv() * gradF                            # gamma^ij xi_j
    p_ = sp.pi
    expected = {
        "scalar_phi": lamb(pi_),
        "scalar_xi_r": lamb((J.T * gradPi)[0]),
        "scalar_xi_z": lamb((J.T * gradPi)[1]),
        "scalar_pi": lamb(at(lapF) - m ** 2 * f_),
        "trK": lamb(4 * p_ * (rhoE + trS)),
        "KTA": lamb(-8 * p_ * (S_rr / psi4 - A * trS / 3)),
        "KTB": lamb(-8 * p_ * (S_zz / psi4 - B * trS / 3)),
        "KTC": lamb(-8 * p_ * (S_rz / (R * psi4) - C * trS / 3)),
        "Delta_r": lamb(16 * p_ * psi4 * pi_ * up[0]),
        "Delta_z": lamb(16 * p_ * psi4 * pi_ * up[1]),
        "alpha": None, "chi": None, "A": None, "B": None, "H": None, "C": None,
    }
    fields = disguised_flat_fields()
    fields.update({"scalar_phi": lamb(f_), "scalar_pi": lamb(pi_),
                   "scalar_xi_r": lamb(xi[0]), "scalar_xi_z": lamb(xi[1])})
    extra = "mattertype = scalar\nscalarmethod = second\nscalarpotential = phi2\nscalar_mass = 0.7\n"
    runs = residuals(oaxis_mod, oracle_mod, extra, fields, expected)
    assert_converges(runs, expected)
