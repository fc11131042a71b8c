"""Pins of the oracle's boundary and box-in-box pieces, and of its initial
data, by checks the mathematics fixes (no use of the oracle's own index
arithmetic or formulas):

  * radiative boundaries (radbound.f90:44-70, radiative_geometry.f90:58-237,
    scalar.f90:244-251): a field f = f0 + u(R)/R is an outgoing spherical
    wave, for which the boundary source -v (rho d_rho f + z d_z f + f - f0)/R
    equals -v u'(R)/R; at i = Nr, j = Nz and j = 1-ghost, corners included,
    for trK (v = sqrt(gauge_f)), KTA, KTB, KTC (v = 1), dtbeta (v =
    sqrt(4 drivercsi / 3)) and the scalar pi (v = 1), converging at the order
    of the one-sided closures;
  * the KTA / KTB coupling to the slicing mode (radiative_geometry.f90:113-131)
    vanishes when the gauge speed is the speed of light (gauge_f = 1);
  * restriction (restrict.f90:88-162) writes exactly the coarse points that
    lie more than 2 ghost fine cells inside the fine level (the axis side of
    a fine level that reaches the axis is open), the set computed here from
    coordinates;
  * finebound (finebound.f90:70-400) writes exactly the fine points within
    the zone depth of the outer boundaries, depth ghost+2 when the coarse
    clock equals the target time and ghost-1 otherwise, with the time
    weights of linear interpolation between t0 and tp;
  * the t = 0 branches of auxiliary (auxiliary_geometry.f90:767-821,
    1085-1114): at clock 0 Delta is overwritten by DelDef of the metric (the
    sympy DelDef of a curvilinear flat metric) and lambda, Alambda by their
    definitions; at a later clock Delta is kept;
  * scalar-pulse and Brill-wave initial data: the closed forms of
    idata_scalarpulse.f90:60-103 (the shell with r0 = z0, sr0 = sz0 reduces
    to A [exp(-(R-r0)^2/s^2) + exp(-(R+r0)^2/s^2)]) and the Brill metric
    exp(2q) after the determinant factoring of initial.f90:528-580.
"""
import numpy as np
import pytest
import sympy as sp

from test_oracle_rhs_pins import BASE, R as RS, Z as ZS, disguised_flat, lamb


def make(oaxis_mod, oracle_mod, h, n, extra):
    p = oaxis_mod.params_from_text(BASE % (h, h, n, 2 * n) + extra)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    return o


def wave(a, R0=3.0, w=1.0):
    """u(R) and u'(R) of a Gaussian shell."""
    u = lambda R: a * np.exp(-((R - R0) / w) ** 2)
    du = lambda R: -2.0 * (R - R0) / w ** 2 * u(R)
    return u, du


RADIATIVE = {           # field: (amplitude, speed as a function of (gauge_f, drivercsi))
    "trK": (0.3, lambda f, c: np.sqrt(f)),
    "KTA": (0.2, lambda f, c: 1.0),
    "KTB": (-0.25, lambda f, c: 1.0),
    "KTC": (0.15, lambda f, c: 1.0),
    "dtbeta_r": (0.1, lambda f, c: np.sqrt(4.0 * c / 3.0)),
    "dtbeta_z": (-0.12, lambda f, c: np.sqrt(4.0 * c / 3.0)),
    "scalar_pi": (0.05, lambda f, c: 1.0),
}


def boundary_mask(o):
    r, z = o.coords()
    d = o.dims(0)
    g, Nr, Nz = d["ghost"], d["Nr"], d["Nz"]
    m = np.zeros(r.shape, dtype=bool)
    m[Nr + g - 1, g:] = True          # i = Nr, j = 1 .. Nz
    m[g:, Nz + g - 1] = True          # j = Nz, i = 1 .. Nr
    m[g:, 0] = True                   # j = 1 - ghost (no equatorial symmetry)
    m[Nr + g - 1, :] = True           # i = Nr over the whole column
    m[:g, :] = False                  # axis images are not boundary points
    return m


def radiative_errors(oaxis_mod, oracle_mod, h, n, gauge_f, drivercsi, fields):
    extra = ("boundtype = radiative\nshift = Gammadriver2\ndrivercsi = %r\nmattertype = scalar\n"
             "gauge_f = %r\n" % (drivercsi, gauge_f))
    o = make(oaxis_mod, oracle_mod, h, n, extra)
    r, z = o.coords()
    R = np.sqrt(r * r + z * z)
    for name in fields:
        u, _ = wave(RADIATIVE[name][0])
        o.set(name, u(R) / R)
    o.run_auxiliary(0)
    o.run_sources(0)
    m = boundary_mask(o)
    out = {}
    for name in fields:
        a, speed = RADIATIVE[name]
        _, du = wave(a)
        want = -speed(gauge_f, drivercsi) * du(R) / R
        got = o.view(name, 0, source=True)
        out[name] = (np.abs(got - want)[m].max(), np.abs(want[m]).max())
    o.close()
    return out


@pytest.mark.parametrize("gauge_f, fields", [
    (2.0, ["trK", "KTC", "dtbeta_r", "dtbeta_z", "scalar_pi"]),
    (1.0, ["trK", "KTA", "KTB", "KTC"]),          # the slicing coupling vanishes
    (2.0, ["KTA", "KTB"]),                         # trK = 0: no slicing-mode source
])
def test_radiative_boundary_is_an_outgoing_wave(oaxis_mod, oracle_mod, gauge_f, fields):
    c = radiative_errors(oaxis_mod, oracle_mod, 0.1, 30, gauge_f, 0.75, fields)
    f = radiative_errors(oaxis_mod, oracle_mod, 0.05, 60, gauge_f, 0.75, fields)
    for name in fields:
        print(name, c[name], f[name])
        assert c[name][1] > 1e-3                   # the wave is really on the boundary
        assert c[name][0] / f[name][0] >= 10.0, (name, c[name], f[name])


def test_radiative_slicing_coupling_of_KTA_KTB(oaxis_mod, oracle_mod):
    """gauge_f = 2 with trK = u(R)/R: KTA and KTB carry their own outgoing
    wave plus (1 - 1/sqrt(gauge_f)) times the trK boundary source, with the
    weights 2/3 (the component along the boundary normal) and -1/3."""
    h, n = 0.05, 60
    o = make(oaxis_mod, oracle_mod, h, n, "boundtype = radiative\ngauge_f = 2.0\n")
    r, z = o.coords()
    R = np.sqrt(r * r + z * z)
    uk, duk = wave(0.3)
    ua, dua = wave(0.2)
    o.set("trK", uk(R) / R)
    o.set("KTA", ua(R) / R)
    o.run_auxiliary(0)
    o.run_sources(0)
    d = o.dims(0)
    g, Nr, Nz = d["ghost"], d["Nr"], d["Nz"]
    strk = -np.sqrt(2.0) * duk(R) / R
    c = 1.0 - 1.0 / np.sqrt(2.0)
    own = -dua(R) / R
    sA, sB = o.view("KTA", 0, source=True), o.view("KTB", 0, source=True)
    i = Nr + g - 1
    rows = slice(g + 2, Nz + g - 3)                # the rho boundary away from the z corners
    assert np.abs(sA[i, rows] - (own + 2.0 / 3.0 * c * strk)[i, rows]).max() < 1e-5
    assert np.abs(sB[i, rows] - (-1.0 / 3.0 * c * strk)[i, rows]).max() < 1e-5
    cols = slice(g, Nr + g - 3)                    # the upper z boundary away from the rho corner
    j = Nz + g - 1
    assert np.abs(sA[cols, j] - (own - 1.0 / 3.0 * c * strk)[cols, j]).max() < 1e-5
    assert np.abs(sB[cols, j] - (2.0 / 3.0 * c * strk)[cols, j]).max() < 1e-5
    o.close()


TWO_LEVELS = """
Nl0 = 1
"""


@pytest.mark.parametrize("eqsym", [False, True])
def test_restriction_writes_the_well_inside_points(oaxis_mod, oracle_mod, eqsym):
    n = 24
    extra = TWO_LEVELS + ("eqsym = .true.\n" if eqsym else "")
    p = oaxis_mod.params_from_text((BASE % (0.1, 0.1, n, 2 * n)).replace("eqsym = .false.\n", "") + extra)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    o.set("alpha", np.full(o.shape(1), 7.0), 1)
    o.set("alpha", np.zeros(o.shape(0)), 0)
    o.run_restrict(1)
    rc, zc = o.coords(0)
    rf, zf = o.coords(1)
    g = o.dims(1)["ghost"]
    dr, dz = o.dims(1)["dr"], o.dims(1)["dz"]
    # the fine level's extent, including its ghost points, shrunk by 2 ghost
    # fine cells; the side at the axis (or at the equator with equatorial
    # symmetry) stays open down to 0
    rlo = 0.0 if rf.min() < 0 else rf.min() + 2 * g * dr
    zlo = 0.0 if (eqsym and zf.min() < 0) else zf.min() + 2 * g * dz
    rhi, zhi = rf.max() - 2 * g * dr, zf.max() - 2 * g * dz
    want = (rc > rlo + 1e-9) & (rc < rhi - 1e-9) & (zc > zlo + 1e-9) & (zc < zhi - 1e-9)
    got = np.abs(o.get("alpha", 0) - 7.0) < 1e-12
    changed = o.get("alpha", 0) != 0.0
    assert np.array_equal(changed, got)            # every written point took the fine value
    assert np.array_equal(got, want), (np.argwhere(got != want)[:5])
    assert want.sum() > 50
    o.close()


@pytest.mark.parametrize("same", [True, False])
def test_finebound_zone_depth_and_time_weights(oaxis_mod, oracle_mod, same):
    n = 24
    p = oaxis_mod.params_from_text(BASE % (0.1, 0.1, n, 2 * n) + TWO_LEVELS)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    dtf = o.dims(1)["dt"]
    # coarse level at tp, fine level at t0 = 0 with t-1 = 0 (linear weights);
    # the call interpolates to tl = t0 + dtw
    dtw = dtf if same else 0.5 * dtf
    o.set_time(0, dtf, 0.0, 0.0, 1)
    o.set_time(1, 0.0, 0.0, 0.0, 0)
    o.set("alpha", np.full(o.shape(0), 7.0), 0)
    o.set("alpha", np.zeros(o.shape(1)), 1)
    o.run_saveold(1)                               # var_p = 0 on the fine level
    o.run_finebound(1, dtw)
    d = o.dims(1)
    g, Nr, Nz = d["ghost"], d["Nr"], d["Nz"]
    rf, zf = o.coords(1)
    depth = g + 2 if same else g - 1
    dr, dz = d["dr"], d["dz"]
    # fine points within depth cells of the outer rho boundary or of either
    # z boundary (from coordinates), axis images excluded
    near = ((rf > rf.max() - (depth + 0.5) * dr) | (zf > zf.max() - (depth + 0.5) * dz)
            | (zf < zf.min() + (depth + 0.5) * dz)) & (rf > 0)
    weight = (dtw - 0.0) / (dtf - 0.0)             # (tl - t0) / (tp - t0), var_p = 0
    got = o.get("alpha", 1)
    assert np.all(got[~near & (rf > 0)] == 0.0)
    assert np.abs(got[near] - 7.0 * weight).max() < 1e-12
    o.close()


def test_time_zero_overwrites_delta(oaxis_mod, oracle_mod):
    """Seeded Delta values are replaced by DelDef at clock 0 and kept at a
    later clock; lambda = (A - H)/rho^2 by definition."""
    d = disguised_flat()
    fields = {k: lamb(d[k]) for k in ("A", "B", "C", "H", "phi")}
    res = []
    for h, n in ((0.1, 30), (0.05, 60)):
        o = make(oaxis_mod, oracle_mod, h, n, "")
        r, z = o.coords()
        for k, fn in fields.items():
            o.set(k, fn(r, z))
        o.set("chi", np.exp(-2.0 * fields["phi"](r, z)))
        o.set("Delta_r", np.full(r.shape, 0.37))
        o.set("Delta_z", np.full(r.shape, -0.21))
        o.run_auxiliary(0)                         # clock 0: Delta <- DelDef
        m = (r > 0) & (r <= 2.4) & (np.abs(z) <= 2.4)
        res.append((np.abs(o.get("Delta_r", 0) - o.get("DelDef_r", 0))[m].max(),
                    np.abs(o.get("Delta_z", 0) - o.get("DelDef_z", 0))[m].max()))
        lam = o.get("lambda", 0)
        assert np.abs(lam * r * r - (o.get("A", 0) - o.get("H", 0))).max() < 1e-12
        o.set_time(0, 0.5, 0.25, 0.0, 2)           # a later clock: Delta is kept
        o.set("Delta_r", np.full(r.shape, 0.37))
        o.run_auxiliary(0)
        assert np.all(o.get("Delta_r", 0) == 0.37)
        o.close()
    assert res[0][0] == 0.0 and res[0][1] == 0.0 and res[1][0] == 0.0 and res[1][1] == 0.0


def test_delta_at_time_zero_is_the_flat_connection(oaxis_mod, oracle_mod):
    """On a curvilinear flat metric, DelDef = -g~^jk Gamma~^i_jk + (the
    regularisation of the phi direction), known in closed form through the
    coordinate change: converging at 4th order to the sympy value."""
    d = disguised_flat()
    A, B, C, H = d["A"], d["B"], d["C"], d["H"]
    # conformal metric of the (r, z, phi) space: [[A, r C, 0], [r C, B, 0], [0, 0, r^2 H]]
    P = sp.symbols("p")
    x = [RS, ZS, P]
    gm = sp.Matrix([[A, RS * C, 0], [RS * C, B, 0], [0, 0, RS ** 2 * H]])
    det = gm[0, 0] * gm[1, 1] - gm[0, 1] * gm[1, 0]
    gi = sp.Matrix([[gm[1, 1] / det, -gm[0, 1] / det, 0], [-gm[1, 0] / det, gm[0, 0] / det, 0],
                    [0, 0, 1 / gm[2, 2]]])
    Gam = [[[sum(gi[i, m] * (sp.diff(gm[m, j], x[k]) + sp.diff(gm[m, k], x[j]) - sp.diff(gm[j, k], x[m]))
                 for m in range(3)) / 2 for k in range(3)] for j in range(3)] for i in range(2)]
    contr = [sum(gi[j, k] * Gam[i][j][k] for j in range(3) for k in range(3)) for i in range(2)]
    # BSSN Delta^r carries the flat (cylindrical) connection subtracted: Delta^i
    # = g~^jk (Gamma~^i_jk - Gamma_flat^i_jk); the flat rho-rho-phi term is -r
    flat_r = gi[2, 2] * (-RS)
    fr = sp.lambdify((RS, ZS), contr[0] - flat_r, "numpy", cse=True)
    fz = sp.lambdify((RS, ZS), contr[1], "numpy", cse=True)
    want_r = lambda r, z: fr(r, z) + 0.0 * r
    want_z = lambda r, z: fz(r, z) + 0.0 * r
    fields = {k: lamb(d[k]) for k in ("A", "B", "C", "H", "phi")}
    errs = []
    for h, n in ((0.1, 30), (0.05, 60)):
        o = make(oaxis_mod, oracle_mod, h, n, "")
        r, z = o.coords()
        for k, fn in fields.items():
            o.set(k, fn(r, z))
        o.set("chi", np.exp(-2.0 * fields["phi"](r, z)))
        o.run_auxiliary(0)
        m = (r > 0) & (r <= 2.4) & (np.abs(z) <= 2.4)
        # Delta^r is stored divided by nothing: compare both components
        er = np.abs(o.get("Delta_r", 0) - want_r(r, z))[m].max()
        ez = np.abs(o.get("Delta_z", 0) - want_z(r, z))[m].max()
        errs.append((er, ez))
        assert np.abs(want_r(r, z))[m].max() > 0.05
        o.close()
    assert errs[0][0] / errs[1][0] >= 10.0 and errs[0][1] / errs[1][1] >= 10.0, errs


def test_factordet_only_after_time_zero(oaxis_mod, oracle_mod):
    """auxiliary leaves a conformal metric with det != 1 alone at clock 0 and
    rescales it to unit determinant (A B - rho^2 C^2) H = 1 at a later clock
    (auxiliary_geometry.f90:265-316, factordet.f90:40-64)."""
    o = make(oaxis_mod, oracle_mod, 0.1, 24, "")
    r, z = o.coords()
    bump = 0.2 * np.exp(-(r * r + z * z))
    A, B, H, C = 1.0 + bump, 1.0 + 0.5 * bump, 1.0 - 0.3 * bump, 0.1 * bump
    for name, v in (("A", A), ("B", B), ("H", H), ("C", C)):
        o.set(name, v)
    o.run_auxiliary(0)
    assert np.array_equal(o.get("A", 0), A) and np.array_equal(o.get("C", 0), C)
    o.set_time(0, 0.5, 0.25, 0.0, 2)
    o.run_auxiliary(0)
    det = (o.get("A", 0) * o.get("B", 0) - (r * o.get("C", 0)) ** 2) * o.get("H", 0)
    assert np.abs(det - 1.0).max() <= 4e-16 * 8
    # the rescaling is by the cube root of the old determinant
    old = (A * B - (r * C) ** 2) * H
    assert np.abs(o.get("A", 0) * np.cbrt(old) - A).max() <= 1e-15
    o.close()


SCALAR = """
mattertype = scalar
idata = scalarpulse
scalar_a0 = 0.04
WE_maxiter = 100
"""


@pytest.mark.parametrize("kind", ["shell_r0_eq_z0", "central", "torus"])
def test_scalar_pulse_profiles(oaxis_mod, oracle_mod, kind):
    extra = SCALAR + {
        "shell_r0_eq_z0": "scalarpulse_type = shell\nscalar_r0 = 1.2\nscalar_z0 = 1.2\nscalar_sr0 = 0.6\nscalar_sz0 = 0.6\n",
        "central": "scalarpulse_type = shell\nscalar_r0 = 0.0\nscalar_z0 = 0.0\nscalar_sr0 = 0.9\nscalar_sz0 = 0.6\n",
        "torus": "scalarpulse_type = torus\nscalar_r0 = 1.0\nscalar_z0 = 0.5\nscalar_sr0 = 0.5\nscalar_sz0 = 0.4\n",
    }[kind]
    o = make(oaxis_mod, oracle_mod, 0.1, 24, extra)
    r, z = o.coords()
    R = np.sqrt(r * r + z * z)
    A = 0.04
    if kind == "shell_r0_eq_z0":
        want = A * (np.exp(-(R - 1.2) ** 2 / 0.36) + np.exp(-(R + 1.2) ** 2 / 0.36))
    elif kind == "central":
        want = A * np.exp(-(r * r / 0.81 + z * z / 0.36))
    else:   # torus without equatorial symmetry: a ring at rho = 1, z = 0.5 and its axis image
        want = A * np.exp(-((z - 0.5) / 0.4) ** 2) * (np.exp(-((r - 1.0) / 0.5) ** 2) + np.exp(-((r + 1.0) / 0.5) ** 2))
    assert np.abs(o.get("scalar_phi", 0) - want).max() <= 1e-15 + 1e-13 * np.abs(want).max()
    assert np.all(o.get("scalar_pi", 0) == 0.0)
    o.close()


def test_brill_metric_after_determinant_factoring(oaxis_mod, oracle_mod):
    """A = B = exp(2q), H = 1, C = 0 with q = a rho^2 exp(-rho^2/sr^2 - z^2/sz^2),
    then divided by hdet^(1/3) = exp(4q/3): A = B = exp(2q/3), H = exp(-4q/3),
    and phi = log(psi) + q/3 with psi from the elliptic solve."""
    a, sr, sz = 1.5, 1.0, 0.8
    extra = "idata = BrillWave\nbrill_a0 = %r\nbrill_sr0 = %r\nbrill_sz0 = %r\nWE_maxiter = 100\n" % (a, sr, sz)
    o = make(oaxis_mod, oracle_mod, 0.1, 24, extra)
    r, z = o.coords()
    q = a * r * r * np.exp(-(r / sr) ** 2 - (z / sz) ** 2)
    for name, want in (("A", np.exp(2 * q / 3)), ("B", np.exp(2 * q / 3)), ("H", np.exp(-4 * q / 3))):
        assert np.abs(o.get(name, 0) - want).max() <= 1e-14 * np.abs(want).max(), name
    assert np.all(o.get("C", 0) == 0.0)
    psi = o.view("ell_u", 0)
    assert np.abs(o.get("phi", 0) - (np.log(psi) + q / 3)).max() <= 1e-14
    o.close()
