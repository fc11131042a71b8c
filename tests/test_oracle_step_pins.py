"""Pins of the oracle's time integration and box-in-box transfers.

  * RK4 and ICN (3 iterations) on the homogeneous massive scalar field: far
    from the boundaries the field obeys y' = M y with M = [[0, 1], [-m^2, 0]]
    (alpha = 1, static slicing, vanishing backreaction at amplitude 1e-7),
    so one step multiplies (phi, pi) by the textbook amplification
    polynomials 1 + z + z^2/2 + z^3/6 + z^4/24 and 1 + z + z^2/2 + z^3/4;
  * restriction (restrict.f90): coarse points inside the fine box receive
    the fine data interpolated with the bicubic Lagrange weights, exact on
    polynomials cubic in rho (the coarse nodes sit half way between fine
    nodes in rho and on fine nodes in z);
  * prolongation (finebound.f90): the fine boundary zone is refilled from
    the coarse level exactly on bicubic polynomials;
  * time interpolation (prl/arrays.pl bound_interp): with two saved strips
    the boundary value at the intermediate time is exact on quadratics in t.
"""
import numpy as np
import pytest

STEP = """
order = four
eqsym = .false.
boundtype = none
geodiss = 0.0
scalardiss = 0.0
chimethod = .true.
slicing = static
shift = none
mattertype = scalar
scalarpotential = phi2
scalar_mass = 2.0
scalarmethod = %s
integrator = %s
dtfac = 0.5
dr0 = 0.1
dz0 = 0.1
Nrtotal = 40
Nztotal = 80
"""


def amplification(kind, dtM):
    I = np.eye(2)
    M2 = dtM @ dtM
    M3 = M2 @ dtM
    if kind == "rk4":
        return I + dtM + M2 / 2 + M3 / 6 + (M3 @ dtM) / 24
    return I + dtM + M2 / 2 + M3 / 4


@pytest.mark.parametrize("integrator", ["rk4", "icn"])
@pytest.mark.parametrize("method", ["first", "second"])
def test_one_step_amplification(oaxis_mod, oracle_mod, integrator, method):
    p = oaxis_mod.params_from_text(STEP % (method, integrator))
    o = oracle_mod.Oracle(p)
    o.initial_data()
    phi0, pi0 = 1.0e-7, -0.6e-7
    o.set("scalar_phi", np.full(o.shape(), phi0))
    o.set("scalar_pi", np.full(o.shape(), pi0))
    o.initial_aux()
    dt = o.dims()["dt"]
    o.coarse_step()
    m = 2.0
    y = amplification(integrator, dt * np.array([[0.0, 1.0], [-m * m, 0.0]])) @ np.array([phi0, pi0])
    r, z = o.coords()
    # boundary values spread at most 3 points per substep
    far = (r < r.max() - 13 * 0.1) & (np.abs(z) < z.max() - 13 * 0.1)
    got_phi = o.get("scalar_phi")[far]
    got_pi = o.get("scalar_pi")[far]
    assert np.abs(got_phi - y[0]).max() <= 1e-12 * abs(y[0])
    assert np.abs(got_pi - y[1]).max() <= 1e-12 * abs(y[1])
    # and the amplification is not the identity or a lower-order polynomial
    lower = (np.eye(2) + dt * np.array([[0.0, 1.0], [-m * m, 0.0]])) @ np.array([phi0, pi0])
    assert abs(y[1] - lower[1]) > 1e-3 * abs(y[1])
    o.close()


LEVELS = """
order = four
eqsym = .false.
interporder = bicubic
Nl0 = 1
dr0 = 0.2
dz0 = 0.2
Nrtotal = 16
Nztotal = 32
"""

FIELDS = ["alpha", "A", "trK", "KTC"]


def poly(r, z, k=0):
    return 1.0 + 0.5 * r ** 2 - 0.2 * z + 0.05 * z ** 3 + 0.1 * r ** 2 * z + 0.01 * k


def test_restriction_exact_on_cubics(oaxis_mod, oracle_mod):
    p = oaxis_mod.params_from_text(LEVELS)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    rf, zf = o.coords(1)
    rc, zc = o.coords(0)
    for k, f in enumerate(FIELDS):
        o.set(f, poly(rf, zf, k), 1)
        o.set(f, np.full(rc.shape, 7.0), 0)
    o.run_restrict(1)
    touched = None
    for k, f in enumerate(FIELDS):
        c = o.get(f, 0)
        t = c != 7.0
        touched = t if touched is None else touched
        assert np.array_equal(t, touched)
        assert np.abs(c - poly(rc, zc, k))[t].max() <= 1e-13
    # the restricted set is a non-empty box inside the fine level
    assert touched.sum() > 20
    assert rc[touched].max() < rf.max() and np.abs(zc[touched]).max() < np.abs(zf).max()
    assert rc[touched].min() > 0
    o.close()


def boundary_zone(o, depth):
    d = o.dims(1)
    g, Nr, Nz = d["ghost"], d["Nr"], d["Nz"]
    i = np.arange(1 - g, Nr + 1)[:, None] + 0 * np.arange(Nz + g)[None, :]
    j = np.arange(1 - g, Nz + 1)[None, :] + 0 * i
    return (i > Nr - depth) | (j > Nz - depth) | (j < 1 - g + depth)


def test_prolongation_exact_on_bicubics(oaxis_mod, oracle_mod):
    p = oaxis_mod.params_from_text(LEVELS)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    rf, zf = o.coords(1)
    rc, zc = o.coords(0)
    dtf = o.dims(1)["dt"]
    g = o.dims(1)["ghost"]
    zone = boundary_zone(o, g + 3)          # clocks meet: layers 0 .. g+2 are refilled
    for k, f in enumerate(FIELDS):
        o.set(f, poly(rc, zc, k), 0)
        v = poly(rf, zf, k)
        v[zone] = -3.0
        o.set(f, v, 1)
    o.set_time(0, dtf, 0.0, 0.0, 1)
    o.set_time(1, 0.0, 0.0, 0.0, 0)
    o.run_saveold(1)
    o.run_finebound(1, dtf)
    for k, f in enumerate(FIELDS):
        got = o.get(f, 1)
        assert np.abs(got - poly(rf, zf, k))[zone].max() <= 1e-12, f
    o.close()


def test_time_interpolation_exact_on_quadratics(oaxis_mod, oracle_mod):
    p = oaxis_mod.params_from_text(LEVELS)
    o = oracle_mod.Oracle(p)
    o.initial_data()
    rf, zf = o.coords(1)
    rc, zc = o.coords(0)
    g = o.dims(1)["ghost"]
    q = lambda t: 1.0 + 2.0 * t - 3.0 * t * t
    tm1, t0, tp, tl = 0.1, 0.2, 0.3, 0.25
    f = "alpha"
    o.set(f, q(tm1) * poly(rf, zf), 1)
    o.run_saveold(1)
    o.set(f, q(t0) * poly(rf, zf), 1)
    o.run_saveold(1)
    o.set(f, q(tp) * poly(rc, zc), 0)
    o.set_time(0, tp, 0.1, 0.0, 1)
    o.set_time(1, t0, tm1, 0.0, 2)
    o.run_finebound(1, tl - t0)
    zone = boundary_zone(o, g)              # layers 0 .. g-1
    got = o.get(f, 1)
    assert np.abs(got - q(tl) * poly(rf, zf))[zone].max() <= 1e-12
    # points outside the zone keep their t0 values
    assert np.array_equal(got[~zone], (q(t0) * poly(rf, zf))[~zone])
    o.close()
