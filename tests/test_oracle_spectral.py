"""Pins of the oracle's spectral apparent-horizon finder (ahmethod =
spectral / both; src/geometry/horizon_finder.f90 ahspectral :1113-1972,
ahshoot's projection :872-910) against closed forms.

The fast flow's fixed point is the horizon whatever its start and step
control, so the converged surface is pinned by the same exact horizons as
the shooting method (test_oracle_horizon.py):

  * Schwarzschild in isotropic coordinates: the sphere R = M/2, so the
    cosine coefficients are a_0 = M and a_j = 0 (j > 0); the odd ones vanish
    to rounding by the reflection symmetry z -> -z, the even ones and the
    distance to M/2 fall with the grid spacing (>= 8x when it halves);
  * the area: the reference integrates the area element with the trapezoid
    rule on theta_i = i dtheta.  On a sphere that element is c sin(theta),
    whose trapezoid sum is known exactly: the sum of sin(theta_i) over the
    same points (cot(dtheta/2) dtheta on [0, pi]), not 2.  The mass must
    approach M * sqrt(trapezoid / exact) -- a few 1e-5 below M, the
    reference's own accuracy -- and the distance to that value must fall
    with the grid spacing;
  * the same Schwarzschild slice through a non-trivial coordinate change:
    the surface crosses the axis where zeta(0, z) = M/2 (a bisection of the
    closed form), with every cosine mode present;
  * Painleve-Gullstrand (horizon at r = 2M from the extrinsic curvature
    alone), equatorial symmetry, horizons centred off the origin (from a
    circle and from an ellipse, aheccentric), flat space (none);
  * the fixed point does not depend on the start: ahmethod = both (flow from
    the shooting search's surface, projected onto ack) and spectral (flow
    from a large circle) reach the same surface, the former in fewer
    iterations; a second search starts from the horizon found (ack) and
    stops at once;
  * the projection of the shooting surface onto ack (ahshoot :872-910),
    exposed by a flow allowed one iteration (not found, so ack keeps the
    projection): a_0 = (2 s / pi) * (theta_N - theta_0) * M/2 for the sphere
    (the trapezoid of a constant is exact), modes above 20 zero, odd modes
    zero with equatorial symmetry.

Not pinned beyond these: the step control of the flow (AA halved or grown),
which decides the iteration count only.
"""
import ctypes

import numpy as np
import pytest

from test_oracle_horizon import (BASE, disguised_schwarzschild, oracle_with, painleve_gullstrand,
                                 schwarzschild_fields)

TIGHT = "ahepsilon_spectral = 1e-10\n"


def search(oaxis_mod, oracle_mod, h, fields, method="spectral", k=0, extra="", eqsym=False, calls=1):
    o = oracle_with(oaxis_mod, oracle_mod, h, fields, eqsym=eqsym, extra="ahmethod = %s\n" % method + extra)
    out = [o.ahfind(k, shape=True) for _ in range(calls)]
    ack = o.ah_coefficients(k)
    o.close()
    return out if calls > 1 else out[0], ack


def sphere_mass(M, dtheta, step):
    """Mass the trapezoid rule on theta_i = i dtheta gives for a sphere of
    mass M: the area element is c sin(theta), integrated over [0, pi/step]
    (the area doubled when step = 2)."""
    n = int(round(np.pi / step / dtheta))
    th = np.arange(n + 1) * dtheta
    w = np.full(n + 1, dtheta)
    w[0] = w[-1] = 0.5 * dtheta
    trap = step * np.sum(w * np.sin(th))
    return M * np.sqrt(trap / 2.0)


def ratio_ok(coarse, fine, name):
    ratio = coarse / max(fine, 1e-300)
    assert ratio >= 8.0, "%s: %.3e -> %.3e (ratio %.1f)" % (name, coarse, fine, ratio)


def test_schwarzschild_sphere(oaxis_mod, oracle_mod):
    M = 2.0
    runs = [search(oaxis_mod, oracle_mod, h, schwarzschild_fields(M), extra=TIGHT) for h in (0.1, 0.05)]
    for res, ack in runs:
        assert res["found"] == 1 and res["errorflag"] == 0 and res["method"] == 1
        assert res["ntheta"] == 100 and res["residual"] <= 1e-10 and 1 < res["iterations"] < 1500
        assert np.abs(ack[1::2]).max() < 1e-14                     # reflection symmetry
        assert np.all(ack[11:] == 0.0)                             # only modes 0..ahNmodes evolve
        assert res["theta"][0] == 0.0 and abs(res["theta"][-1] - np.pi) < 1e-14
    e0 = [abs(ack[0] - M) for _, ack in runs]
    ej = [np.abs(ack[2:11:2]).max() for _, ack in runs]
    es = [np.abs(res["rad"] - M / 2).max() for res, _ in runs]
    assert e0[1] < 1e-6 and ej[1] < 1e-5 and es[1] < 1e-5
    ratio_ok(e0[0], e0[1], "a_0")
    ratio_ok(ej[0], ej[1], "a_j")
    mt = sphere_mass(M, np.pi / 100, 1)
    assert 1e-5 * M < M - mt < 1e-4 * M                            # the quadrature's own floor
    em = [abs(res["mass"] - mt) for res, _ in runs]
    assert em[1] < 2e-6 * M
    ratio_ok(em[0], em[1], "mass")


def test_equatorial_symmetry(oaxis_mod, oracle_mod):
    M = 2.0
    runs = [search(oaxis_mod, oracle_mod, h, schwarzschild_fields(M), extra=TIGHT, eqsym=True) for h in (0.1, 0.05)]
    for res, ack in runs:
        assert res["found"] == 1
        assert np.all(ack[1::2] == 0.0)                            # only even modes (step 2)
        assert abs(res["theta"][-1] - np.pi / 2) < 1e-14           # quarter arc, area doubled
    mt = sphere_mass(M, np.pi / 200, 2)
    em = [abs(res["mass"] - mt) for res, _ in runs]
    assert em[1] < 2e-6 * M
    ratio_ok(em[0], em[1], "mass")


@pytest.mark.parametrize("ecc", [0.0, 0.3])
def test_off_centre_horizon(oaxis_mod, oracle_mod, ecc):
    M, z0 = 1.0, 1.2
    extra = TIGHT + "N_ah = 3\nz_ah1 = 1.2\nz_ah2 = -1.2\naheccentric = %r\n" % ecc
    runs = [search(oaxis_mod, oracle_mod, h, schwarzschild_fields(M, z0), k=1, extra=extra) for h in (0.05, 0.025)]
    for res, ack in runs:
        assert res["found"] == 1 and res["zcenter"] == z0
    e0 = [abs(ack[0] - M) for _, ack in runs]
    assert e0[1] < 1e-6
    ratio_ok(e0[0], e0[1], "a_0")
    em = [abs(res["mass"] - sphere_mass(M, np.pi / 100, 1)) for res, _ in runs]
    assert em[1] < 2e-6


def test_disguised_schwarzschild_and_start_independence(oaxis_mod, oracle_mod):
    M = 2
    fields, zstar = disguised_schwarzschild(M)
    ez = []
    for h in (0.1, 0.05):
        (spec, ack_s) = search(oaxis_mod, oracle_mod, h, fields, extra=TIGHT)
        (both, ack_b) = search(oaxis_mod, oracle_mod, h, fields, method="both", extra=TIGHT)
        assert spec["found"] == 1 and both["found"] == 1
        assert both["shots"] > 10 and spec["shots"] == 0
        assert both["iterations"] < spec["iterations"]             # started at the shooting surface
        # one fixed point (modes 0..ahNmodes; both keeps the projection's
        # modes 11..20, which the flow never changes, as the reference does)
        assert np.abs(ack_s[:11] - ack_b[:11]).max() < 1e-9
        assert np.abs(ack_b[11:21]).max() > 0.0 and np.all(ack_s[11:] == 0.0)
        assert np.abs(ack_s[1:11]).max() > 1e-2                    # a surface that is no sphere
        assert abs(spec["mass"] - M) < 2e-4 * M
        ez.append(abs(spec["rad"][0] - zstar))
    assert ez[1] < 2e-6
    ratio_ok(ez[0], ez[1], "axis crossing")


def test_painleve_gullstrand(oaxis_mod, oracle_mod):
    M = 0.5
    runs = [search(oaxis_mod, oracle_mod, h, painleve_gullstrand(M), extra=TIGHT) for h in (0.1, 0.05)]
    for res, _ in runs:
        assert res["found"] == 1
    es = [np.abs(res["rad"] - 2.0 * M).max() for res, _ in runs]
    assert es[1] < 1e-5
    ratio_ok(es[0], es[1], "shape")


def test_flat_space_has_no_horizon(oaxis_mod, oracle_mod):
    res, ack = search(oaxis_mod, oracle_mod, 0.1, {})
    assert res["found"] == 0 and res["area"] == 0.0 and res["mass"] == 0.0 and res["errorflag"] == 1
    assert np.all(res["rad"] == 0.0) and np.all(ack == 0.0)


def test_second_search_starts_from_the_horizon(oaxis_mod, oracle_mod):
    (first, second), _ = search(oaxis_mod, oracle_mod, 0.1, schwarzschild_fields(2.0), calls=2)
    assert first["found"] == 1 and second["found"] == 1
    assert first["iterations"] > 50 and second["iterations"] == 1
    assert abs(second["mass"] - first["mass"]) < 1e-9


@pytest.mark.parametrize("eqsym", [False, True])
def test_projection_of_the_shooting_surface(oaxis_mod, oracle_mod, eqsym):
    M = 2.0
    res, ack = search(oaxis_mod, oracle_mod, 0.05, schwarzschild_fields(M), method="both",
                      extra="ahmaxiter = 1\n", eqsym=eqsym)
    assert res["found"] == 0 and res["iterations"] == 1 and res["shots"] > 10
    # the shooting search alone on the same state: the end points of its
    # last integration (it stops within 1e-3 of its final angle)
    shot, _ = search(oaxis_mod, oracle_mod, 0.05, schwarzschild_fields(M), method="shoot", eqsym=eqsym)
    assert shot["shots"] == res["shots"] and np.abs(shot["rad"] - M / 2).max() < 1e-5
    th0, th1 = shot["theta"][0], shot["theta"][-1]
    s = 2 if eqsym else 1
    assert abs(ack[0] - 2.0 * s / np.pi * (th1 - th0) * M / 2) < 2e-5
    assert np.all(ack[21:] == 0.0)
    if eqsym:
        assert np.all(ack[1::2] == 0.0)


def test_shoot_first(oaxis_mod, oracle_mod):
    (first, second), _ = search(oaxis_mod, oracle_mod, 0.1, schwarzschild_fields(2.0), extra="ahshootfirst = .true.\n",
                                calls=2)
    assert first["found"] == 1 and first["shots"] > 10
    assert second["found"] == 1 and second["shots"] == 0 and second["iterations"] == 1


def test_mode_count_is_checked(oaxis_mod):
    """ac(0:50) holds at most 50 modes (:1165-1168)."""
    buf = ctypes.create_string_buffer(256)
    for n, ok in ((50, True), (51, False), (0, False)):
        p = oaxis_mod.params_from_text(BASE % (".false.", 0.1, 0.1, 30, 60) + "ahmethod = spectral\nahNmodes = %d\n" % n)
        rc = oaxis_mod.lib().oa_params_check_supported(ctypes.byref(p), buf, 256)
        assert (rc == 0) == ok, (n, buf.value)
