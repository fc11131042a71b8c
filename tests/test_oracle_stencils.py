"""Pins of the oracle's grid operators against what the mathematics fixes.

Each stencil of the fourth-order scheme is exact on polynomials up to its
formal degree (derivatives.f90, derivadvect.f90), parity ghosts carry the
images of the interior, Kreiss-Oliger dissipation annihilates polynomials of
degree <= 5 and has the Fourier symbol -(eps/20 dt)(2 - 2 cos kh)^3
(dissipation.f90:150-200), and the bicubic interpolation of interp.f90 is
exact on bicubic polynomials.  A wrong coefficient, a shifted index or a
sign error fails one of these.
"""
import itertools

import numpy as np
import pytest

GRID = """
order = four
eqsym = %s
Nrtotal = 14
Nztotal = 18
dr0 = 0.1
dz0 = 0.125
dtfac = 0.5
"""


@pytest.fixture(params=[".false.", ".true."], ids=["noeqsym", "eqsym"])
def orc(request, oaxis_mod, oracle_mod):
    p = oaxis_mod.params_from_text(GRID % request.param)
    o = oracle_mod.Oracle(p)
    o.eqsym = request.param == ".true."
    yield o
    o.close()


def monomials(maxdeg, rpar, zpar=None):
    """(a, b) exponents of r^a z^b with total degree <= maxdeg and the given
    parity in r (and in z when zpar is set)."""
    out = []
    for a, b in itertools.product(range(maxdeg + 1), repeat=2):
        if a + b > maxdeg or (a % 2) != (0 if rpar > 0 else 1):
            continue
        if zpar is not None and (b % 2) != (0 if zpar > 0 else 1):
            continue
        out.append((a, b))
    return out


def dpoly(a, b, r, z, dr, dz):
    """d^dr/dr^dr d^dz/dz^dz of r^a z^b."""
    ca, pa = 1.0, a
    for _ in range(dr):
        ca *= pa
        pa -= 1
    cb, pb = 1.0, b
    for _ in range(dz):
        cb *= pb
        pb -= 1
    if ca == 0.0 or cb == 0.0:
        return np.zeros_like(r)
    return ca * cb * r ** pa * z ** pb


KIND = {"diff1r": (1, 0), "diff1z": (0, 1), "diff2r": (2, 0), "diff2z": (0, 2), "diff2rz": (1, 1)}


def zpar_for(o):
    return [1, -1] if o.eqsym else [None]


@pytest.mark.parametrize("kind", ["diff1r", "diff1z", "diff2r", "diff2z"])
@pytest.mark.parametrize("symr", [1, -1])
def test_derivatives_exact_on_quartics(orc, kind, symr):
    r, z = orc.coords()
    for zp in zpar_for(orc):
        symz = zp if zp is not None else 1
        for a, b in monomials(4, symr, zp):
            f = r ** a * z ** b
            got = orc.stencil(kind, f, symr=symr, symz=symz)
            want = dpoly(a, b, r, z, *KIND[kind])
            scale = max(1.0, np.abs(want).max())
            assert np.abs(got - want).max() <= 1e-9 * scale, (kind, a, b)


@pytest.mark.parametrize("symr", [1, -1])
def test_mixed_derivative(orc, symr):
    """Fourth order inside, second-order edges and corners: exact on
    quartics inside and on biquadratics everywhere."""
    r, z = orc.coords()
    d = orc.dims()
    g, Nr, Nz = d["ghost"], d["Nr"], d["Nz"]
    ii = np.arange(1 - g, Nr + 1)[:, None] + 0 * np.arange(Nz + g)[None, :]
    jj = np.arange(1 - g, Nz + 1)[None, :] + 0 * ii
    inner = (ii >= 3 - g) & (ii <= Nr - 2) & (jj >= 3 - g) & (jj <= Nz - 2)
    if orc.eqsym:
        inner &= jj >= 1     # equatorial ghosts are images, checked below
    for zp in zpar_for(orc):
        symz = zp if zp is not None else 1
        for a, b in monomials(4, symr, zp):
            f = r ** a * z ** b
            got = orc.stencil("diff2rz", f, symr=symr, symz=symz)
            want = dpoly(a, b, r, z, 1, 1)
            scale = max(1.0, np.abs(want).max())
            assert np.abs(got - want)[inner].max() <= 1e-9 * scale, (a, b)
            if a <= 2 and b <= 2:
                assert np.abs(got - want).max() <= 1e-9 * scale, (a, b)


@pytest.mark.parametrize("kind", ["diffadvr", "diffadvz"])
def test_upwind_derivatives_exact_on_quartics(orc, kind):
    r, z = orc.coords()
    rng = np.random.default_rng(3)
    beta = rng.choice([-1.0, 0.0, 1.0], size=r.shape)
    sym = -1 if kind == "diffadvr" else 1
    for zp in zpar_for(orc):
        symz = zp if zp is not None else 1
        for a, b in monomials(4, 1, zp):
            f = r ** a * z ** b
            got = orc.stencil(kind, f, symr=1, symz=symz, beta=beta)
            want = dpoly(a, b, r, z, *((1, 0) if kind == "diffadvr" else (0, 1)))
            scale = max(1.0, np.abs(want).max())
            assert np.abs(got - want).max() <= 1e-9 * scale, (kind, a, b)
    del sym


def test_upwind_side(orc):
    """beta > 0 reads the points at larger r (derivadvect.f90:105-175):
    values below the stencil's reach do not change the result."""
    r, z = orc.coords()
    d = orc.dims()
    g, Nr = d["ghost"], d["Nr"]
    f = r ** 2 + 0.3 * z
    rng = np.random.default_rng(5)
    i0 = 6                                 # array row of reference index i0 + 1 - g
    f2 = f.copy()
    f2[:i0 - 1, :] = rng.normal(size=f2[:i0 - 1, :].shape)   # rows i < i0 - 1 scrambled
    beta = np.ones_like(r)
    a = orc.stencil("diffadvr", f, beta=beta)
    b = orc.stencil("diffadvr", f2, beta=beta)
    # rows i0 .. Nr-3 use i-1 .. i+3 only
    assert np.array_equal(a[i0:Nr - 3 + g - 1, :], b[i0:Nr - 3 + g - 1, :])


def test_upwind_z_row_three_minus_ghost(oaxis_mod, oracle_mod):
    """Without equatorial symmetry the row j = 3-ghost takes the right-biased
    form (j-1 .. j+3) for beta_z <= 0 as derivadvect.f90 writes it for that
    row, and the centred form for beta_z > 0: with every row below j = 2-ghost
    scrambled, the result there is unchanged for beta_z <= 0 and still exact
    on quartics (the stencil reaches no further than one row down)."""
    p = oaxis_mod.params_from_text(GRID % ".false.")
    o = oracle_mod.Oracle(p)
    r, z = o.coords()
    g = o.dims()["ghost"]
    row = 2                                 # array column of j = 3 - g
    f = r ** 2 * z ** 3 + z ** 4
    rng = np.random.default_rng(9)
    f2 = f.copy()
    f2[:, :row - 1] = rng.normal(size=f2[:, :row - 1].shape)   # j = 1-g scrambled
    beta = -np.ones_like(r)
    a = o.stencil("diffadvz", f, beta=beta)
    b = o.stencil("diffadvz", f2, beta=beta)
    assert np.array_equal(a[g:, row], b[g:, row])
    want = 3 * r ** 2 * z ** 2 + 4 * z ** 3
    assert np.abs(a[g:, row] - want[g:, row]).max() <= 1e-9 * np.abs(want).max()
    # beta_z > 0 at that row: the centred 4th-order form, which reads j-2
    c = o.stencil("diffadvz", f2, beta=-beta)
    assert not np.array_equal(c[g:, row], o.stencil("diffadvz", f, beta=-beta)[g:, row])
    o.close()
    del g


def test_ko_annihilates_quintics(orc):
    r, z = orc.coords()
    for a, b in monomials(5, 1):
        f = r ** a * z ** b
        src = orc.dissipation(f, np.zeros_like(f), symr=1, symz=1, diss=0.3)
        assert np.abs(src).max() <= 1e-7 * max(1.0, np.abs(f).max()), (a, b)


@pytest.mark.parametrize("theta", [np.pi, 2 * np.pi / 3, np.pi / 4])
def test_ko_symbol(orc, theta):
    d = orc.dims()
    g, Nr, Nz = d["ghost"], d["Nr"], d["Nz"]
    eps, dt = 0.3, d["dt"]
    i = np.arange(1 - g, Nr + 1)[:, None] * np.ones((1, Nz + g))
    f = np.cos(theta * i + 0.2)
    src = orc.dissipation(f, np.zeros_like(f), symr=1, symz=1, diss=eps)
    want = -(eps / (20.0 * dt)) * (2.0 - 2.0 * np.cos(theta)) ** 3 * f
    rows = slice(3, Nr - 3 + g)              # reference i = 4-g .. Nr-3
    assert np.allclose(src[rows], want[rows], rtol=0, atol=1e-12 * eps / dt)
    # no r-operator beyond, images of the interior on the axis ghosts
    assert np.abs(src[Nr - 3 + g:]).max() <= 1e-12 * eps / dt
    for k in range(1, g + 1):
        assert np.array_equal(src[g - k], src[g + k - 1])


def test_interp_exact_on_bicubics(oaxis_mod, oracle_mod):
    p = oaxis_mod.params_from_text(GRID % ".false.")
    o = oracle_mod.Oracle(p)
    r, z = o.coords()
    d = o.dims()
    rng = np.random.default_rng(11)
    for a, b in [(0, 0), (1, 2), (3, 3), (2, 1), (3, 0)]:
        f = r ** a * z ** b
        rmin, rmax = r.min(), r.max()
        zmin, zmax = z.min(), z.max()
        for _ in range(40):
            r0 = rng.uniform(rmin + 1.01 * d["dr"], rmax - 1.01 * d["dr"])
            z0 = rng.uniform(zmin + 1.01 * d["dz"], zmax - 1.01 * d["dz"])
            v = o.interp(f, r0, z0)
            assert abs(v - r0 ** a * z0 ** b) <= 1e-11 * max(1.0, abs(r0 ** a * z0 ** b)), (a, b, r0, z0)
    # near the edges the reference falls back to bilinear: exact on r*z
    f = 1.0 + 2.0 * r - 0.5 * z + 0.7 * r * z
    for _ in range(20):
        r0 = rng.uniform(r.max() - 0.99 * d["dr"], r.max())
        z0 = rng.uniform(z.min() + d["dz"], z.max() - d["dz"])
        assert abs(o.interp(f, r0, z0) - (1 + 2 * r0 - 0.5 * z0 + 0.7 * r0 * z0)) <= 1e-12
    # a grid node is copied
    assert o.interp(f, r[5, 7], z[5, 7]) == f[5, 7]
    o.close()
