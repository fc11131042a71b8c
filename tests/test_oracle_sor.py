"""Pins of the oracle's SOR solve of the Hamiltonian constraint
(ELL_solver = sor; oracle/or_elliptic.c sor_level / sor_elliptic, which
restate src/elliptic/sor_elliptic.f90).

The iteration's fixed point is rebuilt here from the continuum problem,
independently of the oracle's coefficient formulas:

  * interior points: lap u + S1 u = S0 with lap = d_rr + d_r / rho + d_zz,
    each derivative the standard 2nd-order centred difference;
  * the base level's outer and lower boundaries: the Robin condition
    rho u_rho + z u_z + u - 1 = 0, with the derivative normal to the
    boundary one-sided (2nd order, three points) and the tangential one
    centred, taking the neighbour before a point along the boundary as the
    value the boundary sweep has just written (the sweep runs through the
    axis ghost columns before the images overwrite them);
  * axis (and equator) images, even.

The linear system is assembled from these residuals (applied to unit
vectors) and solved directly; the converged SOR solution must match it.
A finer level is checked the same way with its boundary values (Dirichlet,
from the parent's interpolation) taken as given.  Zero amplitude: u = 1
solves the equation and the Robin condition exactly, so the first
iteration already meets the tolerance.  Not pinned beyond the converged
solution: the Chebyshev schedule of omega and the residual norm, which
decide the iteration count only.  (The reference's norm adds the second
pass's squared residuals to the first pass's already rooted norm, so it
levels off near sqrt(sqrt(S/N)/N): ELL_epsilon = 1e-8 on these grids means
a true RMS residual near 1e-13; a much smaller ELL_epsilon is never met and
the loop runs to ELL_maxiter, in the reference as here.)
"""
import numpy as np

BRILL = """
order = four
idata = BrillWave
ELL_solver = sor
ELL_epsilon = %g
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


def solved(oaxis_mod, oracle_mod, text):
    o = oracle_mod.Oracle(oaxis_mod.params_from_text(text))
    o.initial_data()
    return o


def test_zero_amplitude_stops_at_the_first_iteration(oaxis_mod, oracle_mod):
    o = solved(oaxis_mod, oracle_mod, BRILL % (1e-10, 0.0, 0.3, 0.3, 16, 32, 0))
    st = o.ell_stats()
    assert st["steps0"] == 1 and st["steps1"] == 0, st
    assert np.all(o.view("ell_u", 0) == 1.0)
    o.close()


def _grid(o, lev):
    d = o.dims(lev)
    r, z = o.coords(lev)
    return d["Nr"], d["Nz"], d["ghost"], d["dr"], d["dz"], r[:, 0].copy(), z[0, :].copy()


def _direct_base(o, s1, s0):
    """The fixed point of the base level's iteration (Robin boundaries, no
    equatorial symmetry), solved directly."""
    Nr, Nz, g, dr, dz, rr, zz = _grid(o, 0)
    ni, nj = Nr + g, Nz + g
    nghost = g - 1                       # axis ghost columns the boundary rows sweep: i = 2-g .. 0
    nu = ni * nj
    n = nu + 2 * nghost

    def I(i):
        return i + g - 1

    def J(j):
        return j + g - 1

    def residual(x):
        u = x[:nu].reshape(ni, nj)
        Rup = {i: x[nu + (i - (2 - g))] for i in range(2 - g, 1)}
        Rlo = {i: x[nu + nghost + (i - (2 - g))] for i in range(2 - g, 1)}
        U = lambda i, j: u[I(i), J(j)]
        res = np.zeros(n)

        def lap(i, j):
            r = rr[I(i)]
            d_rr = (U(i + 1, j) - 2.0 * U(i, j) + U(i - 1, j)) / dr ** 2
            d_r = (U(i + 1, j) - U(i - 1, j)) / (2.0 * dr)
            d_zz = (U(i, j + 1) - 2.0 * U(i, j) + U(i, j - 1)) / dz ** 2
            return d_rr + d_r / r + d_zz

        def robin(r, z, ur, uz, u0):
            return r * ur + z * uz + u0 - 1.0

        def row_robin(i, jb, inward, prev, after, u0):
            # boundary row jb (upper: inward = -1, lower: inward = +1)
            uz = -inward * (3.0 * u0 - 4.0 * U(i, jb + inward) + U(i, jb + 2 * inward)) / (2.0 * dz)
            ur = (after - prev) / (2.0 * dr)
            return robin(rr[I(i)], zz[J(jb)], ur, uz, u0)

        for i in range(1 - g, Nr + 1):
            for j in range(1 - g, Nz + 1):
                q = I(i) * nj + J(j)
                if i <= 0:
                    res[q] = U(i, j) - U(1 - i, j)
                elif i == Nr and j in (Nz, 1 - g):
                    inward = -1 if j == Nz else 1
                    ur = (3.0 * U(Nr, j) - 4.0 * U(Nr - 1, j) + U(Nr - 2, j)) / (2.0 * dr)
                    uz = -inward * (3.0 * U(Nr, j) - 4.0 * U(Nr, j + inward) + U(Nr, j + 2 * inward)) / (2.0 * dz)
                    res[q] = robin(rr[I(Nr)], zz[J(j)], ur, uz, U(Nr, j))
                elif i == Nr:
                    ur = (3.0 * U(Nr, j) - 4.0 * U(Nr - 1, j) + U(Nr - 2, j)) / (2.0 * dr)
                    uz = (U(Nr, j + 1) - U(Nr, j - 1)) / (2.0 * dz)
                    res[q] = robin(rr[I(Nr)], zz[J(j)], ur, uz, U(Nr, j))
                elif j == Nz or j == 1 - g:
                    R = Rup if j == Nz else Rlo
                    inward = -1 if j == Nz else 1
                    prev = R[0] if i == 1 else U(i - 1, j)
                    res[q] = row_robin(i, j, inward, prev, U(i + 1, j), U(i, j))
                else:
                    res[q] = lap(i, j) + s1[I(i), J(j)] * U(i, j) - s0[I(i), J(j)]
        # the values the boundary rows write into the axis ghost columns
        for R, jb, inward, off in ((Rup, Nz, -1, nu), (Rlo, 1 - g, 1, nu + nghost)):
            for i in range(2 - g, 1):
                prev = U(1 - g, jb) if i == 2 - g else R[i - 1]
                after = U(i + 1, jb)
                res[off + (i - (2 - g))] = row_robin(i, jb, inward, prev, after, R[i])
        return res

    b = -residual(np.zeros(n))
    A = np.empty((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = residual(e) + b
    x = np.linalg.solve(A, b)
    return x[:nu].reshape(ni, nj)


def test_base_level_converges_to_the_discrete_robin_problem(oaxis_mod, oracle_mod):
    o = solved(oaxis_mod, oracle_mod, BRILL % (1e-8, 2.0, 0.3, 0.3, 16, 32, 0))
    st = o.ell_stats()
    print("SOR iterations %s" % st)
    assert 1 < st["steps0"] < 100000 and st["residual"] < 1e-8
    u = o.get("ell_u", 0)
    ud = _direct_base(o, o.get("ell_S1", 0), o.get("ell_S0", 0))
    err = np.abs(u - ud).max()
    print("max |u_sor - u_direct| = %.3e, u in [%.4f, %.4f]" % (err, u.min(), u.max()))
    assert u.min() < 0.999                       # a solution away from the guess
    assert err < 1e-10
    o.close()


def test_fine_level_converges_to_the_dirichlet_problem(oaxis_mod, oracle_mod):
    o = solved(oaxis_mod, oracle_mod, BRILL % (1e-8, 2.0, 0.4, 0.4, 16, 32, 1))
    st = o.ell_stats()
    assert st["steps1"] > 1, st
    Nr, Nz, g, dr, dz, rr, zz = _grid(o, 1)
    u = o.get("ell_u", 1)
    s1 = o.get("ell_S1", 1)
    ni, nj = Nr + g, Nz + g
    n = ni * nj
    A = np.zeros((n, n))
    b = np.zeros(n)
    for i in range(1 - g, Nr + 1):
        for j in range(1 - g, Nz + 1):
            q = (i + g - 1) * nj + (j + g - 1)
            col = lambda ii, jj: (ii + g - 1) * nj + (jj + g - 1)
            if i <= 0:
                A[q, q] = 1.0
                A[q, col(1 - i, j)] -= 1.0
            elif i == Nr or j == Nz or j == 1 - g:
                A[q, q] = 1.0
                b[q] = u[i + g - 1, j + g - 1]           # boundary values as the solve left them
            else:
                r = rr[i + g - 1]
                # d_rr + d_r / r + d_zz + S1
                A[q, col(i + 1, j)] += 1.0 / dr ** 2 + 1.0 / (2.0 * dr * r)
                A[q, col(i - 1, j)] += 1.0 / dr ** 2 - 1.0 / (2.0 * dr * r)
                A[q, col(i, j + 1)] += 1.0 / dz ** 2
                A[q, col(i, j - 1)] += 1.0 / dz ** 2
                A[q, q] += -2.0 / dr ** 2 - 2.0 / dz ** 2 + s1[i + g - 1, j + g - 1]
    ud = np.linalg.solve(A, b).reshape(ni, nj)
    err = np.abs(u - ud).max()
    print("fine level: max |u_sor - u_direct| = %.3e" % err)
    assert err < 1e-10
    o.close()
