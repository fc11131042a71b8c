"""GPU parity: the HIP evolution hot path against the oracle, 1e-10 relative
max norm on every evolved array of every level (plus the derived KT_H,
lambda, Alambda), after several coarse steps.  Cases cover the axis and
equator parity, both integrators, box-in-box levels (fine-boundary time
interpolation of both orders, restriction), the Gamma-driver shift with and
without upwinded advection, and the real scalar field with its potentials."""
import pytest

from parity_util import TOL, run_pair, worst

pytestmark = pytest.mark.gpu

BASE = """
order = four
dtfac = 0.25
boundtype = radiative
geodiss = 0.1
"""

CASES = {
    "schwarzschild_rk4_noeqsym": (BASE + """
idata = schwarzschild
BH1mass = 1.0
ilapse = psiminus2
slicing = 1+log
gauge_f = 2.0
integrator = rk4
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 24
Nztotal = 48
""", 4),
    "schwarzschild_rk4_eqsym": (BASE + """
idata = schwarzschild
BH1mass = 1.0
ilapse = psiminus2
slicing = 1+log
gauge_f = 2.0
integrator = rk4
eqsym = .true.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 24
Nztotal = 24
""", 4),
    "schwarzschild_2levels_rk4": (BASE + """
idata = schwarzschild
BH1mass = 1.0
ilapse = psiminus2
slicing = 1+log
gauge_f = 2.0
integrator = rk4
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 24
Nztotal = 48
Nl0 = 1
""", 3),
    "schwarzschild_3levels_icn_eqsym": (BASE + """
idata = schwarzschild
BH1mass = 1.0
ilapse = psiminus2
slicing = harmonic
gauge_f = 1.0
integrator = icn
eqsym = .true.
dr0 = 0.2
dz0 = 0.2
Nrtotal = 20
Nztotal = 20
Nl0 = 2
""", 2),
    "gauge_pulse_gammadriver2": (BASE + """
idata = minkowski
lapsepert = gauss
ilapse = shell
lapse_a0 = 0.1
lapse_sr0 = 1.0
lapse_sz0 = 1.0
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
drivercsi = 0.75
drivereta = 0.5
integrator = rk4
eqsym = .true.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 24
Nztotal = 24
Nl0 = 1
""", 3),
    "gammadriver2_advect_noeqsym": (BASE + """
idata = schwarzschild
BH1mass = 1.0
ilapse = psiminus2
slicing = shockavoid
gauge_f = 2.0
shift = Gammadriver2
driveradv = .true.
shiftadvect = .false.
integrator = rk4
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 20
Nztotal = 40
""", 3),
    "scalar_pulse_gd2_levels": (BASE + """
mattertype = scalar
idata = scalarpulse
scalarpulse_type = shell
scalar_a0 = 0.05
scalar_r0 = 1.0
scalar_z0 = 1.0
scalar_sr0 = 0.5
scalar_sz0 = 0.5
scalardiss = 0.05
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
integrator = rk4
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 20
Nztotal = 40
Nl0 = 1
""", 3),
    "brill_gd2_odd_sizes_icn": (BASE + """
idata = BrillWave
brill_a0 = 2.0
brill_sr0 = 1.0
brill_sz0 = 1.0
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
integrator = icn
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 21
Nztotal = 62
Nl0 = 1
""", 3),
    "scalar_phi2_first_icn": (BASE + """
mattertype = scalar
idata = scalarpulse
scalarpulse_type = torus
scalarpotential = phi2
scalar_mass = 1.0
scalarmethod = first
scalar_a0 = 0.02
scalar_r0 = 1.0
scalar_z0 = 0.5
scalar_sr0 = 0.5
scalar_sz0 = 0.5
slicing = 1+log
gauge_f = 2.0
integrator = icn
eqsym = .true.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 24
Nztotal = 24
""", 3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_parity_with_oracle(case, oaxis_mod, oracle_mod):
    text, nsteps = CASES[case]
    errs0, errs, times = run_pair(oaxis_mod, oracle_mod, text, nsteps)
    w0, k0 = worst(errs0)
    w, k = worst(errs)
    print("%s: initial worst %.3e at %s; after %d steps worst %.3e at %s" % (case, w0, k0, nsteps, w, k))
    for (tg, to) in times:
        assert tg == to, "level clocks differ: %s vs %s" % (tg, to)
    assert w0 <= TOL, "initial auxiliary differs: %.3e at %s" % (w0, k0)
    assert w <= TOL, "evolution differs: %.3e at %s" % (w, k)
