"""Seeded generators of the benchmark workloads as reference par-file text.

Every workload is expressed with the reference's par keys so that the same
text drives the reference Fortran code, the oracle and this library.  Values
drawn at random come from numpy.random.default_rng(seed).

  plain_small   Schwarzschild puncture M=1, 128x256, dr0=0.05, RK4, 100 steps
  schwarzschild the same physics for any mass, grid and level count
  brill         Brill (Holz) wave a~U[1,4], 5 levels of 256x512, ICN,
                1+log + Gamma-driver 2, 512 coarse steps
  scalar_pulse  one member of the batched scalar-collapse family
                (A~U[.01,.05], r0~U[3,6], s~U[.5,1.5], 64x128..256x512,
                Nl0 in 1..3, RK4, 1+log + Gamma-driver 2)
"""
import numpy as np

PLAIN_SMALL = """
idata = schwarzschild
BH1mass = 1.0
eqsym = .false.
dr0 = 0.05
dz0 = 0.05
Nrtotal = 128
Nztotal = 256
Nl0 = 0
order = four
integrator = rk4
dtfac = 0.25
Nt = 100
slicing = 1+log
gauge_f = 2.0
ilapse = psiminus2
shift = none
geodiss = 0.1
boundtype = radiative
chimethod = .true.
ahfind = .false.
"""


def plain_small(seed=0, Nrtotal=128, Nztotal=256):
    text = PLAIN_SMALL.replace("Nrtotal = 128", "Nrtotal = %d" % Nrtotal)
    return text.replace("Nztotal = 256", "Nztotal = %d" % Nztotal)


def schwarzschild(mass, Nrtotal=128, Nztotal=256, Nl0=0, extent=6.4):
    """The plain configuration's physics with mass ``mass`` on any grid,
    rho in (0, extent] on the coarsest level."""
    text = plain_small(0, Nrtotal, Nztotal)
    dr = extent / Nrtotal
    text = text.replace("BH1mass = 1.0", "BH1mass = %.17g" % mass)
    text = text.replace("dr0 = 0.05", "dr0 = %.17g" % dr).replace("dz0 = 0.05", "dz0 = %.17g" % dr)
    return text.replace("Nl0 = 0", "Nl0 = %d" % Nl0)


def brill(seed=0, Nrtotal=256, Nztotal=512, Nl0=4, dr0=0.1):
    rng = np.random.default_rng(seed)
    a = rng.uniform(1.0, 4.0)
    return """
idata = BrillWave
brilltype = Holz
brill_a0 = %.17g
brill_sr0 = 1.0
brill_sz0 = 1.0
ELL_solver = wave
WE_epsilon = 1.d-8
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
drivercsi = 0.75
Nl0 = %d
Nrtotal = %d
Nztotal = %d
dr0 = %.17g
dz0 = %.17g
eqsym = .false.
integrator = icn
icniter = 3
order = four
dtfac = 0.25
Nt = 512
geodiss = 0.1
boundtype = radiative
""" % (a, Nl0, Nrtotal, Nztotal, dr0, dr0)


SCALAR_GRIDS = [(64, 128), (128, 256), (256, 512)]


def scalar_pulse(seed=0, index=0):
    """Member ``index`` of the batched scalar-collapse family for ``seed``."""
    rng = np.random.default_rng([seed, index])
    A = rng.uniform(0.01, 0.05)
    r0 = rng.uniform(3.0, 6.0)
    s = rng.uniform(0.5, 1.5)
    Nr, Nz = SCALAR_GRIDS[int(rng.integers(0, len(SCALAR_GRIDS)))]
    nl0 = int(rng.integers(1, 4))
    dr0 = 12.0 / Nr        # rho in (0, 12]: the pulse starts well inside
    return """
mattertype = scalar
scalarpotential = none
idata = scalarpulse
scalarpulse_type = shell
scalar_a0 = %.17g
scalar_r0 = %.17g
scalar_z0 = %.17g
scalar_sr0 = %.17g
scalar_sz0 = %.17g
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
drivercsi = 0.75
integrator = rk4
order = four
eqsym = .false.
Nrtotal = %d
Nztotal = %d
Nl0 = %d
dr0 = %.17g
dz0 = %.17g
dtfac = 0.25
geodiss = 0.1
scalardiss = 0.1
boundtype = radiative
""" % (A, r0, r0, s, s, Nr, Nz, nl0, dr0, dr0)


def collapse_pulse(amplitude, Nrtotal=32, Nztotal=32, Nl0=1, dr0=0.125):
    """A central massless scalar pulse phi = A exp(-(rho^2 + z^2)) with the
    Gamma-driver, equatorial symmetry, RK4, on 2 levels of 32x32 (dr0 =
    0.125) by default: the reduced bisection's member."""
    return """
mattertype = scalar
scalarpotential = none
idata = scalarpulse
scalarpulse_type = shell
scalar_a0 = %.17g
scalar_r0 = 0.0
scalar_z0 = 0.0
scalar_sr0 = 1.0
scalar_sz0 = 1.0
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
drivercsi = 0.75
integrator = rk4
order = four
eqsym = .true.
Nrtotal = %d
Nztotal = %d
Nl0 = %d
dr0 = %.17g
dz0 = %.17g
dtfac = 0.25
geodiss = 0.1
scalardiss = 0.1
boundtype = radiative
""" % (amplitude, Nrtotal, Nztotal, Nl0, dr0, dr0)


def bisection_pulse(amplitude, Nrtotal=256, Nztotal=512, Nl0=6, dr0=0.1):
    """A member of the critical-collapse bisection: a shell pulse of the real
    scalar field with r0 = 5, s = 1 and amplitude A, 1+log + Gamma-driver,
    RK4, on 7 box-in-box levels of 256x512 (rho in (0, 25.6] on the coarsest,
    no equatorial symmetry)."""
    return """
mattertype = scalar
scalarpotential = none
idata = scalarpulse
scalarpulse_type = shell
scalar_a0 = %.17g
scalar_r0 = 5.0
scalar_z0 = 5.0
scalar_sr0 = 1.0
scalar_sz0 = 1.0
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
drivercsi = 0.75
integrator = rk4
order = four
eqsym = .false.
Nrtotal = %d
Nztotal = %d
Nl0 = %d
dr0 = %.17g
dz0 = %.17g
dtfac = 0.25
geodiss = 0.1
scalardiss = 0.1
boundtype = radiative
""" % (amplitude, Nrtotal, Nztotal, Nl0, dr0, dr0)


def evolved_count(shift, scalar):
    """N of the metric definition: evolved fields the RHS reads and writes."""
    return 12 + (4 if shift else 0) + (4 if scalar else 0)
