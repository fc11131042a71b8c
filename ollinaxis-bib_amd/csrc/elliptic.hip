// The WaveElliptic solve of the Hamiltonian constraint for Brill-wave and
// scalar-pulse initial data, on the device: wave_elliptic('flat', 'one')
// (src/elliptic/wave_elliptic.f90:2-532) and its recursive step wavestep
// (:542-1164), for box 0.
//
// The equation  lap u + S1 u + S5 u^5 = S0  (flat Laplacian in (rho, z)) is
// relaxed as  du/dt = v,  dv/dt = lap u + S1 u + S5 u^5 - S0 - WE_eta v,
// with Kreiss-Oliger dissipation and radiative boundaries on v, until the
// RMS of dv/dt on the base level drops below WE_epsilon (at least 100
// steps, at most WE_maxiter).  With refinement levels the base level is
// solved alone first, its solution is interpolated onto every finer level,
// and all levels are relaxed together with Berger-Oliger subcycling.
//
// Execution: every piece of a coarse step is a kernel whose arguments do not
// change from step to step -- the level clocks, the time-interpolation
// coefficients and the finebound depth switch (tp == tl) are formed on the
// device from clocks kept in device memory -- so one coarse step is captured
// once as a HIP graph and replayed.  The stopping rule is evaluated on the
// device at the end of each coarse step; once it fires, every kernel of the
// later replays returns at once, so the host may run ahead by a whole chunk
// of steps and reads the outcome after each chunk.
//
// ELL_solver = sor selects the reference's other solver, sor_elliptic('flat',
// 'one') (src/elliptic/sor_elliptic.f90): red-black successive
// over-relaxation with Chebyshev acceleration, level by level (Robin
// boundaries on the base level, Dirichlet from the parent's solution on the
// finer ones), then restriction.  One iteration -- two colour sweeps, their
// norms, the boundary recurrences, the symmetries and the stopping rule --
// is captured and replayed the same way.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "oa_state.h"
#define OA_HD __device__ __forceinline__
#define OA_HHD __host__ __device__ inline
#include "oaxis_device.h"

namespace oax {

namespace {

constexpr int kBlk = 256;

// Device clocks of one level (onestep.f90:351-357 as wavestep :997-1004 keeps them).
struct EllClock {
  double t, t1, t2;
  int s;
};

// Outcome of the solve so far, evaluated after every coarse step.
constexpr int kEllMaxLevels = 16;
struct EllCtl {
  int step;        // coarse steps of the current phase
  int done;        // the stopping rule fired: later kernels return at once
  int stationary;  // 1: the all-levels phase may end on the stationary rule
  double gres;     // residual of the last step
  double wmax, wprev;   // stationary rule: window maxima of the relative change of u
  // a level's clock advances at the end of its step (:997-1004).  The last
  // iteration of the step marks it here, and the next saveold launched (of
  // any level) applies it: every kernel that reads a level's clock (the
  // iterations and finebounds of that level and of the next finer one)
  // comes after a saveold, and no other kernel reads the clocks.
  int tick[kEllMaxLevels];
  double tick_dt[kEllMaxLevels];
};

// One level's planes for the solve.  u, v: the state at the end of a step
// (also the input of the first iteration); up, vp: the state at the start of
// the step (saveold); t1/t2: outputs of the intermediate iterations; ua, va:
// RK4 accumulators; s0, s1, s5: the coefficients; sv: the last source of v
// (base level, for the residual).  Strips: 3 time levels of the fine-boundary
// zone for u and v, in the layout of saveold_point (rR, then zL, then zR per
// field), time level k at strip + k * strip_tl.
struct EllPlanes {
  double *u, *v, *up, *vp, *t1u, *t1v, *t2u, *t2v, *ua, *va, *s0, *s1, *s5, *sv;
  double* strip;
  long strip_tl;        // doubles of one time level (both fields)
};
constexpr int kEllPlanes = 14;

struct EllLevelArgs {
  KGrid G;
  EllPlanes P;
  EllClock* clk;
};

__device__ __forceinline__ double pow5(double x) {
  const double x2 = x * x;
  return x2 * x2 * x;
}

// ---------------------------------------------------------------- setup
// coefficients of the equation at every point (ghosts included):
// Brill waves idata_brillwave.f90:88-105 (S0 = S5 = 0), the time-symmetric
// scalar pulse idata_scalarpulse.f90:143-166 (smallpi = acos(-1d0))
__global__ void __launch_bounds__(kBlk) k_ell_coeffs(oa_params p, KGrid G, const double* __restrict__ slab,
                                                     int sphi, int sxr, int sxz, EllPlanes E) {
  const long q = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x;
  const long n = static_cast<long>(G.Nr + G.g) * (G.Nz + G.g);
  if (q >= n) return;
  const int nzt = G.Nz + G.g;
  const int i = 1 - G.g + static_cast<int>(q / nzt), j = 1 - G.g + static_cast<int>(q % nzt);
  const long k = IDX(G, i, j);
  const double r = coord_r(G, i), z = coord_z(G, j);
  double s0 = 0.0, s1 = 0.0, s5 = 0.0;
  if (p.idata == OA_IDATA_BRILLWAVE) {
    const double sr = p.brill_sr0, sz = p.brill_sz0;
    const double q0 = p.brill_a0 * r * r * exp(-(r / sr) * (r / sr) - (z / sz) * (z / sz));
    s1 = 0.5 * q0 * (1.0 / (r * r) + 2.0 * ((r / (sr * sr)) * (r / (sr * sr)) + (z / (sz * sz)) * (z / (sz * sz)))
                     - 5.0 / (sr * sr) - 1.0 / (sz * sz));
  } else {
    const double smallpi = M_PI;
    const double xr = slab[sxr * G.plane + k], xz = slab[sxz * G.plane + k];
    s1 = smallpi * (xr * xr + xz * xz);
    if (p.scalarpotential != OA_POT_NONE) {
      const double ph = slab[sphi * G.plane + k];
      double V = 0.5 * p.scalar_mass * p.scalar_mass * ph * ph;
      if (p.scalarpotential == OA_POT_PHI4) V = V + 0.25 * p.scalar_lambda * ph * ph * ph * ph;
      s5 = 2.0 * smallpi * V;
    }
  }
  E.s0[k] = s0;
  E.s1[k] = s1;
  E.s5[k] = s5;
}

// initial guess u = 1, v = 0 everywhere (:107-141); strips and the other
// planes cleared; clock and control reset
__global__ void __launch_bounds__(kBlk) k_ell_init(KGrid G, EllPlanes E, EllClock* clk) {
  const long q = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x;
  if (q == 0) *clk = EllClock{0.0, 0.0, 0.0, 0};
  if (q < G.plane) {
    E.u[q] = 1.0; E.v[q] = 0.0; E.up[q] = 1.0; E.vp[q] = 0.0;
    E.t1u[q] = 0.0; E.t1v[q] = 0.0; E.t2u[q] = 0.0; E.t2v[q] = 0.0;
    E.ua[q] = 0.0; E.va[q] = 0.0; E.sv[q] = 0.0;
  }
  if (q < 3 * E.strip_tl) E.strip[q] = 0.0;
}

__global__ void k_ell_reset(EllCtl* ctl, int stationary, EllClock* clocks, int nlev) {
  if (threadIdx.x == 0) *ctl = EllCtl{0, 0, stationary, 1.0, 0.0, 0.0, {}, {}};
  if (static_cast<int>(threadIdx.x) < nlev) clocks[threadIdx.x] = EllClock{0.0, 0.0, 0.0, 0};
}

// ------------------------------------------------------------ one step
// saveold of wavestep (:631-679): up = u, vp = v over the whole level, and
// the strips shift one time level back (t-2 <- t-1 <- t0 <- now)
__global__ void __launch_bounds__(kBlk) k_ell_saveold(EllLevelArgs a, int fine, EllCtl* ctl, EllClock* clocks,
                                                     int nlev) {
  if (ctl->done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int l = 0; l < nlev; ++l)
      if (ctl->tick[l]) {                                     // the level clock's step (:997-1004)
        EllClock c = clocks[l];
        c.t2 = c.t1;
        c.t1 = c.t;
        c.s = c.s + 1;
        c.t = c.t + ctl->tick_dt[l];
        clocks[l] = c;
        ctl->tick[l] = 0;
      }
  const KGrid& G = a.G;
  const EllPlanes& E = a.P;
  const long q = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x;
  if (q < G.plane) {
    E.up[q] = E.u[q];
    E.vp[q] = E.v[q];
  }
  if (!fine) return;
  const int nrt = G.Nr + G.g, nzt = G.Nz + G.g;
  const long per_r = static_cast<long>(G.g) * nzt, per_z = static_cast<long>(G.g) * nrt;
  const long per_f = per_r + 2 * per_z;
  if (q >= 2 * per_f) return;
  const int f = static_cast<int>(q / per_f);
  const long e = q % per_f;
  const double* y = (f == 0) ? E.u : E.v;
  double now;
  if (e < per_r) {
    const int m = static_cast<int>(e / nzt), jj = static_cast<int>(e % nzt);
    now = y[IDX(G, G.Nr - m, jj + 1 - G.g)];
  } else if (e < per_r + per_z) {
    const long e2 = e - per_r;
    const int n = static_cast<int>(e2 / nrt), ii = static_cast<int>(e2 % nrt);
    now = y[IDX(G, ii + 1 - G.g, 1 - G.g + n)];
  } else {
    const long e2 = e - per_r - per_z;
    const int n = static_cast<int>(e2 / nrt), ii = static_cast<int>(e2 % nrt);
    now = y[IDX(G, ii + 1 - G.g, G.Nz - n)];
  }
  double* s = E.strip + q;
  s[2 * E.strip_tl] = s[E.strip_tl];
  s[E.strip_tl] = s[0];
  s[0] = now;
}

struct EllIterArgs {
  EllLevelArgs L;
  const double* uin; const double* vin;
  double* uout; double* vout;
  double dtw, w, idt, eta, diss;
  int iter, niter, rk4, base;
  int level;
  int s5;                            // S5 is not identically zero (a scalar potential)
};

// the last iteration of a level's step marks its clock for advancing
__device__ __forceinline__ void ell_mark_tick(const EllIterArgs& a, EllCtl* ctl) {
  if (a.iter == a.niter && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    ctl->tick_dt[a.level] = a.L.G.dt;
    ctl->tick[a.level] = 1;
  }
}

// One internal iteration of one level (:691-915): sources of u and v, the
// Kreiss-Oliger terms on v, the radiative boundary of v, the RK4
// accumulators and the update, at every computed point with its parity
// images (both fields are even on the axis and the equator).
__device__ __forceinline__ void ell_iter_point(const EllIterArgs& a, int i, int j) {
  const KGrid& G = a.L.G;
  const EllPlanes& E = a.L.P;
  const long k = IDX(G, i, j);
  const Plane U{a.uin, G}, V{a.vin, G};
  const double r = coord_r(G, i), z = coord_z(G, j);
  const double u = U(i, j), v = V(i, j);
  double su = v;
  // flat Laplacian, the coefficients and the damping (:791-851)
  const double Dr_u = s_d1r(U, G, i, j), Drr_u = s_d2r(U, G, i, j), Dzz_u = s_d2z(U, G, i, j);
  double sv = Drr_u + Dzz_u + Dr_u / r;
  // S0 is zero for every supported initial data (subtracting +0 changes
  // nothing) and S5 without a scalar potential (adding a zero changes the
  // value of nothing either): their planes are read only when they matter
  sv = sv + E.s1[k] * u;
  if (a.s5) sv = sv + E.s5[k] * pow5(u);
  sv = sv - a.eta * v;
  // dissipation on v (dissipation.f90:181-199): rho terms for i <= Nr-3,
  // z terms for 1 <= j <= Nz-3
  if (i <= G.Nr - 3)
    sv = sv - a.diss * a.idt * (20.0 * V(i, j) - 15.0 * (V(i + 1, j) + V(i - 1, j)) + 6.0 * (V(i + 2, j) + V(i - 2, j))
                                - (V(i + 3, j) + V(i - 3, j))) / 20.0;
  if (j >= 1 && j <= G.Nz - 3)
    sv = sv - a.diss * a.idt * (20.0 * V(i, j) - 15.0 * (V(i, j + 1) + V(i, j - 1)) + 6.0 * (V(i, j + 2) + V(i, j - 2))
                                - (V(i, j + 3) + V(i, j - 3))) / 20.0;
  // radiative boundary on v with v = 1, v0 = 0 (radbound.f90:44-70)
  if (i == G.Nr || j == G.Nz || (!G.eqsym && j == 1 - G.g)) {
    const double rr = sqrt(r * r + z * z);
    sv = -1.0 * (r * s_d1r(V, G, i, j) + z * s_d1z(V, G, i, j) + (v - 0.0)) / rr;
  }
  // RK4 accumulators (:895-907)
  if (a.rk4) {
    if (a.iter == 1) {
      E.ua[k] = a.w * su;
      E.va[k] = a.w * sv;
    } else if (a.iter < a.niter) {
      E.ua[k] = E.ua[k] + a.w * su;
      E.va[k] = E.va[k] + a.w * sv;
    } else {
      su = E.ua[k] + a.w * su;
      sv = E.va[k] + a.w * sv;
    }
  }
  // update (:914-915)
  put(a.uout, G, i, j, E.up[k] + a.dtw * su, 1.0, 1.0);
  put(a.vout, G, i, j, E.vp[k] + a.dtw * sv, 1.0, 1.0);
  if (a.base && a.iter == a.niter) E.sv[k] = sv;
}

// the base level's iteration (no fine boundary)
__global__ void __launch_bounds__(kBlk) k_ell_iter(EllIterArgs a, EllCtl* ctl) {
  if (ctl->done) return;
  ell_mark_tick(a, ctl);
  const KGrid& G = a.L.G;
  const int j = G.jlo + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = 1 + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
  if (j > G.Nz || i > G.Nr) return;
  ell_iter_point(a, i, j);
}

struct EllFbArgs {
  KGrid Gf, Gc;
  EllPlanes F;                       // fine level
  const double* uc; const double* vc;   // coarse state (already advanced)
  double* uout; double* vout;           // the fine iteration's output
  const FbPoint* tab[2]; long n[2];  // zones of depth g+2 (tp == tl) and g-1
  const EllClock* cclk; const EllClock* fclk;
  double dtw;
};

// finebound(box, level, dtw, .false.) for u, then v (finebound.f90:70-400,
// single-variable branches): the depth and the order of the time
// interpolation come from the device clocks as the reference forms them.
__device__ __forceinline__ void ell_fb_point(const EllFbArgs& a, long q) {
  const double tp = a.cclk->t, t0 = a.fclk->t, tm1 = a.fclk->t1, tl = t0 + a.dtw;
  const int tab = (tp == tl) ? 0 : 1;
  if (q >= a.n[tab]) return;
  const FbPoint fp = a.tab[tab][q];
  if (fp.which < 0) return;
  const KGrid& Gf = a.Gf;
  const int g = Gf.g, i = fp.i, j = fp.j;
  const long k = IDX(Gf, i, j);
  const int border = (tm1 == 0.0) ? 1 : 2;
  const long per_r = static_cast<long>(g) * (Gf.Nz + g), per_z = static_cast<long>(g) * (Gf.Nr + g);
  const long per_f = per_r + 2 * per_z;
  const long off = (fp.which == 0) ? fp.sidx : (fp.which == 2) ? per_r + fp.sidx : per_r + per_z + fp.sidx;
  const bool inside = fp.strip_ok != 0;
  for (int f = 0; f < 2; ++f) {
    const double aux2 = apply(fp.c, f == 0 ? a.uc : a.vc, a.Gc);
    double val;
    if (border == 1) {
      const double fpv = (f == 0) ? a.F.up[k] : a.F.vp[k];
      val = (tl - t0) / (tp - t0) * aux2 + (tl - tp) / (t0 - tp) * fpv;
    } else {
      const double b0 = inside ? a.F.strip[f * per_f + off] : 0.0;
      const double b1 = inside ? a.F.strip[a.F.strip_tl + f * per_f + off] : 0.0;
      val = (tl - t0) * (tl - tm1) / ((tp - t0) * (tp - tm1)) * aux2 +
            (tl - tp) * (tl - tm1) / ((t0 - tp) * (t0 - tm1)) * b0 +
            (tl - tp) * (tl - t0) / ((tm1 - tp) * (tm1 - t0)) * b1;
    }
    put(f == 0 ? a.uout : a.vout, Gf, i, j, val, 1.0, 1.0);
  }
}

// One iteration of a fine level and its finebound in one launch.  The
// finebound overwrites the zone of depth D (g+2 when the clocks meet, else
// g-1: the rows i >= Nr-D, j >= Nz-D and, without equatorial symmetry,
// j <= 1-g+D) of the iteration's output, point and parity images alike, and
// reads nothing the iteration writes; so the iteration's blocks skip the
// zone and the finebound's blocks fill it, with the same final values as
// the two launches in sequence.
struct EllStepArgs {
  EllIterArgs it;
  EllFbArgs fb;
  int nbx, nb_iter;                  // iteration blocks: nbx along z (64 points), rows of 4 along rho
};

__global__ void __launch_bounds__(kBlk) k_ell_iter_fb(EllStepArgs a, EllCtl* ctl) {
  if (ctl->done) return;
  ell_mark_tick(a.it, ctl);
  const int b = static_cast<int>(blockIdx.x);
  if (b < a.nb_iter) {
    const KGrid& G = a.it.L.G;
    const int j = G.jlo + (b % a.nbx) * 64 + static_cast<int>(threadIdx.x & 63u);
    const int i = 1 + (b / a.nbx) * 4 + static_cast<int>(threadIdx.x >> 6);
    if (j > G.Nz || i > G.Nr) return;
    const int g = G.g;
    const double tp = a.fb.cclk->t, tl = a.fb.fclk->t + a.fb.dtw;
    const int D = (tp == tl) ? g + 2 : g - 1;
    if (i >= G.Nr - D || j >= G.Nz - D || (!G.eqsym && j <= 1 - g + D)) return;
    ell_iter_point(a.it, i, j);
  } else {
    ell_fb_point(a.fb, static_cast<long>(b - a.nb_iter) * kBlk + threadIdx.x);
  }
}


struct EllRsArgs {
  KGrid Gf, Gc;
  const double* uf; const double* vf;
  double* uc; double* vc;
  int ic0, nic, jc0, njc, interporder;
  int images;            // also write the coarse level's parity images
};

// restrict(box, level, .false.) for u and v (restrict.f90:88-162); with
// images, the coarse level's symmetries that wavestep applies right after
// it (:1136-1149) as images of each written point (the SOR solve applies
// none)
__global__ void __launch_bounds__(kBlk) k_ell_restrict(EllRsArgs a, const EllCtl* __restrict__ ctl) {
  if (ctl->done) return;
  const long q = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x;
  if (q >= static_cast<long>(a.nic) * a.njc) return;
  const int ic = a.ic0 + static_cast<int>(q / a.njc), jc = a.jc0 + static_cast<int>(q % a.njc);
  const KGrid& Gf = a.Gf;
  const KGrid& Gc = a.Gc;
  const int g = Gc.g;
  const double r0 = coord_r(Gc, 1 - g) + static_cast<double>(ic + g - 1) * Gc.dr;
  const double z0 = coord_z(Gc, 1 - g) + static_cast<double>(jc + g - 1) * Gc.dz;
  const double deltar = static_cast<double>(2 * g) * Gf.dr + 1.0e-10;
  const double deltaz = static_cast<double>(2 * g) * Gf.dz + 1.0e-10;
  const double rminf = coord_r(Gf, 1 - g), rmaxf = coord_r(Gf, Gf.Nr);
  const double zminf = coord_z(Gf, 1 - g), zmaxf = coord_z(Gf, Gf.Nz);
  const double rmin = (rminf < 0.0) ? 0.0 : rminf + deltar;
  const double zmin = (Gf.eqsym && zminf < 0.0) ? 0.0 : zminf + deltaz;
  const double rmax = rmaxf - deltar, zmax = zmaxf - deltaz;
  if (!(r0 > rmin && r0 < rmax && z0 > zmin && z0 < zmax)) return;
  const Cell c = locate(Gf, r0, z0, a.interporder);
  if (a.images) {
    put(a.uc, Gc, ic, jc, apply(c, a.uf, Gf), 1.0, 1.0);
    put(a.vc, Gc, ic, jc, apply(c, a.vf, Gf), 1.0, 1.0);
  } else {
    a.uc[IDX(Gc, ic, jc)] = apply(c, a.uf, Gf);
  }
}

// Residual and change of the base level after a coarse step, per block:
// sum of sv^2 (i in [1, Nr-g], j in [1, Nz-g], wave_elliptic.f90:321-341),
// max |u - up| and max |u| over the same range.
__global__ void __launch_bounds__(kBlk) k_ell_check(KGrid G, EllPlanes E, double* __restrict__ part,
                                                    const EllCtl* __restrict__ ctl) {
  if (ctl->done) return;
  __shared__ double ssq[kBlk], sdu[kBlk], smu[kBlk];
  const int ni = G.Nr - G.g, nj = G.Nz - G.g;
  const long n = static_cast<long>(ni) * nj;
  double sq = 0.0, du = 0.0, mu = 0.0;
  for (long q = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x; q < n; q += static_cast<long>(gridDim.x) * kBlk) {
    const int i = 1 + static_cast<int>(q / nj), j = 1 + static_cast<int>(q % nj);
    const long k = IDX(G, i, j);
    const double s = E.sv[k];
    sq += s * s;
    du = fmax(du, fabs(E.u[k] - E.up[k]));
    mu = fmax(mu, fabs(E.u[k]));
  }
  const int t = static_cast<int>(threadIdx.x);
  ssq[t] = sq; sdu[t] = du; smu[t] = mu;
  __syncthreads();
  for (int h = kBlk / 2; h > 0; h >>= 1) {
    if (t < h) {
      ssq[t] += ssq[t + h];
      sdu[t] = fmax(sdu[t], sdu[t + h]);
      smu[t] = fmax(smu[t], smu[t + h]);
    }
    __syncthreads();
  }
  if (t == 0) {
    part[3 * blockIdx.x] = ssq[0];
    part[3 * blockIdx.x + 1] = sdu[0];
    part[3 * blockIdx.x + 2] = smu[0];
  }
}

// The stopping rule after a coarse step.  Reference (:303): continue while
// step < 100 or (gres > WE_epsilon and step < WE_maxiter) -- a NaN residual
// stops.  Stationary rule (all-levels phase only, when enabled): stop once
// the window maximum over 100 steps of max|u - up| / max|u| is at most
// 1e-12 and above half the previous window's.
__global__ void k_ell_decide(int nparts, const double* __restrict__ part, double nntot, double eps, int maxiter,
                             EllCtl* ctl) {
  if (threadIdx.x != 0 || ctl->done) return;
  double sq = 0.0, du = 0.0, mu = 0.0;
  for (int b = 0; b < nparts; ++b) {
    sq += part[3 * b];
    du = fmax(du, part[3 * b + 1]);
    mu = fmax(mu, part[3 * b + 2]);
  }
  EllCtl c = *ctl;
  c.step += 1;
  c.gres = sqrt(sq / nntot);
  const bool more = (c.step < 100) || (c.gres > eps && c.step < maxiter);
  int done = more ? 0 : 1;
  if (!done && c.stationary) {
    const double rel = du / mu;
    if (rel > c.wmax) c.wmax = rel;
    if (c.step % 100 == 0) {
      if (c.wmax <= 1.0e-12 && c.wmax > 0.5 * c.wprev) done = 1;
      c.wprev = c.wmax;
      c.wmax = 0.0;
    }
  }
  c.done = done;
  *ctl = c;
}

// interpolate the base-level solution onto every point of a finer level
// (:450-490), and set v = 0
__global__ void __launch_bounds__(kBlk) k_ell_inject(KGrid Gf, KGrid G0, const double* __restrict__ u0, double* uf,
                                                     double* vf, int interporder) {
  const long q = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x;
  const int nzt = Gf.Nz + Gf.g;
  if (q >= static_cast<long>(Gf.Nr + Gf.g) * nzt) return;
  const int m = static_cast<int>(q / nzt), n = static_cast<int>(q % nzt);
  const int g = Gf.g;
  const double r0 = coord_r(Gf, 1 - g) + static_cast<double>(m) * Gf.dr;
  const double z0 = coord_z(Gf, 1 - g) + static_cast<double>(n) * Gf.dz;
  const Cell c = locate(G0, r0, z0, interporder);
  const long k = IDX(Gf, m + 1 - g, n + 1 - g);
  uf[k] = apply(c, u0, G0);
  (void)vf;
}

__global__ void __launch_bounds__(kBlk) k_ell_zero(long n, double* __restrict__ f) {
  const long q = static_cast<long>(blockIdx.x) * kBlk + threadIdx.x;
  if (q < n) f[q] = 0.0;
}

// ------------------------------------------------------------------ SOR
// sor (sor_elliptic.f90:535-802) for the flat Laplacian on one level:
// red-black successive over-relaxation with Chebyshev acceleration.  The
// reference's arrays run over 1 .. N+ghost; the level's own indices are used
// here, so its sweep over 2 .. imax-1 is i = 2-g .. Nr-1 (j alike) and its
// colour (i+j) mod 2 is the same.  The flat coefficients are formed from r
// at each point (the mixed-derivative ones are zero).
struct SorCtl {
  int n;                 // iterations done
  int done;              // converged or ELL_maxiter reached: later kernels return at once
  double omega;          // over-relaxation factor of the next pass
  double anorm;          // running residual norm of the iteration (both passes)
  double ganorm;         // its value after the second pass
};

__global__ void k_sor_reset(SorCtl* ctl) {
  if (threadIdx.x == 0) *ctl = SorCtl{0, 0, 1.0, 0.0, 0.0};
}

// One colour of one iteration: sor_e = aux + S5 u^4 with u at the start of
// the iteration (the other colour's pass does not touch this point), the
// residual, the update; per block the sum of resid^2 (the reference sums in
// sweep order, here by blocks in a fixed order).
__global__ void __launch_bounds__(kBlk) k_sor_sweep(KGrid G, EllPlanes E, int color, double* __restrict__ part,
                                                    const SorCtl* __restrict__ ctl) {
  __shared__ double ssq[kBlk];
  const int t = static_cast<int>(threadIdx.x);
  double sq = 0.0;
  if (!ctl->done) {
    const double one = 1.0, half = 0.5, two = 2.0;
    const double omega = ctl->omega;
    const int ni = G.Nr - 1 - (2 - G.g) + 1, nj = G.Nz - 1 - (2 - G.g) + 1;
    const long n = static_cast<long>(ni) * nj;
    double* __restrict__ u = E.u;
    for (long q = static_cast<long>(blockIdx.x) * kBlk + t; q < n; q += static_cast<long>(gridDim.x) * kBlk) {
      const int i = 2 - G.g + static_cast<int>(q / nj), j = 2 - G.g + static_cast<int>(q % nj);
      if (((i + j) & 1) != color) continue;
      const long k = IDX(G, i, j);
      const double r = coord_r(G, i);
      const double a = (one / G.dr + half / r) / G.dr;
      const double b = (one / G.dr - half / r) / G.dr;
      const double c = one / (G.dz * G.dz);
      const double d = one / (G.dz * G.dz);
      const double u0 = u[k];
      const double u2 = u0 * u0;
      const double e = (-two * (one / (G.dr * G.dr) + one / (G.dz * G.dz)) + E.s1[k]) + E.s5[k] * (u2 * u2);
      const double resid = a * u[IDX(G, i + 1, j)] + b * u[IDX(G, i - 1, j)] + c * u[IDX(G, i, j + 1)]
                           + d * u[IDX(G, i, j - 1)] + e * u0 - E.s0[k];
      sq += resid * resid;
      u[k] = u0 - omega * resid / e;
    }
  }
  ssq[t] = sq;
  __syncthreads();
  for (int h = kBlk / 2; h > 0; h >>= 1) {
    if (t < h) ssq[t] += ssq[t + h];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = ssq[0];
}

// after a pass: anorm = sqrt(anorm / (imax jmax)) (the second pass adds its
// squares to the first pass's norm), then the next omega (:585-600)
__global__ void k_sor_norm(int nparts, const double* __restrict__ part, double nn, double rjac, int pass,
                           SorCtl* ctl) {
  if (threadIdx.x != 0 || ctl->done) return;
  double sq = 0.0;
  for (int b = 0; b < nparts; ++b) sq += part[b];
  SorCtl c = *ctl;
  if (pass == 0) {
    c.n += 1;
    c.anorm = sqrt(sq / nn);
  } else {
    c.anorm = sqrt((c.anorm + sq) / nn);
    c.ganorm = c.anorm;
  }
  if (c.n == 1 && pass == 0)
    c.omega = 1.0 / (1.0 - 0.5 * rjac * rjac);
  else
    c.omega = 1.0 / (1.0 - 0.25 * rjac * rjac * c.omega);
  *ctl = c;
}

// Robin boundary of the base level (:655-739), uinf = 1: the outer rho
// boundary, then the upper and (no equatorial symmetry) the lower z
// boundary, then the corners.  Each line is a recurrence in the order of the
// reference (the point before is already new, the point after still old),
// run by one thread over values the block stages in LDS in chunks.
constexpr int kSorChunk = kBlk;
constexpr size_t kSorLds = sizeof(double) * (3 * kSorChunk + 1);

struct SorLine {
  long k0, dk;           // plane offset of the line's first point and the step along it
  long d1;               // offset from a line point to its interior neighbour (d2 = 2 d1)
  int n;                 // points on the line
  double cfix;           // cr (rho boundary) or cz (z boundaries) of the line
  int sgn;               // +1: outer boundary, -1: lower z boundary
  bool along_z;          // the line runs along z (the rho boundary)
  int first;             // index (j or i) of its first point
};

__device__ void sor_line(const KGrid& G, double* __restrict__ u, const SorLine L, double uinf, double* lds) {
  double* cur = lds;                     // old values of points q0 .. q0 + chunk
  double* in1 = lds + kSorChunk + 1;     // the interior neighbours
  double* in2 = in1 + kSorChunk;
  const int t = static_cast<int>(threadIdx.x);
  double prev = 0.0;
  for (int q0 = 0; q0 < L.n; q0 += kSorChunk) {
    const int m = min(kSorChunk, L.n - q0);
    for (int q = t; q <= m; q += kBlk) cur[q] = u[L.k0 + (q0 + q) * L.dk];
    for (int q = t; q < m; q += kBlk) {
      const long k = L.k0 + (q0 + q) * L.dk;
      in1[q] = u[k + L.d1];
      in2[q] = u[k + 2 * L.d1];
    }
    __syncthreads();
    if (t == 0) {
      // the point before the line keeps its old value
      if (q0 == 0) prev = u[L.k0 - L.dk];
      const double c = L.cfix;
      for (int q = 0; q < m; ++q) {
        const int x = L.first + q0 + q;
        const double w = L.along_z ? coord_z(G, x) / G.dz : coord_r(G, x) / G.dr;
        const double after = cur[q + 1];
        double v;
        if (L.sgn > 0)
          v = (uinf + 0.5 * c * (4.0 * in1[q] - in2[q]) - 0.5 * w * (after - prev)) / (1.0 + 1.5 * c);
        else
          v = (uinf - 0.5 * c * (4.0 * in1[q] - in2[q]) - 0.5 * w * (after - prev)) / (1.0 - 1.5 * c);
        cur[q] = v;
        prev = v;
      }
    }
    __syncthreads();
    for (int q = t; q < m; q += kBlk) u[L.k0 + (q0 + q) * L.dk] = cur[q];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlk) k_sor_robin(KGrid G, EllPlanes E, const SorCtl* __restrict__ ctl) {
  if (ctl->done) return;
  extern __shared__ double lds[];
  double* __restrict__ u = E.u;
  const int g = G.g, Nr = G.Nr, Nz = G.Nz;
  const double uinf = 1.0;
  // the rho boundary i = Nr, j = 2-g .. Nz-1
  sor_line(G, u, SorLine{IDX(G, Nr, 2 - g), 1, -static_cast<long>(G.ld), Nz - 1 - (2 - g) + 1,
                         coord_r(G, Nr) / G.dr, +1, true, 2 - g}, uinf, lds);
  // the upper z boundary j = Nz, i = 2-g .. Nr-1
  sor_line(G, u, SorLine{IDX(G, 2 - g, Nz), G.ld, -1, Nr - 1 - (2 - g) + 1, coord_z(G, Nz) / G.dz, +1, false, 2 - g},
           uinf, lds);
  // the lower z boundary j = 1-g, i = 2-g .. Nr-1
  if (!G.eqsym)
    sor_line(G, u, SorLine{IDX(G, 2 - g, 1 - g), G.ld, 1, Nr - 1 - (2 - g) + 1, coord_z(G, 1 - g) / G.dz, -1, false,
                           2 - g}, uinf, lds);
  if (threadIdx.x == 0) {
    const double cr = coord_r(G, Nr) / G.dr;
    double cz = coord_z(G, Nz) / G.dz;
    u[IDX(G, Nr, Nz)] = (uinf + 0.5 * cr * (4.0 * u[IDX(G, Nr - 1, Nz)] - u[IDX(G, Nr - 2, Nz)])
                         + 0.5 * cz * (4.0 * u[IDX(G, Nr, Nz - 1)] - u[IDX(G, Nr, Nz - 2)])) / (1.0 + 1.5 * (cr + cz));
    if (!G.eqsym) {
      cz = coord_z(G, 1 - g) / G.dz;
      u[IDX(G, Nr, 1 - g)] = (uinf + 0.5 * cr * (4.0 * u[IDX(G, Nr - 1, 1 - g)] - u[IDX(G, Nr - 2, 1 - g)])
                              - 0.5 * cz * (4.0 * u[IDX(G, Nr, 2 - g)] - u[IDX(G, Nr, 3 - g)]))
                             / (1.0 + 1.5 * (cr - cz));
    }
  }
}

// the images on the axis, then (eqsym) the equator (:741-758), and the
// stopping rule of the iteration (:766-775)
__global__ void __launch_bounds__(kBlk) k_sor_sym(KGrid G, EllPlanes E, double eps, int maxiter, SorCtl* ctl) {
  if (ctl->done) return;
  double* __restrict__ u = E.u;
  const int g = G.g;
  const int t = static_cast<int>(threadIdx.x);
  const int nzt = G.Nz + g, nrt = G.Nr + g;
  for (int q = t; q < g * nzt; q += kBlk) {
    const int i = 1 + q / nzt, j = 1 - g + q % nzt;
    u[IDX(G, 1 - i, j)] = u[IDX(G, i, j)];
  }
  __syncthreads();
  if (G.eqsym) {
    for (int q = t; q < g * nrt; q += kBlk) {
      const int j = 1 + q / nrt, i = 1 - g + q % nrt;
      u[IDX(G, i, 1 - j)] = u[IDX(G, i, j)];
    }
  }
  __syncthreads();
  if (t == 0) {
    SorCtl c = *ctl;
    if (c.ganorm < eps || c.n >= maxiter) c.done = 1;
    *ctl = c;
  }
}

int blocks(long n) { return static_cast<int>((n + kBlk - 1) / kBlk); }

int launched(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return OA_OK;
  oa_set_error(std::string("elliptic solve: ") + what + ": " + hipGetErrorString(e));
  return OA_ERR_HIP;
}

#define ELLRC(call)                   \
  do {                                \
    const int rc_ = (call);           \
    if (rc_ != OA_OK) return rc_;     \
  } while (0)
#define ELLHIP(call, what)                                                        \
  do {                                                                            \
    const hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                       \
      oa_set_error(std::string("elliptic solve: ") + what + ": " + hipGetErrorString(e_)); \
      return OA_ERR_HIP;                                                          \
    }                                                                             \
  } while (0)

}  // namespace

// Host side of one solve.
struct EllSolver {
  oa_handle* h = nullptr;
  hipStream_t st = nullptr;
  int nlev = 0;
  std::vector<EllLevelArgs> lev;
  double* arena = nullptr;
  EllClock* clocks = nullptr;
  EllCtl* ctl = nullptr;
  double* part = nullptr;
  int nparts = 0;
  EllCtl* host_ctl = nullptr;    // pinned
  SorCtl* sctl = nullptr;        // the SOR solve's control block
  SorCtl* host_sctl = nullptr;   // pinned
  double dt0 = 0.0;
  int stationary = 1;

  ~EllSolver() {
    if (st) (void)hipStreamSynchronize(st);
    if (arena) (void)hipFree(arena);
    if (host_ctl) (void)hipHostFree(host_ctl);
    if (host_sctl) (void)hipHostFree(host_sctl);
  }

  int setup(oa_handle* hh, int stationary_exit) {
    h = hh;
    st = h->stream;
    nlev = h->nlev;
    stationary = stationary_exit;
    const oa_params& p = h->p;
    dt0 = p.WE_dtfac * std::fmin(p.dr0, p.dz0);   // flat Laplacian (:178-183)
    size_t total = 0;
    std::vector<size_t> lev_off(nlev), strip_off(nlev);
    for (int l = 0; l < nlev; ++l) {
      const KGrid& G = h->L[l].G;
      lev_off[l] = total;
      total += static_cast<size_t>(kEllPlanes) * G.plane;
      strip_off[l] = total;
      const long per_f = static_cast<long>(G.g) * (G.Nz + G.g) + 2L * G.g * (G.Nr + G.g);
      total += 3 * 2 * static_cast<size_t>(per_f);
    }
    nparts = 256;
    const size_t ctl_doubles = (sizeof(EllCtl) + sizeof(double) - 1) / sizeof(double);
    const size_t clk_doubles = (sizeof(EllClock) * nlev + sizeof(double) - 1) / sizeof(double);
    const size_t sctl_doubles = (sizeof(SorCtl) + sizeof(double) - 1) / sizeof(double);
    const size_t bytes = sizeof(double) * (total + 3 * nparts + ctl_doubles + clk_doubles + sctl_doubles);
    ELLHIP(hipMalloc(&arena, bytes), "device allocation");
    part = arena + total;
    ctl = reinterpret_cast<EllCtl*>(part + 3 * nparts);
    clocks = reinterpret_cast<EllClock*>(part + 3 * nparts + ctl_doubles);
    sctl = reinterpret_cast<SorCtl*>(part + 3 * nparts + ctl_doubles + clk_doubles);
    ELLHIP(hipHostMalloc(reinterpret_cast<void**>(&host_ctl), sizeof(EllCtl)), "pinned allocation");
    ELLHIP(hipHostMalloc(reinterpret_cast<void**>(&host_sctl), sizeof(SorCtl)), "pinned allocation");
    lev.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
      EllLevelArgs& a = lev[l];
      a.G = h->L[l].G;
      a.G.dt = dt0 / static_cast<double>(1 << l);
      double* b = arena + lev_off[l];
      const long pl = a.G.plane;
      double** planes[kEllPlanes] = {&a.P.u, &a.P.v, &a.P.up, &a.P.vp, &a.P.t1u, &a.P.t1v, &a.P.t2u,
                                     &a.P.t2v, &a.P.ua, &a.P.va, &a.P.s0, &a.P.s1, &a.P.s5, &a.P.sv};
      for (int k = 0; k < kEllPlanes; ++k) *planes[k] = b + k * pl;
      a.P.strip = arena + strip_off[l];
      a.P.strip_tl = 2 * (static_cast<long>(a.G.g) * (a.G.Nz + a.G.g) + 2L * a.G.g * (a.G.Nr + a.G.g));
      a.clk = clocks + l;
    }
    return OA_OK;
  }

  // the coefficients from the state slabs that the first initial-data pass filled
  int coefficients() {
    const bool shift = h->shift;
    for (int l = 0; l < nlev; ++l) {
      const EllLevelArgs& a = lev[l];
      const long n = static_cast<long>(a.G.Nr + a.G.g) * (a.G.Nz + a.G.g);
      hipLaunchKernelGGL(k_ell_coeffs, dim3(blocks(n)), dim3(kBlk), 0, st, h->p, a.G,
                         h->L[l].slab[h->L[l].cur], scalar_slot(shift, SC_phi), scalar_slot(shift, SC_xi_r),
                         scalar_slot(shift, SC_xi_z), a.P);
      ELLRC(launched("coefficients"));
      const long m = std::max(a.G.plane, 3 * a.P.strip_tl);
      hipLaunchKernelGGL(k_ell_init, dim3(blocks(m)), dim3(kBlk), 0, st, a.G, a.P, a.clk);
      ELLRC(launched("initial guess"));
    }
    return OA_OK;
  }

  // wavestep(level) as launches (recursion into the levels up to nlmax)
  int wavestep(int level, int nlmax) {
    const oa_params& p = h->p;
    EllLevelArgs& a = lev[level];
    const KGrid& G = a.G;
    const bool rk4 = (p.WE_method == OA_RK4);
    const int niter = rk4 ? 4 : p.icniter;
    {
      const long per_f = static_cast<long>(G.g) * (G.Nz + G.g) + 2L * G.g * (G.Nr + G.g);
      const long n = std::max(G.plane, level > 0 ? 2 * per_f : 0L);
      hipLaunchKernelGGL(k_ell_saveold, dim3(blocks(n)), dim3(kBlk), 0, st, a, level > 0 ? 1 : 0, ctl, clocks,
                         nlev);
      ELLRC(launched("saveold"));
    }
    const dim3 blk(64, 4, 1);
    const dim3 grd((G.Nz - G.jlo + 1 + 63) / 64, (G.Nr + 3) / 4, 1);
    for (int iter = 1; iter <= niter; ++iter) {
      EllIterArgs it{};
      it.L = a;
      double dtw = 0.0, w = 0.0;
      if (rk4) {
        switch (iter) {
          case 1: dtw = 0.5 * G.dt; w = 1.0 / 6.0; break;
          case 2: dtw = 0.5 * G.dt; w = 1.0 / 3.0; break;
          case 3: dtw = G.dt; w = 1.0 / 3.0; break;
          default: dtw = G.dt; w = 1.0 / 6.0; break;
        }
      } else {
        dtw = (iter < niter) ? 0.5 * G.dt : G.dt;
      }
      // inputs and outputs: up -> t1 -> t2 -> t1 ... -> u.  The first
      // iteration reads up, which saveold has just made equal to u; the
      // last writes u, which no iteration reads, so no launch reads the
      // plane it writes
      double* tu[2] = {a.P.t1u, a.P.t2u};
      double* tv[2] = {a.P.t1v, a.P.t2v};
      it.uin = (iter == 1) ? a.P.up : tu[(iter - 2) % 2];
      it.vin = (iter == 1) ? a.P.vp : tv[(iter - 2) % 2];
      it.uout = (iter == niter) ? a.P.u : tu[(iter - 1) % 2];
      it.vout = (iter == niter) ? a.P.v : tv[(iter - 1) % 2];
      it.dtw = dtw;
      it.w = w;
      it.idt = 1.0 / G.dt;
      it.eta = p.WE_eta;
      it.diss = p.WE_diss;
      it.iter = iter;
      it.niter = niter;
      it.rk4 = rk4 ? 1 : 0;
      it.base = (level == 0) ? 1 : 0;
      it.level = level;
      it.s5 = (p.idata == OA_IDATA_SCALARPULSE && p.scalarpotential != OA_POT_NONE) ? 1 : 0;
      if (level == 0) {
        hipLaunchKernelGGL(k_ell_iter, grd, blk, 0, st, it, ctl);
        ELLRC(launched("iteration"));
      } else {
        const EllLevelArgs& c = lev[level - 1];
        const Level& L = h->L[level];
        EllFbArgs fb{};
        fb.Gf = G;
        fb.Gc = c.G;
        fb.F = a.P;
        fb.uc = c.P.u;
        fb.vc = c.P.v;
        fb.uout = it.uout;
        fb.vout = it.vout;
        fb.tab[0] = L.fbtab[0];
        fb.tab[1] = L.fbtab[1];
        fb.n[0] = L.fbn[0];
        fb.n[1] = L.fbn[1];
        fb.cclk = c.clk;
        fb.fclk = a.clk;
        fb.dtw = dtw;
        EllStepArgs sa{};
        sa.it = it;
        sa.fb = fb;
        sa.nbx = static_cast<int>(grd.x);
        sa.nb_iter = static_cast<int>(grd.x * grd.y);
        const int nb = sa.nb_iter + blocks(std::max(L.fbn[0], L.fbn[1]));
        hipLaunchKernelGGL(k_ell_iter_fb, dim3(nb), dim3(kBlk), 0, st, sa, ctl);
        ELLRC(launched("iteration and finebound"));
      }
    }
    if (level < nlmax) {
      ELLRC(wavestep(level + 1, nlmax));
      ELLRC(wavestep(level + 1, nlmax));
      // the second step of level+1 brings its clock to an even step count:
      // restrict onto this level (:1106-1157)
      const EllLevelArgs& f = lev[level + 1];
      EllRsArgs rs{};
      rs.Gf = f.G;
      rs.Gc = G;
      rs.uf = f.P.u;
      rs.vf = f.P.v;
      rs.uc = a.P.u;
      rs.vc = a.P.v;
      rs.interporder = p.interporder;
      rs.images = 1;
      restrict_box(f.G, G, &rs.ic0, &rs.nic, &rs.jc0, &rs.njc);
      hipLaunchKernelGGL(k_ell_restrict, dim3(blocks(static_cast<long>(rs.nic) * rs.njc)), dim3(kBlk), 0, st, rs,
                         ctl);
      ELLRC(launched("restrict"));
    }
    return OA_OK;
  }

  int coarse_step(int nlmax) {
    ELLRC(wavestep(0, nlmax));
    const EllLevelArgs& b = lev[0];
    hipLaunchKernelGGL(k_ell_check, dim3(nparts), dim3(kBlk), 0, st, b.G, b.P, part, ctl);
    ELLRC(launched("residual"));
    const double nntot = static_cast<double>(b.G.Nr) * static_cast<double>(b.G.Nz);
    hipLaunchKernelGGL(k_ell_decide, dim3(1), dim3(64), 0, st, nparts, part, nntot, h->p.WE_epsilon,
                       h->p.WE_maxiter, ctl);
    return launched("stopping rule");
  }

  // one phase: coarse steps over levels 0..nlmax until the rule fires.  The
  // coarse step is captured once as a graph and replayed, unless the handle's
  // caller turned graphs off (oa_set_graphs): then every step is launched
  // kernel by kernel, with the same kernels in the same order.
  int phase(int nlmax, int stationary_rule, int* steps, double* gres) {
    hipLaunchKernelGGL(k_ell_reset, dim3(1), dim3(64), 0, st, ctl, stationary_rule, clocks, nlev);
    ELLRC(launched("reset"));
    const bool use_graph = h->plan.graphs || h->plan.tuner.enabled;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
      ELLHIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
      const int rc = coarse_step(nlmax);
      hipGraph_t captured = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &captured);
      if (rc != OA_OK) {
        if (captured) (void)hipGraphDestroy(captured);
        return rc;
      }
      ELLHIP(ec, "graph capture");
      graph = captured;
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        ELLHIP(ei, "graph instantiation");
      }
    }
    int status = OA_OK;
    const int maxiter = h->p.WE_maxiter;
    long launched_steps = 0;
    for (;;) {
      // at least 100 steps are always taken; then chunks of 64
      const int chunk = (launched_steps == 0) ? 100 : 64;
      for (int k = 0; k < chunk && status == OA_OK; ++k) {
        if (!use_graph) {
          status = coarse_step(nlmax);
          continue;
        }
        const hipError_t e = hipGraphLaunch(exec, st);
        if (e != hipSuccess) {
          oa_set_error(std::string("elliptic solve: graph launch: ") + hipGetErrorString(e));
          status = OA_ERR_HIP;
        }
      }
      launched_steps += chunk;
      if (status != OA_OK) break;
      const hipError_t e1 = hipMemcpyAsync(host_ctl, ctl, sizeof(EllCtl), hipMemcpyDeviceToHost, st);
      const hipError_t e2 = hipStreamSynchronize(st);
      if (e1 != hipSuccess || e2 != hipSuccess) {
        oa_set_error("elliptic solve: reading the stopping rule failed");
        status = OA_ERR_HIP;
        break;
      }
      if (host_ctl->done) break;
      if (launched_steps > static_cast<long>(maxiter) + 200) {
        oa_set_error("elliptic solve: the stopping rule did not fire within WE_maxiter steps");
        status = OA_ERR_STATE;
        break;
      }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (status != OA_OK) return status;
    *steps = host_ctl->step;
    *gres = host_ctl->gres;
    return OA_OK;
  }

  // ------------------------------------------------------------ SOR
  // one iteration of sor() on a level (sor_elliptic.f90:575-775) as launches
  int sor_iteration(int level, bool robin, double rjac) {
    const EllLevelArgs& a = lev[level];
    const KGrid& G = a.G;
    const double nn = static_cast<double>(G.Nr + G.g) * static_cast<double>(G.Nz + G.g);
    for (int color = 0; color < 2; ++color) {
      hipLaunchKernelGGL(k_sor_sweep, dim3(nparts), dim3(kBlk), 0, st, G, a.P, color, part, sctl);
      ELLRC(launched("SOR sweep"));
      hipLaunchKernelGGL(k_sor_norm, dim3(1), dim3(64), 0, st, nparts, part, nn, rjac, color, sctl);
      ELLRC(launched("SOR norm"));
    }
    if (robin) {
      hipLaunchKernelGGL(k_sor_robin, dim3(1), dim3(kBlk), kSorLds, st, G, a.P, sctl);
      ELLRC(launched("SOR Robin boundary"));
    }
    hipLaunchKernelGGL(k_sor_sym, dim3(1), dim3(kBlk), 0, st, G, a.P, h->p.ELL_epsilon, h->p.ELL_maxiter, sctl);
    return launched("SOR symmetries");
  }

  // sor() on one level until its residual falls below ELL_epsilon or
  // ELL_maxiter iterations; the iteration is captured once as a graph and
  // replayed in chunks, unless graphs are switched off
  int sor_level(int level, bool robin, double rjac, int* iters, double* res) {
    hipLaunchKernelGGL(k_sor_reset, dim3(1), dim3(64), 0, st, sctl);
    ELLRC(launched("SOR reset"));
    const bool use_graph = h->plan.graphs || h->plan.tuner.enabled;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
      ELLHIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
      const int rc = sor_iteration(level, robin, rjac);
      hipGraph_t captured = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &captured);
      if (rc != OA_OK) {
        if (captured) (void)hipGraphDestroy(captured);
        return rc;
      }
      ELLHIP(ec, "graph capture");
      graph = captured;
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        ELLHIP(ei, "graph instantiation");
      }
    }
    int status = OA_OK;
    long launched_its = 0;
    for (;;) {
      const int chunk = (launched_its == 0) ? 100 : 64;
      for (int k = 0; k < chunk && status == OA_OK; ++k) {
        if (!use_graph) {
          status = sor_iteration(level, robin, rjac);
          continue;
        }
        const hipError_t e = hipGraphLaunch(exec, st);
        if (e != hipSuccess) {
          oa_set_error(std::string("elliptic solve: graph launch: ") + hipGetErrorString(e));
          status = OA_ERR_HIP;
        }
      }
      launched_its += chunk;
      if (status != OA_OK) break;
      const hipError_t e1 = hipMemcpyAsync(host_sctl, sctl, sizeof(SorCtl), hipMemcpyDeviceToHost, st);
      const hipError_t e2 = hipStreamSynchronize(st);
      if (e1 != hipSuccess || e2 != hipSuccess) {
        oa_set_error("elliptic solve: reading the SOR state failed");
        status = OA_ERR_HIP;
        break;
      }
      if (host_sctl->done) break;
      if (launched_its > static_cast<long>(h->p.ELL_maxiter) + 200) {
        oa_set_error("elliptic solve: SOR did not stop within ELL_maxiter iterations");
        status = OA_ERR_STATE;
        break;
      }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (status != OA_OK) return status;
    *iters = host_sctl->n;
    *res = host_sctl->ganorm;
    return OA_OK;
  }

  // sor_elliptic('flat', 'one') (sor_elliptic.f90:2-530): u = 1 on every
  // level (the coefficients pass set it), the base level with Robin
  // boundaries, each finer level from the interpolation of its parent's
  // solution (initial guess and Dirichlet boundary), then every level
  // restricted onto its parent, level 1 first, without images
  int sor_solve(int* steps0, int* steps1, double* gres) {
    const KGrid& G0 = lev[0].G;
    const double smallpi = acos(-1.0);
    const double rjac = (G0.dz / G0.dr * cos(smallpi / static_cast<double>(G0.Nr + G0.g))
                         + G0.dr / G0.dz * cos(smallpi / static_cast<double>(G0.Nz + G0.g)))
                        / (G0.dz / G0.dr + G0.dr / G0.dz);
    ELLRC(sor_level(0, true, rjac, steps0, gres));
    *steps1 = 0;
    for (int l = 1; l < nlev; ++l) {
      const EllLevelArgs& f = lev[l];
      const EllLevelArgs& c = lev[l - 1];
      const long n = static_cast<long>(f.G.Nr + f.G.g) * (f.G.Nz + f.G.g);
      hipLaunchKernelGGL(k_ell_inject, dim3(blocks(n)), dim3(kBlk), 0, st, f.G, c.G, c.P.u, f.P.u, f.P.v,
                         h->p.interporder);
      ELLRC(launched("injection"));
      int it = 0;
      double r = 0.0;
      ELLRC(sor_level(l, false, rjac, &it, &r));
      *steps1 += it;
    }
    // k_ell_restrict runs while the wave solve's control block says not done
    hipLaunchKernelGGL(k_ell_reset, dim3(1), dim3(64), 0, st, ctl, 0, clocks, nlev);
    ELLRC(launched("reset"));
    for (int l = 1; l < nlev; ++l) {
      const EllLevelArgs& f = lev[l];
      const EllLevelArgs& c = lev[l - 1];
      EllRsArgs rs{};
      rs.Gf = f.G;
      rs.Gc = c.G;
      rs.uf = f.P.u;
      rs.vf = f.P.v;
      rs.uc = c.P.u;
      rs.vc = c.P.v;
      rs.interporder = h->p.interporder;
      rs.images = 0;
      restrict_box(f.G, c.G, &rs.ic0, &rs.nic, &rs.jc0, &rs.njc);
      hipLaunchKernelGGL(k_ell_restrict, dim3(blocks(static_cast<long>(rs.nic) * rs.njc)), dim3(kBlk), 0, st, rs,
                         ctl);
      ELLRC(launched("restrict"));
    }
    return OA_OK;
  }

  // wave_elliptic: the base level alone, then (refinement) all levels
  int solve(int* steps0, int* steps1, double* gres) {
    if (h->p.ELL_solver == OA_ELL_SOR) return sor_solve(steps0, steps1, gres);
    const int nlmax_old = nlev - 1;
    *steps1 = 0;
    ELLRC(phase(nlmax_old > 0 ? 0 : nlmax_old, 0, steps0, gres));
    if (nlmax_old > 0 && *steps0 != h->p.ELL_maxiter) {
      const EllLevelArgs& b = lev[0];
      hipLaunchKernelGGL(k_ell_zero, dim3(blocks(b.G.plane)), dim3(kBlk), 0, st, b.G.plane, b.P.v);
      ELLRC(launched("v = 0"));
      for (int l = 1; l <= nlmax_old; ++l) {
        const EllLevelArgs& f = lev[l];
        const long n = static_cast<long>(f.G.Nr + f.G.g) * (f.G.Nz + f.G.g);
        hipLaunchKernelGGL(k_ell_inject, dim3(blocks(n)), dim3(kBlk), 0, st, f.G, b.G, b.P.u, f.P.u, f.P.v,
                           h->p.interporder);
        ELLRC(launched("injection"));
      }
      ELLRC(phase(nlmax_old, stationary, steps1, gres));
    }
    return OA_OK;
  }
};

// ------------------------------------------------------------ interface
int ell_solve_begin(oa_handle* h, int stationary_exit, EllSolver** out) {
  EllSolver* s = new EllSolver();
  int rc = s->setup(h, stationary_exit);
  if (rc == OA_OK) rc = s->coefficients();
  if (rc != OA_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return OA_OK;
}

int ell_solve_run(EllSolver* s, int* steps0, int* steps1, double* gres) { return s->solve(steps0, steps1, gres); }

const double* ell_solution(const EllSolver* s, int level) { return s->lev[level].P.u; }

void ell_solve_end(EllSolver* s) { delete s; }

}  // namespace oax
