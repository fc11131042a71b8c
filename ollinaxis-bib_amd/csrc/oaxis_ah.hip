// Apparent-horizon finder on the device: the shooting and the spectral
// method.
//
// Reference: src/geometry/horizon_finder.f90 -- horizon_finder (:2-237),
// ahshoot (:246-1103), ahspectral (:1113-1972), ah_source (:1983-2752) --
// with no angular momentum (no C1/C2 arrays in this hot path), for box 0.
// The spectral method is described beside k_ah_spectral; the rest of this
// note is about shooting.
//
// The reference integrates the horizon ODE h'' = sdh(h, h', theta) once per
// "shot" from an axis point zinit, marching zinit down the axis by half a
// local cell until the end slope dh/dtheta changes sign, then bisecting.
// Every shot is a serial adaptive RK4 over ~10^2-10^3 steps whose sources
// interpolate 20 metric quantities, so one shot is a long dependent chain:
// this file runs one shot per wave and many shots per launch.
//
//  1. k_ah_fields: one pass per level forms the 20 quantities the ODE
//     interpolates (psi, the conformal-factor and metric first derivatives,
//     A, B, C, H, trK, the physical extrinsic curvature), with their parity
//     images in the ghost zones as the auxiliary arrays hold them, stored
//     point by point (20 consecutive doubles per grid point): one bicubic
//     stencil then reads 16 runs of 160 bytes instead of 320 scattered
//     values.
//  2. k_ah_shoot: one wave integrates one shot (zinit) to the end; its lanes
//     share the 320 loads of every interpolation, the rest is uniform.
//  3. The host walks the reference's control flow (march, bracket, bisection,
//     restart; ahshoot :434-855) over a table of shot results.  When the walk
//     needs a shot it has not got, it launches the whole speculative
//     continuation at once: during the march, the next positions down the
//     axis; during the bisection, the full tree of the next bisection steps,
//     as deep as keeps every search's tree resident on the device at once
//     (kAhResident waves in all, at most kAhMaxDepth steps).  The walk then takes exactly the reference's path
//     through those results, so the sequence of shots and every decision
//     are the reference's; unused shots are discarded.
//
// Floating-point contraction is off in this file, so the sources and the
// interpolation round like the oracle's IEEE build of the same expressions.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "oa_internal.h"
#include "oa_state.h"
#define OA_HD __device__ __forceinline__
#define OA_HHD __host__ __device__ inline
#include "oaxis_device.h"

namespace oax {
namespace {

constexpr int kAhMaxLevels = 16;
constexpr int kAhNthetaMax = 50000;                 // Nthetamax (:350)
// Accepted plus rejected steps of one shot.  The reference bounds only the
// accepted ones; rejections halve the step until it is accepted, so this
// bound is reached only by a shot the reference would integrate for hours.
constexpr int kAhMaxSteps = 4 * kAhNthetaMax;
constexpr int kAhScanChunk = 1024;                  // march positions per launch
constexpr int kAhMaxDepth = 11;                     // bisection steps per launch, at most
constexpr int kAhResident = 2048;                   // shot waves resident at once (256 CUs x 8)
constexpr int kShotBlock = 64;

// interpolated quantities, in ah_source's order
enum AhField {
  AF_psi = 0, AF_Drphi, AF_Dzphi,
  AF_A, AF_DrA, AF_DzA, AF_B, AF_DrB, AF_DzB, AF_C, AF_DrC, AF_DzC, AF_H, AF_DrH, AF_DzH,
  AF_trK, AF_KA, AF_KB, AF_KC, AF_KH, AF_N
};
static_assert(AF_N % 2 == 0, "a point record is read as 16-byte pairs");

struct AhLevel {
  KGrid G;
  double rmin, rmax, zmin, zmax;      // rminl .. zmaxl (evolve.f90:80-95)
  const double* q;                    // AF_N values per point, point index IDX(G, i, j)
};

struct AhCtx {
  int nlev, eqsym, interporder, method;
  double ahrmax;                      // clamped to 0.9 as ahshoot does (:367-373)
  AhLevel L[kAhMaxLevels];
};

struct ShotIn {
  const AhCtx* ctx;
  double zinit, zcenter;
  double* theta;                      // when not null: accepted points (kAhNthetaMax + 1)
  double* rad;
};

struct ShotOut {
  double dradlast, area;
  int err, ntheta;
};

// ------------------------------------------------------------ derived planes
// psi = exp(phi) (auxiliary_geometry.f90:236-242), Dr/Dz of phi (:176-231),
// of A, B, C, H (diff1r/diff1z with the metric's parities), and
// KA..KH = psi4 (KT* + X trK / 3) (:1038-1041) with KTH from tracelessness
// (:747-756).  PHIMODE: the last auxiliary ran at level clock 0, so the
// conformal-factor derivatives are those of phi; otherwise of chi.
template <bool PHIMODE>
__global__ void __launch_bounds__(256) k_ah_fields(KGrid G, const double* __restrict__ y, int phi_from_chi,
                                                   double* __restrict__ out) {
  const int j = G.jlo + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = 1 + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
  if (j > G.Nz || i > G.Nr) return;
  const GlobalView V{y, G, nullptr};
  const double r = (static_cast<double>(i) - 0.5) * G.dr;
  const double chi = V.fld(S_chi)(i, j);
  const double phi = phi_from_chi ? -std::log(chi) / 2.0 : V.fld(S_phi)(i, j);
  const double psi = std::exp(phi);
  double Drp, Dzp;
  if constexpr (PHIMODE) {
    const D1 dp = grad(V.fld(S_phi), G, i, j);
    Drp = dp.r;
    Dzp = dp.z;
  } else {
    const D1 dc = grad(V.fld(S_chi), G, i, j);
    Drp = -(dc.r / chi) / 2.0;
    Dzp = -(dc.z / chi) / 2.0;
  }
  const D1 dA = grad(V.fld(S_A), G, i, j), dB = grad(V.fld(S_B), G, i, j);
  const D1 dC = grad(V.fld(S_C), G, i, j), dH = grad(V.fld(S_H), G, i, j);
  const double A = V.fld(S_A)(i, j), B = V.fld(S_B)(i, j), C = V.fld(S_C)(i, j), H = V.fld(S_H)(i, j);
  const double trK = V.fld(S_trK)(i, j), KTA = V.fld(S_KTA)(i, j), KTB = V.fld(S_KTB)(i, j);
  const double KTC = V.fld(S_KTC)(i, j);
  const double gA = B * H, gB = A * H, gH = 1.0 / H, gC = -C * H;
  const double KTH = -(gA * KTA + gB * KTB + 2.0 * r * r * gC * KTC) / gH;
  const double psi2 = psi * psi, psi4 = psi2 * psi2;
  const double third = 1.0 / 3.0;
  double v[AF_N];
  v[AF_psi] = psi;
  v[AF_Drphi] = Drp;
  v[AF_Dzphi] = Dzp;
  v[AF_A] = A;
  v[AF_DrA] = dA.r;
  v[AF_DzA] = dA.z;
  v[AF_B] = B;
  v[AF_DrB] = dB.r;
  v[AF_DzB] = dB.z;
  v[AF_C] = C;
  v[AF_DrC] = dC.r;
  v[AF_DzC] = dC.z;
  v[AF_H] = H;
  v[AF_DrH] = dH.r;
  v[AF_DzH] = dH.z;
  v[AF_trK] = trK;
  v[AF_KA] = psi4 * (KTA + third * A * trK);
  v[AF_KB] = psi4 * (KTB + third * B * trK);
  v[AF_KC] = psi4 * (KTC + third * C * trK);
  v[AF_KH] = psi4 * (KTH + third * H * trK);
  // parities (arrays.config; a derivative flips its direction's parity)
  constexpr double SR[AF_N] = {1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, 1, 1, 1, 1};
  constexpr double SZ[AF_N] = {1, 1, -1, 1, 1, -1, 1, 1, -1, -1, -1, 1, 1, 1, -1, 1, 1, 1, -1, 1};
  auto write = [&](int ii, int jj, bool fr, bool fz) {
    double* o = out + IDX(G, ii, jj) * AF_N;
#pragma unroll
    for (int f = 0; f < AF_N; ++f) o[f] = v[f] * (fr ? SR[f] : 1.0) * (fz ? SZ[f] : 1.0);
  };
  write(i, j, false, false);
  if (i <= G.g) write(1 - i, j, true, false);
  if (G.eqsym && j >= 1 && j <= G.g) {
    write(i, 1 - j, false, true);
    if (i <= G.g) write(1 - i, 1 - j, true, true);
  }
}

// ------------------------------------------------------------ interpolation
// interp (interp.f90:2-270) of every AF_* quantity at one point, by the
// whole wave of the shot: every lane forms the same Lagrange weights (in the
// reference's order), lane f < AF_N sums quantity f over the stencil in the
// reference's order -- so the 20 lanes read each point record as one
// contiguous 160-byte access -- and the results reach every lane through
// LDS (sv).  The caller (ah_source) has already placed (r0, z0) at least one
// cell inside the level (one cell inside level 0's edges); the stencil rows
// and columns are clamped to the array, which matters only where a weight
// is exactly zero.
//
// The 4 x 4 stencil records of the last cell each lane read stay in its
// registers (AhCell): the adaptive RK4 of a shot takes many small steps, so
// consecutive evaluations mostly fall in the same cell of the same level and
// reuse them instead of waiting for another round of global loads.  The
// values, and so the results, are the same either way.  So do the sine and
// cosine of the last angle (stages 2 and 3 of an RK4 step, and the first
// evaluation after an accepted step, repeat the angle before them).
struct AhCell {
  const double* q;       // level records the cell belongs to (null: empty)
  int i0, j0;            // its lower-left stencil point + (1, 1)
  double fa[4][4];       // quantity f of the records (i0-1+a, j0-1+b), indices clamped
  double th, sth, cth;   // the last angle (NaN: none) and its sine and cosine
};

// How interp.f90 combines the 4 x 4 stencil records around (r0, z0) on one
// level (:146-263): the cell, and a copy at a node, the bilinear formula
// near the level's edges (or with interporder = bilinear), else bicubic
// Lagrange weights.  Formed once per point; ah_combine applies it to the
// records of one quantity with the reference's expressions.
struct AhPlan {
  int lev, i0, j0, mode;              // mode 0: node copy, 1: bilinear, 2: bicubic
  double auxr, auxz;
  double w[4][4];
};

__device__ __forceinline__ void ah_plan(const AhLevel& L, int order, double r0, double z0, AhPlan& P) {
  const KGrid& G = L.G;
  const int g = G.g;
  const double r1g = (static_cast<double>(1 - g) - 0.5) * G.dr;
  const double z1g = (static_cast<double>(1 - g) - G.zoff) * G.dz;
  const int i0 = static_cast<int>((r0 - r1g) / G.dr) + 1 - g;
  const double deltar = r0 - (static_cast<double>(i0) - 0.5) * G.dr;
  const int j0 = static_cast<int>((z0 - z1g) / G.dz) + 1 - g;
  const double deltaz = z0 - (static_cast<double>(j0) - G.zoff) * G.dz;
  P.i0 = i0;
  P.j0 = j0;
  P.auxr = 0.0;
  P.auxz = 0.0;
  if (deltar == 0.0 && deltaz == 0.0) {
    P.mode = 0;
    return;
  }
  const bool lin = ((L.rmax - r0) < G.dr) || ((L.zmax - z0) < G.dz) || ((r0 - L.rmin) < G.dr) ||
                   ((z0 - L.zmin) < G.dz);
  const double auxr = deltar / G.dr, auxz = deltaz / G.dz;
  P.auxr = auxr;
  P.auxz = auxz;
  if (order == OA_INTERP_BILINEAR || lin) {
    P.mode = 1;
    return;
  }
  P.mode = 2;
#pragma unroll
  for (int i = -1; i <= 2; ++i) {
    double wr = 1.0;
#pragma unroll
    for (int m = -1; m <= 2; ++m) {
      if (m == i) continue;
      wr = wr * (auxr - static_cast<double>(m)) / static_cast<double>(i - m);
    }
#pragma unroll
    for (int j = -1; j <= 2; ++j) {
      double aux = wr;
#pragma unroll
      for (int n = -1; n <= 2; ++n) {
        if (n == j) continue;
        aux = aux * (auxz - static_cast<double>(n)) / static_cast<double>(j - n);
      }
      P.w[i + 1][j + 1] = aux;
    }
  }
}

__device__ __forceinline__ double ah_combine(const AhPlan& P, const double (&fa)[4][4]) {
  if (P.mode == 0) return fa[1][1];
  const double auxr = P.auxr, auxz = P.auxz;
  if (P.mode == 1)
    return fa[1][1] * (1.0 - auxr) * (1.0 - auxz) + fa[2][1] * auxr * (1.0 - auxz) + fa[1][2] * (1.0 - auxr) * auxz +
           fa[2][2] * auxr * auxz;
  double out = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) out = out + P.w[i][j] * fa[i][j];
  return out;
}

// quantity f of the 16 records of the plan's cell, indices clamped to the
// level's arrays
__device__ __forceinline__ void ah_records(const AhLevel& L, int i0, int j0, int f, double (&fa)[4][4]) {
  const KGrid& G = L.G;
  const int g = G.g;
  auto ci = [&](int i) { return i < 1 - g ? 1 - g : (i > G.Nr ? G.Nr : i); };
  auto cj = [&](int j) { return j < 1 - g ? 1 - g : (j > G.Nz ? G.Nz : j); };
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) fa[i][j] = L.q[IDX(G, ci(i0 - 1 + i), cj(j0 - 1 + j)) * AF_N + f];
}

// The 20 quantities at (r0, z0) on level L by one half-wave (lane f sums
// quantity f); every lane of the half-wave receives them in v, through its
// slot in sv and two block barriers.
__device__ void ah_interp(const AhLevel& L, int order, double r0, double z0, double (&v)[AF_N], double* sv,
                          AhCell& cell) {
  AhPlan P;
  ah_plan(L, order, r0, z0, P);
  const int f = static_cast<int>(threadIdx.x & 31u);
  double* slot = sv + (threadIdx.x >> 5) * AF_N;
  if (cell.q != L.q || cell.i0 != P.i0 || cell.j0 != P.j0) {
    if (f < AF_N) ah_records(L, P.i0, P.j0, f, cell.fa);
    cell.q = L.q;
    cell.i0 = P.i0;
    cell.j0 = P.j0;
  }
  if (f < AF_N) slot[f] = ah_combine(P, cell.fa);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < AF_N; ++q) v[q] = slot[q];
  __syncthreads();
}

// ------------------------------------------------------------ ah_source
// h'' (sdh) and the area element (sarea) at (h, h', theta) for a horizon
// centred at zcenter (:1983-2752), in three parts: ah_locate (the checks
// that end the reference's routine early, and the level to interpolate
// on), ah_interp (the 20 quantities, one half-wave) and ah_source_math (the
// rest, on one lane's values).
struct AhPos {
  double rc, zc, r0, z0;
  int lev, err;
};

// sth, cth: the sine and cosine of the angle
__device__ AhPos ah_locate_sc(const AhCtx& c, double rad, double sth, double cth, double zcenter) {
  const AhLevel& L0 = c.L[0];
  // the checks that end the reference's ah_source early set err; the
  // interpolation point then moves to the middle of level 0
  int err = (rad < 0.0) ? 1 : 0;                                                 // :2140-2145
  const double rc = rad * sth;                                                   // :2151-2159
  const double zc = rad * cth;
  double r0 = rc;
  double z0 = zc + zcenter;
  if (c.eqsym && z0 < 0.0) z0 = -z0;
  // a non-finite position (never produced by a finite integration) is refused
  // before it could index the grid
  if (!(std::isfinite(r0) && std::isfinite(z0))) err = 1;
  if (r0 > c.ahrmax * L0.rmax || std::fabs(z0) > c.ahrmax * L0.zmax) err = 1;   // :2163-2168
  // :2172-2178 (the z test against zmax subtracts drl, as the reference does)
  if (r0 < L0.rmin + L0.G.dr || r0 > L0.rmax - L0.G.dr || z0 < L0.zmin + L0.G.dz || z0 > L0.zmax - L0.G.dr) err = 1;
  if (z0 * zcenter < 0.0 && c.method == OA_AH_SHOOT) err = 1;                   // :2185-2190
  if (err) {
    r0 = 0.5 * (L0.rmin + L0.rmax);
    z0 = 0.5 * (L0.zmin + L0.zmax);
  }

  int lev = 0;                                                                   // :2199-2213
  for (int ll = 1; ll < c.nlev; ++ll) {
    const AhLevel& L = c.L[ll];
    if (r0 > L.rmin + L.G.dr && r0 < L.rmax - L.G.dr && z0 > L.zmin + L.G.dz && z0 < L.zmax - L.G.dz) lev = ll;
  }
  AhPos P;
  P.rc = rc;
  P.zc = zc;
  P.r0 = r0;
  P.z0 = z0;
  P.lev = lev;
  P.err = err;
  return P;
}

__device__ AhPos ah_locate(const AhCtx& c, double rad, double th, double zcenter) {
  return ah_locate_sc(c, rad, std::sin(th), std::cos(th), zcenter);
}

__device__ int ah_source_math(double rad, double dh, double th, double rc, double zc, const double (&v)[AF_N],
                              double& sdh, double& sarea) {
  const double zero = 0.0, half = 0.5, one = 1.0, two = 2.0;
  const double smallpi = M_PI;

  const double interp_psi4 = (v[AF_psi] * v[AF_psi]) * (v[AF_psi] * v[AF_psi]);
  double Dphi[3];                                                                // :2247-2249
  Dphi[0] = (rc * v[AF_Drphi] + zc * v[AF_Dzphi]) / rad;
  Dphi[1] = (zc * v[AF_Drphi] - rc * v[AF_Dzphi]);
  Dphi[2] = zero;

  const double iA = v[AF_A], iDrA = v[AF_DrA], iDzA = v[AF_DzA];
  const double iB = v[AF_B], iDrB = v[AF_DrB], iDzB = v[AF_DzB];
  const double iC = v[AF_C], iDrC = v[AF_DrC], iDzC = v[AF_DzC];
  const double iH = v[AF_H], iDrH = v[AF_DrH], iDzH = v[AF_DzH];
  const double rc2 = rc * rc, zc2 = zc * zc, rad2 = rad * rad;

  double gd[3][3], gu[3][3];                                                     // :2320-2340
  gd[0][0] = (rc2 * iA + zc2 * iB + two * rc2 * zc * iC) / rad2;
  gd[1][1] = (zc2 * iA + rc2 * iB - two * rc2 * zc * iC);
  gd[0][1] = rc * (zc * (iA - iB) + (zc2 - rc2) * iC) / rad;
  gd[2][2] = rc2 * iH;
  gd[0][2] = zero;
  gd[1][2] = zero;
  gd[1][0] = gd[0][1];
  gd[2][0] = gd[0][2];
  gd[2][1] = gd[1][2];
  if (th != zero && th != smallpi) {                                             // :2345-2378
    const double aux = gd[0][0] * gd[1][1] - gd[0][1] * gd[0][1];
    gu[0][0] = gd[1][1] / aux;
    gu[1][1] = gd[0][0] / aux;
    gu[2][2] = one / gd[2][2];
    gu[0][1] = -gd[0][1] / aux;
  } else {
    gu[0][0] = one / gd[0][0];
    gu[1][1] = one / gd[1][1];
    gu[2][2] = zero;
    gu[0][1] = zero;
  }
  gu[0][2] = zero;
  gu[1][2] = zero;
  gu[1][0] = gu[0][1];
  gu[2][0] = gu[0][2];
  gu[2][1] = gu[1][2];

  double Dg[3][3][3];                                                            // :2388-2472
  double diffr, diffz;
  diffr = (two * rc * iA + rc2 * iDrA + zc2 * iDrB + two * rc * zc * (two * iC + rc * iDrC) - two * rc * gd[0][0]) / rad2;
  diffz = (rc2 * iDzA + two * zc * iB + zc2 * iDzB + two * rc2 * (iC + zc * iDzC) - two * zc * gd[0][0]) / rad2;
  Dg[0][0][0] = (rc * diffr + zc * diffz) / rad;
  Dg[1][0][0] = (zc * diffr - rc * diffz);
  diffr = (two * rc * iB + rc2 * iDrB + zc2 * iDrA - two * rc * zc * (two * iC + rc * iDrC));
  diffz = (rc2 * iDzB + two * zc * iA + zc2 * iDzA - two * rc2 * (iC + zc * iDzC));
  Dg[0][1][1] = (rc * diffr + zc * diffz) / rad;
  Dg[1][1][1] = (zc * diffr - rc * diffz);
  {
    const double q = rc / rad;
    diffr = ((one - q * q) * (zc * (iA - iB) + (zc2 - rc2) * iC) +
             rc * (zc * (iDrA - iDrB) + (zc2 - rc2) * iDrC - two * rc * iC)) / rad;
    diffz = q * (-(zc / rad2) * (zc * (iA - iB) + (zc2 - rc2) * iC) + (iA - iB) + zc * (iDzA - iDzB) +
                 (zc2 - rc2) * iDzC + two * zc * iC);
  }
  Dg[0][0][1] = (rc * diffr + zc * diffz) / rad;
  Dg[1][0][1] = (zc * diffr - rc * diffz);
  diffr = rc * (two * iH + rc * iDrH);
  diffz = rc2 * iDzH;
  Dg[0][2][2] = (rc * diffr + zc * diffz) / rad;
  Dg[1][2][2] = (zc * diffr - rc * diffz);
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    Dg[d][0][2] = zero;
    Dg[d][1][2] = zero;
    Dg[d][1][0] = Dg[d][0][1];
    Dg[d][2][0] = Dg[d][0][2];
    Dg[d][2][1] = Dg[d][1][2];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Dg[2][i][j] = zero;

  double Kd[3][3];                                                               // :2510-2530
  Kd[0][0] = (rc2 * v[AF_KA] + zc2 * v[AF_KB] + two * rc2 * zc * v[AF_KC]) / rad2;
  Kd[1][1] = (zc2 * v[AF_KA] + rc2 * v[AF_KB] - two * rc2 * zc * v[AF_KC]);
  Kd[0][1] = rc * (zc * (v[AF_KA] - v[AF_KB]) + (zc2 - rc2) * v[AF_KC]) / rad;
  Kd[2][2] = rc2 * v[AF_KH];
  Kd[0][2] = zero;
  Kd[1][2] = zero;
  Kd[1][0] = Kd[0][1];
  Kd[2][0] = Kd[0][2];
  Kd[2][1] = Kd[1][2];

  double DF[3];                                                                  // :2537-2552
#pragma unroll
  for (int i = 0; i < 3; ++i) DF[i] = (gu[0][i] - gu[1][i] * dh) / interp_psi4;
  const double u2 = (gu[0][0] - two * gu[0][1] * dh + gu[1][1] * dh * dh) / interp_psi4;
  const double u = std::sqrt(std::fabs(u2));
  const double denom = (gu[0][0] * gu[1][1] - gu[0][1] * gu[0][1]) / (interp_psi4 * interp_psi4);

  // Christoffel symbols (:2575-2608) and the contractions (:2615-2650).  In
  // axisymmetry without rotation g^{rho phi} = g^{z phi} = 0 and D_phi F = 0,
  // so in the sums over m the m = phi terms, and in the contractions every
  // term with one phi index, are products with an exact zero: they are left
  // out (IEEE arithmetic keeps the compiler from dropping them itself), which
  // changes no value, and only the symbols the contractions use are formed:
  // Gam^k_{ij} for (i, j) in the rho-z block and (phi, phi).
  constexpr int kIJ[5][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}, {2, 2}};
  double Gam[2][3][3];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const int i = kIJ[q][0], j = kIJ[q][1];
      double gm = zero;
#pragma unroll
      for (int m = 0; m < 2; ++m) gm = gm + half * gu[k][m] * (Dg[i][j][m] + Dg[j][i][m] - Dg[m][i][j]);
      const double dki = (k == i) ? one : zero, dkj = (k == j) ? one : zero;
      Gam[k][i][j] = gm + two * (dki * Dphi[j] + dkj * Dphi[i] - gd[i][j] * (gu[k][0] * Dphi[0] + gu[k][1] * Dphi[1]));
    }

  double s = zero;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const int i = kIJ[q][0], j = kIJ[q][1];
    s = s + u2 * gu[i][j] * (Gam[0][i][j] - Gam[1][i][j] * dh) / interp_psi4;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = kIJ[q][0], j = kIJ[q][1];
    s = s - DF[i] * DF[j] * (Gam[0][i][j] - Gam[1][i][j] * dh);
  }
  s = s + u * u * u * v[AF_trK];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = kIJ[q][0], j = kIJ[q][1];
    s = s - u * DF[i] * DF[j] * Kd[i][j];
  }
  s = -s / denom;
  if (std::fabs(th) < 1.0e-5 || std::fabs(th - smallpi) < 1.0e-5) s = zero;    // :2663-2667
  if (s != s || std::fabs(s) > 1.0e50) {                                         // :2671-2676
    sdh = 1.0e10;
    return 1;
  }
  sdh = s;
  sarea = two * smallpi * interp_psi4 * std::sqrt(std::fabs(gd[2][2] * (dh * dh * gd[0][0] + two * dh * gd[0][1] + gd[1][1])));
  return 0;
}

// Returns the error flag.  Every lane runs one interpolation, at a harmless
// point when the position checks fail, so that both half-waves of a shot
// always meet its barriers together.
__device__ int ah_source(const AhCtx& c, double rad, double dh, double th, double zcenter, double& sdh,
                         double& sarea, double* sv, AhCell& cell) {
  sdh = 0.0;
  if (!(th == cell.th)) {
    cell.th = th;
    cell.sth = std::sin(th);
    cell.cth = std::cos(th);
  }
  const AhPos P = ah_locate_sc(c, rad, cell.sth, cell.cth, zcenter);
  double v[AF_N];
  ah_interp(c.L[P.lev], c.interporder, P.r0, P.z0, v, sv, cell);
  if (P.err) return 1;
  return ah_source_math(rad, dh, th, P.rc, P.zc, v, sdh, sarea);
}

// ------------------------------------------------------------ one shot
// ahshoot :478-706: adaptive RK4 with step doubling from the axis point
// zinit.  The reference's "local extrapolation" (:672-677) writes the slot
// after the accepted one, which the next step overwrites before reading, so
// it has no effect and is not restated.
__device__ ShotOut ah_integrate(const AhCtx& c, double zinit, double zcenter, double* th_out, double* rad_out,
                                double* sv, double* xs, int* xe) {
  const double zero = 0.0, half = 0.5, two = 2.0;
  const double smallpi = M_PI;
  double rad = std::fabs(zinit - zcenter);
  double drad = 0.0, area = 0.0;
  double dth = smallpi / 100.0;
  double theta;
  if (zcenter >= zero) {
    theta = 1.0e-3;
  } else {
    dth = -dth;
    theta = smallpi - 1.0e-3;
  }
  double thetafinal;
  if (c.eqsym && zcenter == zero) thetafinal = half * smallpi;
  else if (zcenter >= zero) thetafinal = smallpi - 1.0e-3;
  else thetafinal = 1.0e-3;
  const bool rec = th_out != nullptr && threadIdx.x == 0;
  if (rec) {
    th_out[0] = theta;
    rad_out[0] = rad;
  }
  const double sixth = 1.0 / 6.0;
  int i = 0, err = 0;
  double sdh = 0.0, sarea = 0.0;
  AhCell cell;
  cell.q = nullptr;
  cell.i0 = cell.j0 = 0;
  cell.th = __builtin_nan("");
  for (int steps = 0; std::fabs(theta - thetafinal) > 1.0e-3; ++steps) {
    if (steps >= kAhMaxSteps) {
      err = 1;
      break;
    }
    if (zcenter >= zero) {
      if (theta + dth > thetafinal) dth = thetafinal - theta;
    } else {
      if (theta + dth < thetafinal) dth = thetafinal - theta;
    }
    const double dth2 = half * dth;
    double rad1, drad1, area1, drad_aux;
    double k_1, k_2, k_3, k_4, l_1, l_2, l_3, l_4, m_1, m_2, m_3, m_4;
    if ((err = ah_source(c, rad, drad, theta, zcenter, sdh, sarea, sv, cell))) break;
    k_1 = dth * drad;
    l_1 = dth * sdh;
    m_1 = std::fabs(dth2) * sarea;
    // The large step's last three stages and the first small step's last
    // three are independent: the upper half-wave runs the large step, the
    // lower half the small one, in one instruction stream (step h, first
    // stage k1/l1 per half, in the reference's expressions).  An error in
    // either ends the step as the reference's first error would: nothing of
    // the step is kept.
    {
      const bool big = (threadIdx.x >> 5) != 0;
      const double h = big ? dth : dth2;
      const double k1 = big ? k_1 : half * k_1;
      const double l1 = big ? l_1 : half * l_1;
      int e = ah_source(c, rad + 0.5 * k1, drad + 0.5 * l1, theta + 0.5 * h, zcenter, sdh, sarea, sv, cell);
      k_2 = h * (drad + 0.5 * l1);
      l_2 = h * sdh;
      m_2 = std::fabs(dth2) * sarea;
      e |= ah_source(c, rad + 0.5 * k_2, drad + 0.5 * l_2, theta + 0.5 * h, zcenter, sdh, sarea, sv, cell);
      k_3 = h * (drad + 0.5 * l_2);
      l_3 = h * sdh;
      m_3 = std::fabs(dth2) * sarea;
      e |= ah_source(c, rad + k_3, drad + l_3, theta + h, zcenter, sdh, sarea, sv, cell);
      k_4 = h * (drad + l_3);
      l_4 = h * sdh;
      m_4 = std::fabs(dth2) * sarea;
      // exchange: the small step's (rad1, drad1, area1), the large step's
      // drad_aux and both error flags reach every lane
      if (threadIdx.x == 0) {
        xs[0] = rad + sixth * (k1 + 2.0 * k_2 + 2.0 * k_3 + k_4);
        xs[1] = drad + sixth * (l1 + 2.0 * l_2 + 2.0 * l_3 + l_4);
        xs[2] = area + sixth * (m_1 + 2.0 * m_2 + 2.0 * m_3 + m_4);
        xe[0] = e;
      }
      if (threadIdx.x == 32) {
        xs[3] = drad + sixth * (l1 + 2.0 * l_2 + 2.0 * l_3 + l_4);
        xe[1] = e;
      }
      __syncthreads();
      err = xe[0] | xe[1];
      rad1 = xs[0];
      drad1 = xs[1];
      area1 = xs[2];
      drad_aux = xs[3];
      __syncthreads();
    }
    if (err) break;

    if ((err = ah_source(c, rad1, drad1, theta + dth, zcenter, sdh, sarea, sv, cell))) break;
    k_1 = dth2 * drad1;
    l_1 = dth2 * sdh;
    m_1 = std::fabs(dth2) * sarea;
    if ((err = ah_source(c, rad1 + 0.5 * k_1, drad1 + 0.5 * l_1, theta + 1.5 * dth2, zcenter, sdh, sarea, sv, cell))) break;
    k_2 = dth2 * (drad1 + 0.5 * l_1);
    l_2 = dth2 * sdh;
    m_2 = std::fabs(dth2) * sarea;
    if ((err = ah_source(c, rad1 + 0.5 * k_2, drad1 + 0.5 * l_2, theta + 1.5 * dth2, zcenter, sdh, sarea, sv, cell))) break;
    k_3 = dth2 * (drad1 + 0.5 * l_2);
    l_3 = dth2 * sdh;
    m_3 = std::fabs(dth2) * sarea;
    if ((err = ah_source(c, rad1 + k_3, drad1 + l_3, theta + dth, zcenter, sdh, sarea, sv, cell))) break;
    k_4 = dth2 * (drad1 + l_3);
    l_4 = dth2 * sdh;
    m_4 = std::fabs(dth2) * sarea;
    const double rad2 = rad1 + sixth * (k_1 + 2.0 * k_2 + 2.0 * k_3 + k_4);
    const double drad2 = drad1 + sixth * (l_1 + 2.0 * l_2 + 2.0 * l_3 + l_4);
    const double area2 = area1 + sixth * (m_1 + 2.0 * m_2 + 2.0 * m_3 + m_4);

    const double drad_error = std::fabs(drad2 - drad_aux);                       // :652-700
    const double eps = 1.0e-4 * (std::fabs(drad2) + std::fabs(drad_aux));
    if (drad_error > std::fmax(eps, 1.0e-8)) {
      dth = half * dth;
    } else {
      theta = theta + dth;
      rad = rad2;
      drad = drad2;
      area = area2;
      i = i + 1;
      if (rec) {
        th_out[i] = theta;
        rad_out[i] = rad;
      }
      if (i >= kAhNthetaMax) {
        err = 1;
        break;
      }
      if (drad_error < 0.1 * eps) dth = two * dth;
    }
  }
  ShotOut o;
  o.dradlast = drad;
  o.area = area;
  o.err = err;
  o.ntheta = i;
  return o;
}

// One shot per workgroup of one wave (kShotBlock = 64 lanes): the lanes
// share the loads of each interpolation, and the two halves of the wave run
// the independent stages of a step side by side (8 sequential source
// evaluations per step instead of 11); elsewhere every lane runs the same
// integration on the same values.
__global__ void __launch_bounds__(kShotBlock) k_ah_shoot(const ShotIn* __restrict__ in, ShotOut* __restrict__ out,
                                                         int n) {
  __shared__ double sv[2 * AF_N];
  __shared__ double xs[4];
  __shared__ int xe[2];
  const int t = static_cast<int>(blockIdx.x);
  if (t >= n) return;
  const ShotIn s = in[t];
  const ShotOut o = ah_integrate(*s.ctx, s.zinit, s.zcenter, s.theta, s.rad, sv, xs, xe);
  if (threadIdx.x == 0) out[t] = o;
}

// Dz_phi at one point of one level (the moving-centre walk of
// horizon_finder.f90:117-180), by one wave as in the shots.
__global__ void __launch_bounds__(kShotBlock) k_ah_dzphi(const AhCtx* __restrict__ c, int level, double r0,
                                                         double z0, double* __restrict__ out) {
  __shared__ double sv[2 * AF_N];
  double v[AF_N];
  AhCell cell;
  cell.q = nullptr;
  cell.i0 = cell.j0 = 0;
  cell.th = __builtin_nan("");
  ah_interp(c->L[level], c->interporder, r0, z0, v, sv, cell);
  if (threadIdx.x == 0) out[0] = v[AF_Dzphi];
}

// ------------------------------------------------------------ spectral finder
// ahspectral (:1113-1972): the fast flow of the cosine coefficients of the
// surface r = h(theta) (Gundlach, gr-qc/9707050).  One workgroup runs one
// search to the end.  Each stage evaluates ah_source at the Ntheta + 1
// points of the surface in two passes: the interpolation of the 20
// quantities, one point per half-wave (kSpecBlock / 32 points per round,
// the lanes sharing the stencil loads), into LDS; then the rest of the
// source, a long scalar chain, one lane per point, so that it is not run
// 32 times over.  One lane per mode then sums the trapezoid coefficients in
// the reference's order, every lane rebuilds a slice of h and h' from the
// coefficients, and lane 0 runs the residual and the step control.  The
// surface, the coefficients and their previous values stay in LDS for the
// whole search, and nothing returns to the host between iterations.
constexpr int kSpecModes = 50;                      // ac(0:50) (:1165-1168)
constexpr int kSpecThetaMax = 10 * kSpecModes;      // Ntheta = 10 Nmode (:1232)
constexpr int kSpecBlock = 512;                     // 16 surface points per interpolation round
constexpr int kSpecChunk = 128;                     // surface points per source pass (<= kSpecBlock)

struct SpecIn {
  const AhCtx* ctx;
  const double* cs;                   // cos(j theta_i), row j of ntheta + 1 values
  const double* sn;                   // sin(j theta_i)
  double* rad;                        // out: h(theta_i) of the last iteration
  double zcenter, dtheta, AA, BB, eps;
  int nmode, ntheta, step, maxiter;
  int reconstruct;                    // 1: h, h' from ac; 0: the circle h = ac(0)/2
  double ac[kSpecModes + 1];          // initial coefficients
};

struct SpecOut {
  double ac[kSpecModes + 1];
  double res, area;                   // residual; trapezoid of the last area elements
  int iter, err;
};

// The coefficient sums and the rebuilds read the cos / sin tables term after
// term; TAB_LDS copies them into LDS first (when they fit, kSpecTabLds), so
// each term waits for an LDS read instead of an L2 one.
constexpr int kSpecTabLds = 64 * 1024;

template <bool TAB_LDS>
__global__ void __launch_bounds__(kSpecBlock) k_ah_spectral(const SpecIn* __restrict__ in,
                                                            SpecOut* __restrict__ out) {
  extern __shared__ double tab[];          // TAB_LDS: cos(j theta_i), then sin(j theta_i)
  __shared__ double vals[kSpecChunk * AF_N];          // interpolated quantities of a chunk
  __shared__ AhPlan plans[kSpecChunk];                // their interpolation plans
  __shared__ int perr[kSpecChunk];
  __shared__ double rad[kSpecThetaMax + 1], drad[kSpecThetaMax + 1];
  __shared__ double rad_p[kSpecThetaMax + 1], drad_p[kSpecThetaMax + 1];
  __shared__ double Ifunc[kSpecThetaMax + 1], area[kSpecThetaMax + 1];
  __shared__ double ac[kSpecModes + 1], cc[kSpecModes + 1], ac_old[kSpecModes + 1], cc_old[kSpecModes + 1];
  __shared__ double x_res, x_AA;
  __shared__ int x_err, x_rej;
  const SpecIn& s = in[blockIdx.x];
  const AhCtx& c = *s.ctx;
  const int t = static_cast<int>(threadIdx.x);
  const int nt = static_cast<int>(blockDim.x);
  const int N = s.ntheta, nm = s.nmode, step = s.step;
  const double zero = 0.0, half = 0.5, one = 1.0, two = 2.0;
  const double smallpi = M_PI;
  const double dtheta = s.dtheta, BB = s.BB, AA0 = s.AA, zcenter = s.zcenter;
  const int ntab = (nm + 1) * (N + 1);
  if (TAB_LDS)
    for (int k = t; k < ntab; k += nt) {
      tab[k] = s.cs[k];
      tab[ntab + k] = s.sn[k];
    }
  const double* __restrict__ cs_g = s.cs;
  const double* __restrict__ sn_g = s.sn;
  auto CS = [&](int k) { return TAB_LDS ? tab[k] : cs_g[k]; };
  auto SN = [&](int k) { return TAB_LDS ? tab[ntab + k] : sn_g[k]; };
  // h(theta_i), h'(theta_i) from the coefficients (:1320-1327, :1602-1609)
  auto rebuild = [&]() {
    for (int i = t; i <= N; i += nt) {
      double r = half * ac[0], d = zero;
#pragma unroll 4
      for (int j = step; j <= nm; j += step) {
        r = r + ac[j] * CS(j * (N + 1) + i);
        d = d - static_cast<double>(j) * ac[j] * SN(j * (N + 1) + i);
      }
      rad[i] = r;
      drad[i] = d;
    }
  };
  for (int j = t; j <= kSpecModes; j += nt) {
    ac[j] = s.ac[j];
    cc[j] = ac_old[j] = cc_old[j] = zero;
  }
  __syncthreads();
  if (s.reconstruct) {
    rebuild();
  } else {
    for (int i = t; i <= N; i += nt) {
      rad[i] = half * ac[0];
      drad[i] = zero;
    }
  }
  for (int i = t; i <= N; i += nt) Ifunc[i] = area[i] = zero;
  const int hw = t >> 5, nhw = nt >> 5;
  const int jm = t * step;                 // this lane's mode in the coefficient sums
  double res = 1.0, AA = AA0;
  int iter = 0, err = 0;
  __syncthreads();
  while (res > s.eps && iter < s.maxiter) {                                      // :1451
    ++iter;
    for (int j = t; j <= kSpecModes; j += nt) {
      ac_old[j] = ac[j];
      cc_old[j] = cc[j];
    }
    for (int i = t; i <= N; i += nt) {
      rad_p[i] = rad[i];
      drad_p[i] = drad[i];
    }
    if (t == 0) x_err = 0;
    __syncthreads();
    for (int q = 1; q <= 2; ++q) {
      // sources at every point (:1494-1501), kSpecChunk points at a time, in
      // three passes: each point's position checks, level and interpolation
      // plan (one lane per point); its 20 quantities (one half-wave per
      // point, the lanes sharing the stencil loads); the rest of ah_source
      // (one lane per point)
      for (int c0 = 0; c0 <= N; c0 += kSpecChunk) {
        const int cn = (N + 1 - c0 < kSpecChunk) ? N + 1 - c0 : kSpecChunk;
        if (t < cn) {
          const int i = c0 + t;
          const AhPos P = ah_locate(c, rad[i], static_cast<double>(i) * dtheta, zcenter);
          AhPlan pl;
          ah_plan(c.L[P.lev], c.interporder, P.r0, P.z0, pl);
          pl.lev = P.lev;
          plans[t] = pl;
          perr[t] = P.err;
        }
        __syncthreads();
        for (int b = 0; b < cn; b += nhw) {
          const int k = b + hw;
          const int f = t & 31;
          if (k < cn && f < AF_N) {
            const AhPlan& pl = plans[k];
            double fa[4][4];
            ah_records(c.L[pl.lev], pl.i0, pl.j0, f, fa);
            vals[k * AF_N + f] = ah_combine(pl, fa);
          }
        }
        __syncthreads();
        if (t < cn) {
          const int i = c0 + t;
          const double th = static_cast<double>(i) * dtheta;
          double sdh = 0.0, sarea = 0.0;
          int e = perr[t];
          if (!e) {
            double v[AF_N];
#pragma unroll
            for (int f = 0; f < AF_N; ++f) v[f] = vals[t * AF_N + f];
            e = ah_source_math(rad[i], drad[i], th, rad[i] * std::sin(th), rad[i] * std::cos(th), v, sdh, sarea);
          }
          Ifunc[i] = sdh;
          area[i] = sarea;
          if (e) x_err = 1;
        }
        __syncthreads();
      }
      err = x_err;
      if (err) break;                                                            // goto 100
      if (t == 0) {                                                              // :1506-1515
        Ifunc[0] = Ifunc[1];
        area[0] = zero;
        if (step == 1) {
          Ifunc[N] = Ifunc[N - 1];
          area[N] = zero;
        }
      }
      __syncthreads();
      if (jm <= nm) {                                                            // :1539-1563, :1590-1592
        const int cj = jm * (N + 1);
        double cc_j = half * dtheta * Ifunc[0];
        // unrolled so that the reads run ahead of the (serial) sum
#pragma unroll 8
        for (int i = 1; i < N; ++i) cc_j = cc_j + dtheta * Ifunc[i] * CS(cj + i);
        cc_j = cc_j + half * dtheta * Ifunc[N] * CS(cj + N);
        cc_j = two * static_cast<double>(step) * cc_j / smallpi;
        const double jj = static_cast<double>(jm) * static_cast<double>(jm);
        cc_j = cc_j + jj * ac[jm];
        cc[jm] = cc_j;
        if (q == 1) cc_old[jm] = cc_j;
        ac[jm] = ac_old[jm] - half * AA * static_cast<double>(q) / (one + BB * jj) * cc_j;
      }
      __syncthreads();
      rebuild();
      __syncthreads();
    }
    if (err) break;
    if (t == 0) {                                                                // :1636-1682
      double r = zero, aux1 = zero, aux2 = zero;
      for (int j = 0; j <= nm; j += step) r = std::fmax(r, std::fabs(cc[j]));
      for (int j = 0; j <= nm; j += step) {
        const double jj = static_cast<double>(j) * static_cast<double>(j);
        const double tc = ac_old[j] - AA / (one + BB * jj) * cc_old[j];
        aux1 = aux1 + std::fabs(ac[j] - tc);
        aux2 = aux2 + std::fabs(ac_old[j]);
      }
      int rej = 0;
      double AAn = AA;
      if (aux1 > 0.05 * aux2) {
        AAn = half * AA;
        rej = 1;
      } else if (aux1 < 0.001 * aux2) {
        AAn = std::fmin(1.5 * AA, AA0);
      }
      x_res = r;
      x_AA = AAn;
      x_rej = rej;
    }
    __syncthreads();
    res = x_res;
    AA = x_AA;
    if (x_rej) {
      if (jm <= nm) ac[jm] = ac_old[jm];
      for (int i = t; i <= N; i += nt) {
        rad[i] = rad_p[i];
        drad[i] = drad_p[i];
      }
    }
    __syncthreads();
  }
  SpecOut& o = out[blockIdx.x];
  for (int j = t; j <= kSpecModes; j += nt) o.ac[j] = ac[j];
  for (int i = t; i <= N; i += nt) s.rad[i] = rad[i];
  if (t == 0) {                                                                  // :1833-1839
    double a = half * dtheta * area[0];
    for (int i = 1; i < N; ++i) a = a + dtheta * area[i];
    a = a + half * dtheta * area[N];
    o.area = a;
    o.res = res;
    o.iter = iter;
    o.err = err;
  }
}

// ================================================================ host side
int hip_rc(hipError_t e, const char* what) {
  if (e == hipSuccess) return OA_OK;
  oa_set_error(std::string("oa_ahfind: ") + what + ": " + hipGetErrorString(e));
  return OA_ERR_HIP;
}
#define AH_HIP(call)                             \
  do {                                           \
    int rc_ = hip_rc((call), #call);             \
    if (rc_ != OA_OK) return rc_;                \
  } while (0)
#define AH_RC(call)                              \
  do {                                           \
    int rc_ = (call);                            \
    if (rc_ != OA_OK) return rc_;                \
  } while (0)

struct ShotRes {
  double dradlast, area;
  int err, ntheta;
};

uint64_t key_of(double z) {
  uint64_t k;
  std::memcpy(&k, &z, sizeof k);
  return k;
}

}  // namespace

// Per-handle state of the finder: derived planes, the device context, the
// shot tables, ahfound(k) (kept between searches as the reference keeps it)
// and the last integration of each horizon.
struct AhState {
  std::vector<double*> planes;            // point records of every level
  AhCtx hctx{};
  AhCtx* dctx = nullptr;
  ShotIn* h_in = nullptr;                 // pinned
  ShotOut* h_out = nullptr;               // pinned
  ShotIn* d_in = nullptr;
  ShotOut* d_out = nullptr;
  int cap = 0;
  double* d_theta[3] = {nullptr, nullptr, nullptr};   // last integration of horizon k
  double* d_rad[3] = {nullptr, nullptr, nullptr};
  int found[3] = {0, 0, 0};
  // The last integration of horizon k is recorded on demand (oa_horizon_shape)
  // by integrating it again on the point records it was found on: gen counts
  // rebuilds of the records, gen_of[k] is the one horizon k's search used.
  long gen = 0;
  long gen_of[3] = {-1, -1, -1};
  double last_z[3] = {0.0, 0.0, 0.0}, last_zc[3] = {0.0, 0.0, 0.0};
  int n_of[3] = {0, 0, 0};
  double* d_scalar = nullptr;             // one value back from the device
  // spectral finder: ack(k,:) (shared with the shooting search, which
  // projects its surface onto it), its own first-call flags, the surface of
  // its last iteration per horizon, the cos / sin tables of the current
  // (Nmode, step) and the staging of its launches
  double ack[3][kSpecModes + 1] = {};
  int first_sp[3] = {1, 1, 1};
  int shape_spec[3] = {0, 0, 0};          // horizon k's last search ended with the fast flow
  double spec_dtheta[3] = {0.0, 0.0, 0.0};
  double* d_srad[3] = {nullptr, nullptr, nullptr};
  double* d_cs = nullptr;
  double* d_sn = nullptr;
  int tab_nmode = 0, tab_step = 0;
  SpecIn* h_sin = nullptr;                // pinned
  SpecOut* h_sout = nullptr;              // pinned
  SpecIn* d_sin = nullptr;
  SpecOut* d_sout = nullptr;
  int scap = 0;
  hipStream_t search_stream = nullptr;    // the shots of a begun search (oa_*ahfind_begin)
  hipEvent_t fields_done = nullptr;       // the point records of a begun search are formed
  int pending = 0;                        // a begun search awaits its end
};

namespace {

// finest level holding the axis point (0, zinit) one cell inside (:462-476)
double dz_on_axis(const AhCtx& c, double zinit) {
  int level = 0;
  for (int ll = 1; ll < c.nlev; ++ll) {
    const AhLevel& L = c.L[ll];
    if (0.0 > L.rmin + L.G.dr && 0.0 < L.rmax - L.G.dr && zinit > L.zmin + L.G.dz && zinit < L.zmax - L.G.dz)
      level = ll;
  }
  return c.L[level].G.dz;
}

int ensure_capacity(AhState& a, int n) {
  if (n <= a.cap) return OA_OK;
  int cap = a.cap ? a.cap : 256;
  while (cap < n) cap *= 2;
  if (a.h_in) (void)hipHostFree(a.h_in);
  if (a.h_out) (void)hipHostFree(a.h_out);
  if (a.d_in) (void)hipFree(a.d_in);
  if (a.d_out) (void)hipFree(a.d_out);
  a.h_in = nullptr;
  a.h_out = nullptr;
  a.d_in = nullptr;
  a.d_out = nullptr;
  a.cap = 0;
  AH_HIP(hipHostMalloc(reinterpret_cast<void**>(&a.h_in), sizeof(ShotIn) * cap, hipHostMallocDefault));
  AH_HIP(hipHostMalloc(reinterpret_cast<void**>(&a.h_out), sizeof(ShotOut) * cap, hipHostMallocDefault));
  AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_in), sizeof(ShotIn) * cap));
  AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_out), sizeof(ShotOut) * cap));
  a.cap = cap;
  return OA_OK;
}

// Integrate the shots `in` in one launch, staged through buf's tables.
int run_shots(hipStream_t st, AhState& buf, const std::vector<ShotIn>& in, std::vector<ShotRes>& res) {
  const int n = static_cast<int>(in.size());
  res.assign(static_cast<size_t>(n), ShotRes{0.0, 0.0, 0, 0});
  if (n == 0) return OA_OK;
  AH_RC(ensure_capacity(buf, n));
  std::memcpy(buf.h_in, in.data(), sizeof(ShotIn) * n);
  AH_HIP(hipMemcpyAsync(buf.d_in, buf.h_in, sizeof(ShotIn) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_ah_shoot, dim3(n), dim3(kShotBlock), 0, st, buf.d_in, buf.d_out, n);
  AH_HIP(hipGetLastError());
  AH_HIP(hipMemcpyAsync(buf.h_out, buf.d_out, sizeof(ShotOut) * n, hipMemcpyDeviceToHost, st));
  AH_HIP(hipStreamSynchronize(st));
  for (int t = 0; t < n; ++t)
    res[static_cast<size_t>(t)] = ShotRes{buf.h_out[t].dradlast, buf.h_out[t].area, buf.h_out[t].err, buf.h_out[t].ntheta};
  return OA_OK;
}

// One search (ahshoot :377-855) as a walk over a table of shot results.
struct Search {
  const AhCtx* c;
  double zcenter, eps;
  int found_in;
  std::map<uint64_t, ShotRes> table;
  // outcome of a complete walk
  int found = 0, errorflag = 0, shots = 0;
  double zinit = 0.0, dradlast = 0.0, last_zinit = 0.0;
  ShotRes last{0.0, 0.0, 0, 0};
};

// Walk the reference's control flow from the start.  Returns true when the
// search is complete; otherwise `want` holds the shots to integrate next
// (the missing one first, then its speculative continuation).
bool walk(Search& s, int depth, std::vector<double>& want) {
  const AhCtx& c = *s.c;
  const AhLevel& L0 = c.L[0];
  const double zero = 0.0, half = 0.5, two = 2.0;
  const double zcenter = s.zcenter;
  const bool up = zcenter >= zero;
  want.clear();

  double zinit_old = zero, zinit;
  // ahzinit is resolved by the caller into s.zinit's start (see ah_search)
  zinit = s.zinit;
  double dzlocal = L0.G.dz, dzz = half * dzlocal;
  double dradlast = 0.0, dradlast_old = 0.0, drada = 0.0, za, zb;
  int errorflag = 0, shots = 0, found = s.found_in;
  ShotRes last{0.0, 0.0, 0, 0};
  double last_zinit = zinit;

  for (;;) {                                                          // 50 continue
    bool bracket = false, firstiter = true;
    errorflag = 0;
    za = 1.0e10;
    zb = -1.0e10;
    while (std::fabs(zinit - zcenter) > 5.0 * dzlocal && std::fabs(zinit_old - zinit) > s.eps) {
      dzlocal = dz_on_axis(c, zinit);
      dzz = half * dzlocal;
      dradlast_old = firstiter ? 1.0e10 : dradlast;
      const auto it = s.table.find(key_of(zinit));
      if (it == s.table.end()) {
        // speculate from here
        if (!bracket) {
          // the march: zinit, zinit -/+ dzz(zinit), ... while the loop
          // condition would still hold
          double z = zinit, zold = zinit_old, dzl = dzlocal;
          for (int q = 0; q < kAhScanChunk; ++q) {
            if (q > 0 && !(std::fabs(z - zcenter) > 5.0 * dzl && std::fabs(zold - z) > s.eps)) break;
            want.push_back(z);
            dzl = dz_on_axis(c, z);
            zold = z;
            z = up ? z - half * dzl : z + half * dzl;
          }
        } else {
          // the bisection: every node of the next `depth` steps; node
          // (a, b) shoots half*(a+b) and continues in (a, m) or (m, b)
          std::vector<std::pair<double, double>> level{{za, zb}};
          for (int d = 0; d < depth; ++d) {
            std::vector<std::pair<double, double>> next;
            for (const auto& ab : level) {
              const double m = half * (ab.first + ab.second);
              want.push_back(m);
              next.push_back({ab.first, m});
              next.push_back({m, ab.second});
            }
            level.swap(next);
          }
        }
        return false;
      }
      last = it->second;
      last_zinit = zinit;
      ++shots;
      errorflag = last.err;
      dradlast = last.dradlast;
      if (firstiter || (!bracket && dradlast * dradlast_old >= zero)) {  // :739-792
        firstiter = false;
        zinit_old = zinit;
        zinit = up ? zinit - dzz : zinit + dzz;
      } else if (!bracket) {
        bracket = true;
        za = zinit_old;
        zb = zinit;
        drada = dradlast_old;
        zinit_old = zinit;
        zinit = half * (za + zb);
      } else {
        if (drada * dradlast <= zero) {
          zb = zinit;
        } else {
          za = zinit;
          drada = dradlast;
        }
        zinit_old = zinit;
        zinit = half * (za + zb);
      }
    }
    if (std::fabs(zinit - zinit_old) <= s.eps) {                       // :797-855
      if (!errorflag && std::fabs(dradlast) < 1.0e-1) {
        found = 1;
      } else if (std::fabs(zinit - zcenter) > dzlocal) {
        zinit = up ? zinit - two * dzz : zinit + two * dzz;
        continue;
      }
    } else {
      found = 0;
    }
    break;
  }
  s.found = found;
  s.errorflag = errorflag;
  s.shots = shots;
  s.zinit = zinit;
  s.dradlast = dradlast;
  s.last = last;
  s.last_zinit = last_zinit;
  return true;
}

int build_fields(oa_handle* h, AhState& a) {
  const int nlev = h->nlev;
  if (static_cast<int>(a.planes.size()) != nlev) {
    a.planes.assign(static_cast<size_t>(nlev), nullptr);
    for (int l = 0; l < nlev; ++l) {
      const size_t bytes = sizeof(double) * AF_N * static_cast<size_t>(h->L[l].G.plane);
      if (hipMalloc(reinterpret_cast<void**>(&a.planes[static_cast<size_t>(l)]), bytes) != hipSuccess) {
        oa_set_error("oa_ahfind: device allocation of the horizon planes failed");
        return OA_ERR_NOMEM;
      }
      AH_HIP(hipMemsetAsync(a.planes[static_cast<size_t>(l)], 0, bytes, h->stream));
    }
    AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.dctx), sizeof(AhCtx)));
  }
  const oa_params& p = h->p;
  ++a.gen;
  AhCtx& c = a.hctx;
  std::memset(&c, 0, sizeof c);
  c.nlev = nlev;
  c.eqsym = p.eqsym;
  c.interporder = p.interporder;
  c.method = p.ahmethod;
  c.ahrmax = (p.ahrmax > 0.9) ? 0.9 : p.ahrmax;
  for (int l = 0; l < nlev; ++l) {
    Level& L = h->L[l];
    const KGrid& G = L.G;
    const double* y = L.slab[L.in_step ? L.last_out : L.cur];
    double* out = a.planes[static_cast<size_t>(l)];
    const dim3 blk(64, 4, 1);
    const dim3 grd((G.Nz - G.jlo + 1 + 63) / 64, (G.Nr + 3) / 4, 1);
    const int from_chi = L.phi_valid ? 0 : 1;
    if (L.aux_phi) hipLaunchKernelGGL((k_ah_fields<true>), grd, blk, 0, h->stream, G, y, from_chi, out);
    else hipLaunchKernelGGL((k_ah_fields<false>), grd, blk, 0, h->stream, G, y, from_chi, out);
    AH_HIP(hipGetLastError());
    AhLevel& A = c.L[l];
    A.G = G;
    const int g = G.g;
    A.rmin = (static_cast<double>(1 - g) - 0.5) * G.dr;
    A.rmax = (static_cast<double>(G.Nr) - 0.5) * G.dr;
    A.zmin = (static_cast<double>(1 - g) - G.zoff) * G.dz;
    A.zmax = (static_cast<double>(G.Nz) - G.zoff) * G.dz;
    A.q = out;
  }
  AH_HIP(hipMemcpyAsync(a.dctx, &a.hctx, sizeof(AhCtx), hipMemcpyHostToDevice, h->stream));
  return OA_OK;
}

}  // namespace

int ah_record_doubles() { return AF_N; }

// the context, the recorded last integration of each horizon and the shot
// tables at their usual size (one march chunk); with the spectral finder
// also its surfaces, trigonometric tables and one search's staging
double ah_fixed_bytes(int n_ah, int method) {
  double b = static_cast<double>(sizeof(AhCtx)) + 2.0 * n_ah * sizeof(double) * (kAhNthetaMax + 1) +
             static_cast<double>(kAhScanChunk) * 2.0 * (sizeof(ShotIn) + sizeof(ShotOut));
  if (method != OA_AH_SHOOT)
    b += n_ah * sizeof(double) * (kSpecThetaMax + 1) + 2.0 * sizeof(double) * (kSpecModes + 1) * (kSpecThetaMax + 1) +
         static_cast<double>(n_ah) * (sizeof(SpecIn) + sizeof(SpecOut));
  return b;
}

void ah_release(oa_handle* h) {
  AhState* a = h->ah;
  if (!a) return;
  for (double* p : a->planes)
    if (p) (void)hipFree(p);
  if (a->dctx) (void)hipFree(a->dctx);
  if (a->d_scalar) (void)hipFree(a->d_scalar);
  for (int k = 0; k < 3; ++k) {
    if (a->d_theta[k]) (void)hipFree(a->d_theta[k]);
    if (a->d_rad[k]) (void)hipFree(a->d_rad[k]);
  }
  if (a->d_in) (void)hipFree(a->d_in);
  if (a->d_out) (void)hipFree(a->d_out);
  if (a->h_in) (void)hipHostFree(a->h_in);
  if (a->h_out) (void)hipHostFree(a->h_out);
  for (int k = 0; k < 3; ++k)
    if (a->d_srad[k]) (void)hipFree(a->d_srad[k]);
  if (a->d_cs) (void)hipFree(a->d_cs);
  if (a->d_sn) (void)hipFree(a->d_sn);
  if (a->d_sin) (void)hipFree(a->d_sin);
  if (a->d_sout) (void)hipFree(a->d_sout);
  if (a->h_sin) (void)hipHostFree(a->h_sin);
  if (a->h_sout) (void)hipHostFree(a->h_sout);
  if (a->search_stream) (void)hipStreamDestroy(a->search_stream);
  if (a->fields_done) (void)hipEventDestroy(a->fields_done);
  delete a;
  h->ah = nullptr;
}

int ah_pending(const oa_handle* h) { return h->ah ? h->ah->pending : 0; }

}  // namespace oax

using namespace oax;

namespace {

// horizon_finder's checks (:48-95) and this library's scope
int ah_check(const oa_handle* h, int k) {
  const oa_params& p = h->p;
  if (p.N_ah > 3 || p.N_ah < 1 || k < 0 || k >= p.N_ah) {
    oa_set_error("oa_ahfind: horizon index out of range (N_ah <= 3)");
    return OA_ERR_ARG;
  }
  if (p.eqsym && p.N_ah == 3) {
    oa_set_error("oa_ahfind: for equatorial symmetry we can only look for 2 horizons");
    return OA_ERR_PARAM;
  }
  if (p.z_ah1 < 0.0 || p.z_ah2 > 0.0) {
    oa_set_error("oa_ahfind: horizon 1 must be above the equator and horizon 2 below");
    return OA_ERR_PARAM;
  }
  if (p.ahmethod != OA_AH_SHOOT && (p.ahNmodes < 1 || p.ahNmodes > kSpecModes)) {
    oa_set_error("oa_ahfind: ahNmodes must lie in 1..50 for the spectral finder");
    return OA_ERR_PARAM;
  }
  if (h->nlev > kAhMaxLevels) {
    oa_set_error("oa_ahfind: too many levels");
    return OA_ERR_UNSUPPORTED;
  }
  for (const Level& L : h->L)
    if (L.dirty) {
      oa_set_error("oa_ahfind: state not projected (call oa_auxiliary / oa_initial_aux)");
      return OA_ERR_STATE;
    }
  return OA_OK;
}

// Moving centre (horizon_finder.f90:117-180): from *z, step along the axis
// by dz0 towards larger phi until Dz_phi changes sign; the centre is the
// linear zero of Dz_phi.  Dz_phi is interpolated on the device, on the
// finest level holding the starting point, one value per step.  A walk that
// leaves that level returns OA_ERR_STATE (the reference stops in interp).
int move_centre(oa_handle* h, AhState& a, double* z) {
  const AhCtx& c = a.hctx;
  const double r0 = 0.0;
  double z0 = *z, z0_old = 0.0, Dzphi = 0.0, Dzphi_old = 0.0;
  int level = 0;
  for (int ll = 1; ll < c.nlev; ++ll) {
    const AhLevel& L = c.L[ll];
    if (r0 > L.rmin + L.G.dr && r0 < L.rmax - L.G.dr && z0 > L.zmin + L.G.dz && z0 < L.zmax - L.G.dz) level = ll;
  }
  const AhLevel& L = c.L[level];
  if (!a.d_scalar) AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_scalar), sizeof(double)));
  for (int it = 0;; ++it) {
    if (it > 1000000 || z0 < L.zmin || z0 > L.zmax) {
      oa_set_error("oa_ahfind: the moving-centre walk left the grid");
      return OA_ERR_STATE;
    }
    hipLaunchKernelGGL(k_ah_dzphi, dim3(1), dim3(kShotBlock), 0, h->stream, a.dctx, level, r0, z0, a.d_scalar);
    AH_HIP(hipGetLastError());
    AH_HIP(hipMemcpyAsync(&Dzphi, a.d_scalar, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    AH_HIP(hipStreamSynchronize(h->stream));
    if (Dzphi * Dzphi_old >= 0.0) {
      z0_old = z0;
      Dzphi_old = Dzphi;
      z0 = (Dzphi > 0.0) ? z0 + h->p.dz0 : z0 - h->p.dz0;
    } else {
      *z = z0_old - Dzphi_old * (z0 - z0_old) / (Dzphi - Dzphi_old);
      return OA_OK;
    }
  }
}

struct Job {
  oa_handle* h;
  int k;
  Search s;
  double zstart;
  bool done;
  // shooting searches this call runs: before the flow (ahmethod = both,
  // :223-226) and inside it (ahshootfirst on the first call, :1292-1303)
  bool shoot_first = false, shoot_in_flow = false, flow = false;
  int shots = 0;                           // of the last shooting search
  double zinit = 0.0, dradlast = 0.0;
  SpecOut spec{};
};

// First part of a search: the point records of every job's spacetime, formed
// on its stream from the current state (the only part that reads the state),
// and the start of each job's walk.
int search_prepare(std::vector<Job>& jobs) {
  for (Job& j : jobs) {
    AH_RC(ah_check(j.h, j.k));
    if (!j.h->ah) j.h->ah = new AhState();
    if (j.h->ah->pending) {
      oa_set_error("oa_ahfind: a begun search of this spacetime has not ended");
      return OA_ERR_STATE;
    }
  }
  for (Job& j : jobs) {
    AhState& a = *j.h->ah;
    AH_RC(build_fields(j.h, a));
    const oa_params& p = j.h->p;
    double zah[3] = {0.0, 0.0, 0.0};                                         // :104-111
    if (p.N_ah > 1) {
      zah[1] = p.z_ah1;
      zah[2] = p.z_ah2;
    }
    if (p.ahmove && p.N_ah > 1 && j.k > 0) AH_RC(move_centre(j.h, a, &zah[j.k]));
    const double zcenter = zah[j.k];
    const AhLevel& L0 = a.hctx.L[0];
    j.s.c = &a.hctx;
    j.s.zcenter = zcenter;
    j.s.eps = p.ahepsilon_shoot;
    j.s.found_in = a.found[j.k];
    if (p.ahzinit == 0.0) {                                                  // :379-405
      j.zstart = (zcenter >= 0.0) ? 0.8 * L0.zmax : 0.8 * L0.zmin;
    } else if (p.ahzinit > 0.0 && p.ahzinit <= 0.9 * L0.zmax) {
      j.zstart = (zcenter >= 0.0) ? p.ahzinit : -p.ahzinit;
    } else {
      j.zstart = (zcenter >= 0.0) ? 0.8 * L0.zmax : 0.8 * L0.zmin;
    }
    j.flow = p.ahmethod != OA_AH_SHOOT;
    j.shoot_first = p.ahmethod != OA_AH_SPECTRAL;
    j.shoot_in_flow = j.flow && a.first_sp[j.k] && p.ahshootfirst;
    j.done = !j.shoot_first;
  }
  return OA_OK;
}

// The walks of the jobs not yet done, in lockstep on stream st (every round
// integrates the speculative shots of all unfinished searches in one
// launch), then each one's outcome: ahfound(k) and where to integrate its
// last shot again.
int run_walks(std::vector<Job>& jobs, hipStream_t st) {
  AhState& buf = *jobs[0].h->ah;
  std::vector<double> want;
  std::vector<ShotIn> in;
  std::vector<std::pair<size_t, double>> owner;   // (job, zinit) of every launched shot
  std::vector<ShotRes> res;
  std::vector<bool> ran(jobs.size());
  for (size_t q = 0; q < jobs.size(); ++q) ran[q] = !jobs[q].done;
  for (;;) {
    in.clear();
    owner.clear();
    int live = 0;
    for (const Job& j : jobs) live += j.done ? 0 : 1;
    int depth = 1;
    while (depth < kAhMaxDepth && live * ((2 << depth) - 1) <= kAhResident) ++depth;
    for (size_t q = 0; q < jobs.size(); ++q) {
      Job& j = jobs[q];
      if (j.done) continue;
      j.s.zinit = j.zstart;
      if (walk(j.s, depth, want)) {
        j.done = true;
        continue;
      }
      for (double z : want) {
        in.push_back(ShotIn{j.h->ah->dctx, z, j.s.zcenter, nullptr, nullptr});
        owner.push_back({q, z});
      }
    }
    if (in.empty()) break;
    AH_RC(run_shots(st, buf, in, res));
    for (size_t t = 0; t < in.size(); ++t) jobs[owner[t].first].s.table[key_of(owner[t].second)] = res[t];
  }
  for (size_t q = 0; q < jobs.size(); ++q) {
    if (!ran[q]) continue;
    Job& j = jobs[q];
    AhState& a = *j.h->ah;
    a.found[j.k] = j.s.found;
    j.shots = j.s.shots;
    j.zinit = j.s.zinit;
    j.dradlast = j.s.dradlast;
    // where to integrate the last shot again when its points are asked for
    a.gen_of[j.k] = j.s.shots > 0 ? a.gen : -1;
    a.last_z[j.k] = j.s.last_zinit;
    a.last_zc[j.k] = j.s.zcenter;
    a.n_of[j.k] = j.s.shots > 0 ? j.s.last.ntheta + 1 : 0;
    a.shape_spec[j.k] = 0;
  }
  return OA_OK;
}

// ahshoot's projection of its last surface onto ack(k,0..20) (:872-910),
// for every job whose shooting search just found a horizon: the last
// integration of each runs again (one launch) recording its points, and
// the trapezoids over those uneven points run here in the reference's order.
int project_shots(std::vector<Job*>& js, hipStream_t st) {
  if (js.empty()) return OA_OK;
  AhState& buf = *js[0]->h->ah;
  std::vector<ShotIn> in;
  for (Job* j : js) {
    AhState& a = *j->h->ah;
    if (!a.d_theta[j->k]) {
      AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_theta[j->k]), sizeof(double) * (kAhNthetaMax + 1)));
      AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_rad[j->k]), sizeof(double) * (kAhNthetaMax + 1)));
    }
    in.push_back(ShotIn{a.dctx, a.last_z[j->k], a.last_zc[j->k], a.d_theta[j->k], a.d_rad[j->k]});
  }
  std::vector<ShotRes> res;
  AH_RC(run_shots(st, buf, in, res));
  std::vector<double> th, rd;
  const double half = 0.5, two = 2.0, smallpi = M_PI;
  for (size_t q = 0; q < js.size(); ++q) {
    Job& j = *js[q];
    AhState& a = *j.h->ah;
    const int n = res[q].ntheta;
    if (n + 1 != a.n_of[j.k]) {
      oa_set_error("oa_ahfind: the repeated integration differs from the search's");
      return OA_ERR_STATE;
    }
    th.resize(static_cast<size_t>(n) + 1);
    rd.resize(static_cast<size_t>(n) + 1);
    AH_HIP(hipMemcpyAsync(th.data(), a.d_theta[j.k], sizeof(double) * (n + 1), hipMemcpyDeviceToHost, st));
    AH_HIP(hipMemcpyAsync(rd.data(), a.d_rad[j.k], sizeof(double) * (n + 1), hipMemcpyDeviceToHost, st));
    AH_HIP(hipStreamSynchronize(st));
    const int iaux = (j.h->p.eqsym && j.s.zcenter == 0.0) ? 2 : 1;
    double* ack = a.ack[j.k];
    for (int m = 0; m <= kSpecModes; ++m) ack[m] = 0.0;
    for (int m = 0; m <= 20; m += iaux) {
      for (int i = 0; i < n; ++i)
        ack[m] = ack[m] + half * (th[i + 1] - th[i]) *
                              (rd[i + 1] * std::cos(static_cast<double>(m) * th[i + 1]) +
                               rd[i] * std::cos(static_cast<double>(m) * th[i]));
      ack[m] = two * static_cast<double>(iaux) * ack[m] / smallpi;
    }
  }
  return OA_OK;
}

int spec_capacity(AhState& a, int n) {
  if (n <= a.scap) return OA_OK;
  if (a.h_sin) (void)hipHostFree(a.h_sin);
  if (a.h_sout) (void)hipHostFree(a.h_sout);
  if (a.d_sin) (void)hipFree(a.d_sin);
  if (a.d_sout) (void)hipFree(a.d_sout);
  a.h_sin = nullptr;
  a.h_sout = nullptr;
  a.d_sin = nullptr;
  a.d_sout = nullptr;
  a.scap = 0;
  AH_HIP(hipHostMalloc(reinterpret_cast<void**>(&a.h_sin), sizeof(SpecIn) * n, hipHostMallocDefault));
  AH_HIP(hipHostMalloc(reinterpret_cast<void**>(&a.h_sout), sizeof(SpecOut) * n, hipHostMallocDefault));
  AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_sin), sizeof(SpecIn) * n));
  AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_sout), sizeof(SpecOut) * n));
  a.scap = n;
  return OA_OK;
}

// cos(j theta_i) and sin(j theta_i), theta_i = i pi / (step Ntheta), as the
// reference evaluates them (theta(i) first, then j times it)
int spec_tables(AhState& a, int nmode, int step, hipStream_t st) {
  if (a.d_cs && a.tab_nmode == nmode && a.tab_step == step) return OA_OK;
  const int N = 10 * nmode;
  const double dtheta = M_PI / static_cast<double>(step * N);
  std::vector<double> c(static_cast<size_t>(nmode + 1) * (N + 1)), sn(c.size());
  for (int j = 0; j <= nmode; ++j)
    for (int i = 0; i <= N; ++i) {
      const double th = static_cast<double>(i) * dtheta;
      c[static_cast<size_t>(j) * (N + 1) + i] = std::cos(static_cast<double>(j) * th);
      sn[static_cast<size_t>(j) * (N + 1) + i] = std::sin(static_cast<double>(j) * th);
    }
  if (!a.d_cs) {
    const size_t bytes = sizeof(double) * (kSpecModes + 1) * (kSpecThetaMax + 1);
    AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_cs), bytes));
    AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_sn), bytes));
  }
  AH_HIP(hipMemcpyAsync(a.d_cs, c.data(), sizeof(double) * c.size(), hipMemcpyHostToDevice, st));
  AH_HIP(hipMemcpyAsync(a.d_sn, sn.data(), sizeof(double) * sn.size(), hipMemcpyHostToDevice, st));
  AH_HIP(hipStreamSynchronize(st));
  a.tab_nmode = nmode;
  a.tab_step = step;
  return OA_OK;
}

// The fast flow of every job that runs it, one workgroup each, in one
// launch: the starting surface (:1284-1407) here, the iterations on the
// device, then ahfound(k) and ack(k,:) (:1700-1733).
int run_flows(std::vector<Job*>& js, hipStream_t st) {
  if (js.empty()) return OA_OK;
  AhState& buf = *js[0]->h->ah;
  AH_RC(spec_capacity(buf, static_cast<int>(js.size())));
  const double half = 0.5, one = 1.0, two = 2.0, smallpi = M_PI;
  for (size_t q = 0; q < js.size(); ++q) {
    Job& j = *js[q];
    AhState& a = *j.h->ah;
    const oa_params& p = j.h->p;
    const double zcenter = j.s.zcenter;
    const int nmode = p.ahNmodes, N = 10 * nmode;
    const int step = (p.eqsym && zcenter == 0.0) ? 2 : 1;                  // :1250-1258
    const double dtheta = smallpi / static_cast<double>(step * N);
    AH_RC(spec_tables(a, nmode, step, st));
    if (!a.d_srad[j.k]) AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_srad[j.k]), sizeof(double) * (kSpecThetaMax + 1)));
    SpecIn& in = buf.h_sin[q];
    std::memset(&in, 0, sizeof in);
    in.ctx = a.dctx;
    in.cs = a.d_cs;
    in.sn = a.d_sn;
    in.rad = a.d_srad[j.k];
    in.zcenter = zcenter;
    in.dtheta = dtheta;
    in.AA = p.ahAA;
    in.BB = p.ahBB;
    in.eps = p.ahepsilon_spectral;
    in.nmode = nmode;
    in.ntheta = N;
    in.step = step;
    in.maxiter = p.ahmaxiter;
    if (a.found[j.k]) {                                                      // :1311-1330
      for (int m = 0; m <= kSpecModes; ++m) in.ac[m] = a.ack[j.k][m];
      in.reconstruct = 1;
    } else {                                                                 // :1340-1356
      const double ahrmax = (p.ahrmax > 0.9) ? 0.9 : p.ahrmax;
      if (zcenter == 0.0) in.ac[0] = (p.ahzinit == 0.0) ? 1.8 * ahrmax * a.hctx.L[0].zmax : 2.0 * p.ahzinit;
      else in.ac[0] = 1.95 * std::fabs(zcenter);
      in.reconstruct = 0;
      if (zcenter != 0.0 && p.aheccentric != 0.0) {                         // :1365-1407
        std::vector<double> rad(static_cast<size_t>(N) + 1);
        for (int i = 0; i <= N; ++i) {
          const double e = p.aheccentric * std::sin(static_cast<double>(i) * dtheta);
          rad[i] = half * in.ac[0] / std::sqrt(one - e * e);
        }
        for (int m = 0; m <= nmode; ++m) {
          double c = half * dtheta * rad[0];
          for (int i = 1; i < N; ++i) c = c + dtheta * rad[i] * std::cos(static_cast<double>(m) * (static_cast<double>(i) * dtheta));
          c = c + half * dtheta * rad[N] * std::cos(static_cast<double>(m) * (static_cast<double>(N) * dtheta));
          in.ac[m] = two * static_cast<double>(step) * c / smallpi;
        }
        in.reconstruct = 1;
      }
    }
  }
  const int n = static_cast<int>(js.size());
  size_t tab_bytes = 0;
  for (int q = 0; q < n; ++q) {
    const SpecIn& in = buf.h_sin[q];
    tab_bytes = std::max(tab_bytes, 2 * sizeof(double) * static_cast<size_t>(in.nmode + 1) * (in.ntheta + 1));
  }
  AH_HIP(hipMemcpyAsync(buf.d_sin, buf.h_sin, sizeof(SpecIn) * n, hipMemcpyHostToDevice, st));
  if (tab_bytes <= static_cast<size_t>(kSpecTabLds)) {
    // the tables and the kernel's own LDS together pass the 64 KiB default
    AH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_ah_spectral<true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, kSpecTabLds));
    hipLaunchKernelGGL(k_ah_spectral<true>, dim3(n), dim3(kSpecBlock), tab_bytes, st, buf.d_sin, buf.d_sout);
  } else {
    hipLaunchKernelGGL(k_ah_spectral<false>, dim3(n), dim3(kSpecBlock), 0, st, buf.d_sin, buf.d_sout);
  }
  AH_HIP(hipGetLastError());
  AH_HIP(hipMemcpyAsync(buf.h_sout, buf.d_sout, sizeof(SpecOut) * n, hipMemcpyDeviceToHost, st));
  AH_HIP(hipStreamSynchronize(st));
  for (size_t q = 0; q < js.size(); ++q) {
    Job& j = *js[q];
    AhState& a = *j.h->ah;
    const oa_params& p = j.h->p;
    SpecOut& o = buf.h_sout[q];
    if (o.ac[0] < p.dr0 + p.dz0) o.err = 1;                                 // :1700-1702
    const int found = (o.iter < p.ahmaxiter && !o.err) ? 1 : 0;
    a.found[j.k] = found;
    if (found)
      for (int m = 0; m <= kSpecModes; ++m) a.ack[j.k][m] = o.ac[m];
    a.first_sp[j.k] = 0;                                                     // :1748-1754
    a.shape_spec[j.k] = 1;
    a.spec_dtheta[j.k] = buf.h_sin[q].dtheta;
    a.n_of[j.k] = buf.h_sin[q].ntheta + 1;
    j.spec = o;
  }
  return OA_OK;
}

// Second part: the searches of `jobs` on stream st and their results.
int search_run(std::vector<Job>& jobs, hipStream_t st, oa_horizon* out) {
  // shooting searches that precede the flow (ahmethod shoot and both)
  AH_RC(run_walks(jobs, st));
  std::vector<Job*> proj, flows;
  for (Job& j : jobs)
    if (j.flow && j.shoot_first && j.h->ah->found[j.k] && j.shots > 0) proj.push_back(&j);
  AH_RC(project_shots(proj, st));
  // the flow's own start (:1286-1303): ack cleared on the first call unless a
  // horizon is known, then the shooting search of ahshootfirst
  bool again = false;
  for (Job& j : jobs) {
    if (!j.flow) continue;
    AhState& a = *j.h->ah;
    if (a.first_sp[j.k] && !a.found[j.k])
      for (int m = 0; m <= kSpecModes; ++m) a.ack[j.k][m] = 0.0;
    j.done = !j.shoot_in_flow;
    if (j.shoot_in_flow) {
      j.s.found_in = a.found[j.k];
      again = true;
    }
  }
  if (again) {
    AH_RC(run_walks(jobs, st));
    proj.clear();
    for (Job& j : jobs)
      if (j.shoot_in_flow && j.h->ah->found[j.k] && j.shots > 0) proj.push_back(&j);
    AH_RC(project_shots(proj, st));
    for (Job& j : jobs)
      if (j.shoot_in_flow) j.h->ah->ack[j.k][0] = 1.05 * j.h->ah->ack[j.k][0];
  }
  for (Job& j : jobs)
    if (j.flow) flows.push_back(&j);
  AH_RC(run_flows(flows, st));

  const double smallpi = M_PI, half = 0.5, two = 2.0;
  for (size_t q = 0; q < jobs.size(); ++q) {
    Job& j = jobs[q];
    const Search& s = j.s;
    AhState& a = *j.h->ah;
    oa_horizon& o = out[q];
    std::memset(&o, 0, sizeof o);
    o.method = j.h->p.ahmethod;
    o.found = a.found[j.k];
    o.zcenter = s.zcenter;
    o.shots = j.shots;
    o.zinit = j.zinit;
    o.dradlast = j.dradlast;
    double ah_area = 0.0;
    if (!j.flow) {
      o.ntheta = s.last.ntheta;
      o.errorflag = s.errorflag;
      if (o.found) ah_area = s.last.area;                                    // :989-997
    } else {
      o.ntheta = 10 * j.h->p.ahNmodes;
      o.errorflag = j.spec.err;
      o.iterations = j.spec.iter;
      o.residual = j.spec.res;
      if (o.found) ah_area = j.spec.area;                                    // :1826-1843
    }
    if (o.found && j.h->p.eqsym && s.zcenter == 0.0) ah_area = two * ah_area;
    o.area = ah_area;
    o.radius = half * std::sqrt(ah_area / smallpi);
    o.mass = o.found ? half * o.radius : 0.0;                                // :1065-1076, :1935-1945
  }
  return OA_OK;
}

int search_many(std::vector<Job>& jobs, hipStream_t st, oa_horizon* out) {
  AH_RC(search_prepare(jobs));
  return search_run(jobs, st, out);
}

// A begun search: its jobs run their walks on the search stream of the first
// job's spacetime, which waits for the point records.
int search_begin(std::vector<Job>& jobs, hipStream_t state_stream, oa_ahsearch** out) {
  AH_RC(search_prepare(jobs));
  AhState& a = *jobs[0].h->ah;
  if (!a.search_stream) AH_HIP(hipStreamCreateWithFlags(&a.search_stream, hipStreamNonBlocking));
  if (!a.fields_done) AH_HIP(hipEventCreateWithFlags(&a.fields_done, hipEventDisableTiming));
  AH_HIP(hipEventRecord(a.fields_done, state_stream));
  AH_HIP(hipStreamWaitEvent(a.search_stream, a.fields_done, 0));
  for (Job& j : jobs) j.h->ah->pending += 1;
  *out = reinterpret_cast<oa_ahsearch*>(new std::vector<Job>(std::move(jobs)));
  return OA_OK;
}

}  // namespace

extern "C" int oa_ahfind(oa_handle* h, int k, oa_horizon* out) {
  if (!h || !out) {
    oa_set_error("oa_ahfind: null argument");
    return OA_ERR_ARG;
  }
  std::vector<Job> jobs(1);
  jobs[0].h = h;
  jobs[0].k = k;
  return search_many(jobs, h->stream, out);
}

extern "C" int oa_batch_ahfind(oa_batch* b, int n, oa_handle* const* members, const int* k, oa_horizon* out) {
  if (!b || n < 0 || (n > 0 && (!members || !k || !out))) {
    oa_set_error("oa_batch_ahfind: bad argument");
    return OA_ERR_ARG;
  }
  if (n == 0) return OA_OK;
  std::vector<Job> jobs(static_cast<size_t>(n));
  for (int q = 0; q < n; ++q) {
    oa_handle* h = members[q];
    if (!h || h->batch != b) {
      oa_set_error("oa_batch_ahfind: every spacetime must be a member of the batch");
      return OA_ERR_ARG;
    }
    for (int r = 0; r < q; ++r)
      if (members[r] == h && k[r] == k[q]) {
        oa_set_error("oa_batch_ahfind: the same horizon of a spacetime twice");
        return OA_ERR_ARG;
      }
    jobs[static_cast<size_t>(q)].h = h;
    jobs[static_cast<size_t>(q)].k = k[q];
  }
  return search_many(jobs, members[0]->stream, out);
}

extern "C" int oa_ahfind_begin(oa_handle* h, int k, oa_ahsearch** search) {
  if (!h || !search) {
    oa_set_error("oa_ahfind_begin: null argument");
    return OA_ERR_ARG;
  }
  *search = nullptr;
  std::vector<Job> jobs(1);
  jobs[0].h = h;
  jobs[0].k = k;
  return search_begin(jobs, h->stream, search);
}

extern "C" int oa_batch_ahfind_begin(oa_batch* b, int n, oa_handle* const* members, const int* k,
                                     oa_ahsearch** search) {
  if (!b || n <= 0 || !members || !k || !search) {
    oa_set_error("oa_batch_ahfind_begin: bad argument");
    return OA_ERR_ARG;
  }
  *search = nullptr;
  std::vector<Job> jobs(static_cast<size_t>(n));
  for (int q = 0; q < n; ++q) {
    oa_handle* h = members[q];
    if (!h || h->batch != b) {
      oa_set_error("oa_batch_ahfind_begin: every spacetime must be a member of the batch");
      return OA_ERR_ARG;
    }
    for (int r = 0; r < q; ++r)
      if (members[r] == h) {
        oa_set_error("oa_batch_ahfind_begin: a spacetime twice");
        return OA_ERR_ARG;
      }
    jobs[static_cast<size_t>(q)].h = h;
    jobs[static_cast<size_t>(q)].k = k[q];
  }
  return search_begin(jobs, members[0]->stream, search);
}

extern "C" int oa_ahfind_end(oa_ahsearch* search, oa_horizon* out) {
  if (!search || !out) {
    oa_set_error("oa_ahfind_end: null argument");
    return OA_ERR_ARG;
  }
  std::vector<Job>* jobs = reinterpret_cast<std::vector<Job>*>(search);
  for (Job& j : *jobs) j.h->ah->pending -= 1;
  const int rc = search_run(*jobs, (*jobs)[0].h->ah->search_stream, out);
  delete jobs;
  return rc;
}

extern "C" int oa_horizon_shape(oa_handle* h, int k, double* theta, double* rad, int cap, int* n) {
  if (!h || !n || k < 0 || k > 2) {
    oa_set_error("oa_horizon_shape: bad argument");
    return OA_ERR_ARG;
  }
  if (!h->ah || h->ah->n_of[k] == 0) {
    oa_set_error("oa_horizon_shape: no search of this horizon yet (call oa_ahfind)");
    return OA_ERR_STATE;
  }
  AhState& a = *h->ah;
  const int m = a.n_of[k];
  *n = m;
  if (!theta || !rad || cap < m) {
    oa_set_error("oa_horizon_shape: arrays too small (need *n entries)");
    return OA_ERR_ARG;
  }
  if (a.shape_spec[k]) {                                                     // ah_radius_SP (:1766-1790)
    AH_HIP(hipMemcpyAsync(rad, a.d_srad[k], sizeof(double) * m, hipMemcpyDeviceToHost, h->stream));
    AH_HIP(hipStreamSynchronize(h->stream));
    for (int i = 0; i < m; ++i) {
      theta[i] = static_cast<double>(i) * a.spec_dtheta[k];
      if (!a.found[k]) rad[i] = 0.0;
    }
    return OA_OK;
  }
  if (a.gen_of[k] != a.gen) {
    oa_set_error("oa_horizon_shape: a later search rebuilt the horizon records; search this horizon again");
    return OA_ERR_STATE;
  }
  // the search's last integration again, recording its accepted points
  if (!a.d_theta[k]) {
    AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_theta[k]), sizeof(double) * (kAhNthetaMax + 1)));
    AH_HIP(hipMalloc(reinterpret_cast<void**>(&a.d_rad[k]), sizeof(double) * (kAhNthetaMax + 1)));
  }
  std::vector<ShotRes> res;
  AH_RC(run_shots(h->stream, a, std::vector<ShotIn>{ShotIn{a.dctx, a.last_z[k], a.last_zc[k], a.d_theta[k], a.d_rad[k]}},
                  res));
  if (res[0].ntheta + 1 != m) {
    oa_set_error("oa_horizon_shape: the repeated integration differs from the search's");
    return OA_ERR_STATE;
  }
  AH_HIP(hipMemcpyAsync(theta, a.d_theta[k], sizeof(double) * m, hipMemcpyDeviceToHost, h->stream));
  AH_HIP(hipMemcpyAsync(rad, a.d_rad[k], sizeof(double) * m, hipMemcpyDeviceToHost, h->stream));
  AH_HIP(hipStreamSynchronize(h->stream));
  if (!a.found[k])                                                           // :962-966
    for (int i = 0; i < m; ++i) rad[i] = 0.0;
  return OA_OK;
}
