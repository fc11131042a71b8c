// Per-point device code of the evolution hot path: finite-difference
// stencils with the reference's closures, the pointwise projection and the
// full right-hand side + RK update of one grid point.  Included by
// oaxis_kernels.hip, where OA_HD is __device__ __forceinline__.
#pragma once

#include <cmath>

#include "oa_internal.h"

#ifndef OA_HD
#define OA_HD inline
#endif
// functions the host side calls too
#ifndef OA_HHD
#define OA_HHD inline
#endif

namespace oax {

#define IDX(G, i, j) (static_cast<long>((i) + (G).g - 1) * (G).ld + ((j) + (G).g - 1))

constexpr double c12 = 1.0 / 12.0;
constexpr double third = 1.0 / 3.0;

using std::exp; using std::log; using std::sqrt; using std::cbrt; using std::fabs;

// ----------------------------------------------------------------- stencils
// The closures follow derivatives.f90:67-88 (diff1r), 179-212 (diff1z),
// 308-330 (diff2r), 425-459 (diff2z) and 606-691 (diff2rz) for the point
// classes that are computed (i >= 1; j >= 1 with equatorial symmetry).
// F is any callable returning the value at (ii, jj).  With IN = true the
// point is known to lie in i in [1, Nr-3], j in [1, Nz-3], where every
// stencil takes its centred (interior) form, and the closures compile out.

template <bool IN = false, class F>
OA_HD double s_d1r(const F& f, const KGrid& G, int i, int j) {
  if (IN || i <= G.Nr - 2)
    return G.idr * (8.0 * (f(i + 1, j) - f(i - 1, j)) - (f(i + 2, j) - f(i - 2, j))) * c12;
  if (i == G.Nr - 1)
    return G.idr * (3.0 * f(i + 1, j) + 10.0 * f(i, j) - 18.0 * f(i - 1, j) + 6.0 * f(i - 2, j)
                    - f(i - 3, j)) * c12;
  return G.idr * (25.0 * f(i, j) - 48.0 * f(i - 1, j) + 36.0 * f(i - 2, j) - 16.0 * f(i - 3, j)
                  + 3.0 * f(i - 4, j)) * c12;
}

template <bool IN = false, class F>
OA_HD double s_d1z(const F& f, const KGrid& G, int i, int j) {
  if (IN || (j >= 0 && j <= G.Nz - 2))
    return G.idz * (8.0 * (f(i, j + 1) - f(i, j - 1)) - (f(i, j + 2) - f(i, j - 2))) * c12;
  if (j == G.Nz - 1)
    return G.idz * (3.0 * f(i, j + 1) + 10.0 * f(i, j) - 18.0 * f(i, j - 1) + 6.0 * f(i, j - 2)
                    - f(i, j - 3)) * c12;
  if (j == G.Nz)
    return G.idz * (25.0 * f(i, j) - 48.0 * f(i, j - 1) + 36.0 * f(i, j - 2) - 16.0 * f(i, j - 3)
                    + 3.0 * f(i, j - 4)) * c12;
  if (j == -1)
    return -G.idz * (3.0 * f(i, j - 1) + 10.0 * f(i, j) - 18.0 * f(i, j + 1) + 6.0 * f(i, j + 2)
                     - f(i, j + 3)) * c12;
  return -G.idz * (25.0 * f(i, j) - 48.0 * f(i, j + 1) + 36.0 * f(i, j + 2) - 16.0 * f(i, j + 3)
                   + 3.0 * f(i, j + 4)) * c12;
}

template <bool IN = false, class F>
OA_HD double s_d2r(const F& f, const KGrid& G, int i, int j) {
  const double idr2 = G.idr * G.idr;
  if (IN || i <= G.Nr - 2)
    return -idr2 * (30.0 * f(i, j) - 16.0 * (f(i + 1, j) + f(i - 1, j)) + f(i + 2, j) + f(i - 2, j)) * c12;
  if (i == G.Nr - 1)
    return idr2 * (10.0 * f(i + 1, j) - 15.0 * f(i, j) - 4.0 * f(i - 1, j) + 14.0 * f(i - 2, j)
                   - 6.0 * f(i - 3, j) + f(i - 4, j)) * c12;
  return idr2 * (45.0 * f(i, j) - 154.0 * f(i - 1, j) + 214.0 * f(i - 2, j) - 156.0 * f(i - 3, j)
                 + 61.0 * f(i - 4, j) - 10.0 * f(i - 5, j)) * c12;
}

template <bool IN = false, class F>
OA_HD double s_d2z(const F& f, const KGrid& G, int i, int j) {
  const double idz2 = G.idz * G.idz;
  if (IN || (j >= 0 && j <= G.Nz - 2))
    return -idz2 * (30.0 * f(i, j) - 16.0 * (f(i, j + 1) + f(i, j - 1)) + f(i, j + 2) + f(i, j - 2)) * c12;
  if (j == G.Nz - 1)
    return idz2 * (10.0 * f(i, j + 1) - 15.0 * f(i, j) - 4.0 * f(i, j - 1) + 14.0 * f(i, j - 2)
                   - 6.0 * f(i, j - 3) + f(i, j - 4)) * c12;
  if (j == G.Nz)
    return idz2 * (45.0 * f(i, j) - 154.0 * f(i, j - 1) + 214.0 * f(i, j - 2) - 156.0 * f(i, j - 3)
                   + 61.0 * f(i, j - 4) - 10.0 * f(i, j - 5)) * c12;
  if (j == -1)
    return idz2 * (10.0 * f(i, j - 1) - 15.0 * f(i, j) - 4.0 * f(i, j + 1) + 14.0 * f(i, j + 2)
                   - 6.0 * f(i, j + 3) + f(i, j + 4)) * c12;
  return idz2 * (45.0 * f(i, j) - 154.0 * f(i, j + 1) + 214.0 * f(i, j + 2) - 156.0 * f(i, j + 3)
                 + 61.0 * f(i, j + 4) - 10.0 * f(i, j + 5)) * c12;
}

// mixed derivative: 4th order inside, the reference's 2nd-order edge rows /
// columns and corners, resolved per point
template <bool IN = false, class F>
OA_HD double s_d2rz(const F& f, const KGrid& G, int i, int j) {
  const double idrz = G.idr * G.idz;
  if constexpr (IN)
    return idrz * ((f(i + 1, j + 1) + f(i - 1, j - 1) - f(i + 1, j - 1) - f(i - 1, j + 1)) * third
                   - (f(i + 2, j + 2) + f(i - 2, j - 2) - f(i + 2, j - 2) - f(i - 2, j + 2)) * (1.0 / 48.0));
  const bool zedge = (j == -2 || j == G.Nz);
  const bool znear = (j == -1 || j == G.Nz - 1);
  if (i == G.Nr) {
    if (j == -2)
      return -0.25 * idrz * (9.0 * f(i, j) + 16.0 * f(i - 1, j + 1) + f(i - 2, j + 2)
                             - 12.0 * (f(i, j + 1) + f(i - 1, j)) + 3.0 * (f(i, j + 2) + f(i - 2, j))
                             - 4.0 * (f(i - 1, j + 2) + f(i - 2, j + 1)));
    if (j == G.Nz)
      return 0.25 * idrz * (9.0 * f(i, j) + 16.0 * f(i - 1, j - 1) + f(i - 2, j - 2)
                            - 12.0 * (f(i, j - 1) + f(i - 1, j)) + 3.0 * (f(i, j - 2) + f(i - 2, j))
                            - 4.0 * (f(i - 1, j - 2) + f(i - 2, j - 1)));
    return 0.25 * idrz * ((3.0 * f(i, j + 1) - 4.0 * f(i - 1, j + 1) + f(i - 2, j + 1))
                          - (3.0 * f(i, j - 1) - 4.0 * f(i - 1, j - 1) + f(i - 2, j - 1)));
  }
  if (zedge) {
    if (j == -2)
      return -0.25 * idrz * ((3.0 * f(i + 1, j) - 4.0 * f(i + 1, j + 1) + f(i + 1, j + 2))
                             - (3.0 * f(i - 1, j) - 4.0 * f(i - 1, j + 1) + f(i - 1, j + 2)));
    return 0.25 * idrz * ((3.0 * f(i + 1, j) - 4.0 * f(i + 1, j - 1) + f(i + 1, j - 2))
                          - (3.0 * f(i - 1, j) - 4.0 * f(i - 1, j - 1) + f(i - 1, j - 2)));
  }
  if (znear || i == G.Nr - 1)
    return 0.25 * idrz * (f(i + 1, j + 1) + f(i - 1, j - 1) - f(i + 1, j - 1) - f(i - 1, j + 1));
  return idrz * ((f(i + 1, j + 1) + f(i - 1, j - 1) - f(i + 1, j - 1) - f(i - 1, j + 1)) * third
                 - (f(i + 2, j + 2) + f(i - 2, j - 2) - f(i + 2, j - 2) - f(i - 2, j + 2)) * (1.0 / 48.0));
}

// upwinded advective derivatives, derivadvect.f90:105-175 / 304-374
template <bool IN = false, class F>
OA_HD double s_adr(const F& f, const KGrid& G, int i, int j, double beta) {
  const double q = 0.25 * G.idr * third;
  if constexpr (IN) {
    // both upwind sides in one expression: s = +1 reads i+1..i-3, s = -1 reads
    // i-1..i+3 with the overall sign flipped (s*q is exact)
    const int s = (beta <= 0.0) ? 1 : -1;
    return (s * q) * (3.0 * f(i + s, j) + 10.0 * f(i, j) - 18.0 * f(i - s, j) + 6.0 * f(i - 2 * s, j)
                      - f(i - 3 * s, j));
  }
  if (i <= G.Nr - 3) {
    if (beta <= 0.0)
      return q * (3.0 * f(i + 1, j) + 10.0 * f(i, j) - 18.0 * f(i - 1, j) + 6.0 * f(i - 2, j) - f(i - 3, j));
    return -q * (3.0 * f(i - 1, j) + 10.0 * f(i, j) - 18.0 * f(i + 1, j) + 6.0 * f(i + 2, j) - f(i + 3, j));
  }
  if (i == G.Nr - 2) {
    if (beta <= 0.0)
      return q * (3.0 * f(i + 1, j) + 10.0 * f(i, j) - 18.0 * f(i - 1, j) + 6.0 * f(i - 2, j) - f(i - 3, j));
    return q * (8.0 * (f(i + 1, j) - f(i - 1, j)) - (f(i + 2, j) - f(i - 2, j)));
  }
  if (i == G.Nr - 1)
    return q * (3.0 * f(i + 1, j) + 10.0 * f(i, j) - 18.0 * f(i - 1, j) + 6.0 * f(i - 2, j) - f(i - 3, j));
  return q * (25.0 * f(i, j) - 48.0 * f(i - 1, j) + 36.0 * f(i - 2, j) - 16.0 * f(i - 3, j) + 3.0 * f(i - 4, j));
}

template <bool IN = false, class F>
OA_HD double s_adz(const F& f, const KGrid& G, int i, int j, double beta) {
  const double q = 0.25 * G.idz * third;
  if constexpr (IN) {
    const int s = (beta <= 0.0) ? 1 : -1;
    return (s * q) * (3.0 * f(i, j + s) + 10.0 * f(i, j) - 18.0 * f(i, j - s) + 6.0 * f(i, j - 2 * s)
                      - f(i, j - 3 * s));
  }
  if (j >= 1 && j <= G.Nz - 3) {
    if (beta <= 0.0)
      return q * (3.0 * f(i, j + 1) + 10.0 * f(i, j) - 18.0 * f(i, j - 1) + 6.0 * f(i, j - 2) - f(i, j - 3));
    return -q * (3.0 * f(i, j - 1) + 10.0 * f(i, j) - 18.0 * f(i, j + 1) + 6.0 * f(i, j + 2) - f(i, j + 3));
  }
  if (j == 0) {   // the reference takes the right-biased form for beta_z <= 0 here
    if (beta <= 0.0)
      return -q * (3.0 * f(i, j - 1) + 10.0 * f(i, j) - 18.0 * f(i, j + 1) + 6.0 * f(i, j + 2) - f(i, j + 3));
    return q * (8.0 * (f(i, j + 1) - f(i, j - 1)) - (f(i, j + 2) - f(i, j - 2)));
  }
  if (j == -1)
    return -q * (3.0 * f(i, j - 1) + 10.0 * f(i, j) - 18.0 * f(i, j + 1) + 6.0 * f(i, j + 2) - f(i, j + 3));
  if (j == -2)
    return -q * (25.0 * f(i, j) - 48.0 * f(i, j + 1) + 36.0 * f(i, j + 2) - 16.0 * f(i, j + 3) + 3.0 * f(i, j + 4));
  if (j == G.Nz - 2) {
    if (beta <= 0.0)
      return q * (3.0 * f(i, j + 1) + 10.0 * f(i, j) - 18.0 * f(i, j - 1) + 6.0 * f(i, j - 2) - f(i, j - 3));
    return q * (8.0 * (f(i, j + 1) - f(i, j - 1)) - (f(i, j + 2) - f(i, j - 2)));
  }
  if (j == G.Nz - 1)
    return q * (3.0 * f(i, j + 1) + 10.0 * f(i, j) - 18.0 * f(i, j - 1) + 6.0 * f(i, j - 2) - f(i, j - 3));
  return q * (25.0 * f(i, j) - 48.0 * f(i, j - 1) + 36.0 * f(i, j - 2) - 16.0 * f(i, j - 3) + 3.0 * f(i, j - 4));
}

// Kreiss-Oliger D^6 terms of dissipation.f90:185-199: returns the amount to
// subtract from the source (already scaled by diss/dt)
template <bool IN = false, class F>
OA_HD double s_ko(const F& f, const KGrid& G, int i, int j, double dscale) {
  double s = 0.0;
  if (IN || i <= G.Nr - 3)
    s = dscale * (20.0 * f(i, j) - 15.0 * (f(i + 1, j) + f(i - 1, j)) + 6.0 * (f(i + 2, j) + f(i - 2, j))
                  - (f(i + 3, j) + f(i - 3, j))) * 0.05;
  double t = 0.0;
  if (IN || (j >= 1 && j <= G.Nz - 3))
    t = dscale * (20.0 * f(i, j) - 15.0 * (f(i, j + 1) + f(i, j - 1)) + 6.0 * (f(i, j + 2) + f(i, j - 2))
                  - (f(i, j + 3) + f(i, j - 3))) * 0.05;
  return s + t;
}

struct Plane {
  const double* __restrict__ p;
  KGrid G;
  OA_HD double operator()(int i, int j) const { return p[IDX(G, i, j)]; }
};

// Field views of the right-hand side: every stencil of rhs_point reads its
// fields through V.fld(slot)(i, j).  GlobalView reads the level slab in HBM;
// TileView reads a block's tile (with a 3-point halo) staged in LDS, laid out
// row-major as (i - ib) * tw + (j - jb), one tile of ts values per field.
// prefld(k) reads plane k (0: dtbeta_r, 1: dtbeta_z) of the dtbeta state that
// the first substep after the initial auxiliary pass differentiates
// (KStep::dtb_pre).
struct GlobalView {
  const double* __restrict__ y;
  KGrid G;
  const double* pre;
  OA_HD Plane fld(int s) const { return Plane{y + s * G.plane, G}; }
  OA_HD Plane prefld(int k) const { return Plane{pre + k * G.plane, G}; }
};

struct TilePlane {
  const double* p;      // the tile value of point (ib, jb)
  int ib, jb, tw;
  OA_HD double operator()(int i, int j) const { return p[(i - ib) * tw + (j - jb)]; }
};

struct TileView {
  const double* t;      // slot 0's value of point (ib, jb)
  int ib, jb, tw, ts;
  int npre;             // tile slot of prefld(0): staged after the evolved fields
  OA_HD TilePlane fld(int s) const { return TilePlane{t + s * ts, ib, jb, tw}; }
  OA_HD TilePlane prefld(int k) const { return TilePlane{t + (npre + k) * ts, ib, jb, tw}; }
};

struct D1 { double r, z; };
struct D2 { double r, z, rr, zz, rz; };

template <bool IN = false, class F>
OA_HD D1 grad(const F& f, const KGrid& G, int i, int j) {
  return {s_d1r<IN>(f, G, i, j), s_d1z<IN>(f, G, i, j)};
}
template <bool IN = false, class F>
OA_HD D2 hess(const F& f, const KGrid& G, int i, int j) {
  return {s_d1r<IN>(f, G, i, j), s_d1z<IN>(f, G, i, j), s_d2r<IN>(f, G, i, j), s_d2z<IN>(f, G, i, j),
          s_d2rz<IN>(f, G, i, j)};
}

// ------------------------------------------------------------- projection
// Pointwise part of auxiliary_geometry.f90 applied to freshly updated values:
// factordet (:299-316, factordet.f90:40-64) after the first step, the
// phi <-> chi relation (:178-203) and, with matter, stressenergy.f90:65-67.
struct Proj {
  double chi, phi, A, B, H, C;
};

OA_HD void project(Proj& q, double r, bool time0, int chimethod, bool matter) {
  if (!chimethod || time0) {
    q.chi = exp(-2.0 * q.phi);
    if (chimethod && matter) q.phi = -0.5 * log(q.chi);
  } else {
    q.phi = -0.5 * log(q.chi);
  }
  if (!time0) {
    const double rc = r * q.C;
    double aux = (q.A * q.B - rc * rc) * q.H;
    aux = cbrt(aux);
    const double ia = 1.0 / aux;
    q.A *= ia; q.B *= ia; q.C *= ia; q.H *= ia;
  }
}

// write v at (i,j) and, inside the symmetry bands, its parity images
OA_HD void put(double* __restrict__ plane, const KGrid& G, int i, int j, double v,
                                    double symr, double symz) {
  plane[IDX(G, i, j)] = v;
  if (i <= G.g) plane[IDX(G, 1 - i, j)] = symr * v;
  if (G.eqsym && j >= 1 && j <= G.g) {
    plane[IDX(G, i, 1 - j)] = symz * v;
    if (i <= G.g) plane[IDX(G, 1 - i, 1 - j)] = symr * symz * v;
  }
}


// axis / equator parity of every device slot (arrays.config SYMMETRYR/Z)
OA_HD void slot_parity(bool shift, bool scalar, int s, double& symr, double& symz) {
  symr = 1.0;
  symz = 1.0;
  if (s == S_C || s == S_KTC || s == S_Delta_z) symz = -1.0;
  if (s == S_Delta_r) symr = -1.0;
  if (shift && (s == shift_slot(SH_beta_r) || s == shift_slot(SH_dtbeta_r))) symr = -1.0;
  if (shift && (s == shift_slot(SH_beta_z) || s == shift_slot(SH_dtbeta_z))) symz = -1.0;
  if (scalar && s == scalar_slot(shift, SC_xi_r)) symr = -1.0;
  if (scalar && s == scalar_slot(shift, SC_xi_z)) symz = -1.0;
}

// ------------------------------------------------------------- RHS kernel
// PHIMODE: the auxiliary arrays the sources read were last computed while the
// level clock was 0, so the conformal-factor derivatives are those of phi
// (auxiliary_geometry.f90:178-199) and the chi advection derivatives were
// never formed (they keep their initial zero).  This holds for the first step
// of a level and for the first internal iteration of its second step.
// S.proj_t0: the auxiliary call that ends this iteration runs at clock 0.
template <bool SHIFT, bool SCALAR, bool PHIMODE, bool IN, int STAGE, class View>
OA_HD void rhs_point(const KGrid& G, const KPhys& P, const KStep& S, const View& V,
                     const double* __restrict__ yp, double* __restrict__ acc, double* __restrict__ yout,
                     double* __restrict__ gaux, int i, int j) {

  const long pl = G.plane;
  auto fld = [&](int s) { return V.fld(s); };
  const long k = IDX(G, i, j);

  // coordinates (evolve.f90:52-71)
  const double r = (static_cast<double>(i) - 0.5) * G.dr;
  const double z = (static_cast<double>(j) - G.zoff) * G.dz;
  const double ir = 1.0 / r;
  const double r2 = r * r;

  const auto Fa = fld(S_alpha), Fchi = fld(S_chi), Fphi = fld(S_phi);
  const auto FA = fld(S_A), FB = fld(S_B), FH = fld(S_H), FC = fld(S_C);
  const auto FK = fld(S_trK), FKA = fld(S_KTA), FKB = fld(S_KTB), FKC = fld(S_KTC);
  const auto FDr = fld(S_Delta_r), FDz = fld(S_Delta_z);

  const double alpha = Fa(i, j), chi = Fchi(i, j);
  const double A = FA(i, j), B = FB(i, j), H = FH(i, j), C = FC(i, j);
  const double trK = FK(i, j), KTA = FKA(i, j), KTB = FKB(i, j), KTC = FKC(i, j);
  const double Delta_r = FDr(i, j), Delta_z = FDz(i, j);

  // shift values
  double br = 0.0, bz = 0.0, dbr = 0.0, dbz = 0.0;
  if constexpr (SHIFT) {
    br = fld(shift_slot(SH_beta_r))(i, j);
    bz = fld(shift_slot(SH_beta_z))(i, j);
    dbr = fld(shift_slot(SH_dtbeta_r))(i, j);
    dbz = fld(shift_slot(SH_dtbeta_z))(i, j);
  }
  // advective derivative: upwinded (shiftadvect) or the centred one
  auto adv = [&](const auto& f, const D1& d, double& DAr, double& DAz) {
    if (P.shiftadvect) {
      DAr = s_adr<IN>(f, G, i, j, br);
      DAz = s_adz<IN>(f, G, i, j, bz);
    } else {
      DAr = d.r;
      DAz = d.z;
    }
  };

  // ---- lapse (auxiliary_geometry.f90:55-72)
  const D2 da = hess<IN>(Fa, G, i, j);
  double DAr_alpha = 0.0, DAz_alpha = 0.0;
  if constexpr (SHIFT) adv(Fa, D1{da.r, da.z}, DAr_alpha, DAz_alpha);

  // ---- conformal factor (:178-236); psi4 = chi^-2
  const double ipsi4 = chi * chi;
  D2 dp;
  double DAr_phi = 0.0, DAz_phi = 0.0, DAr_chi = 0.0, DAz_chi = 0.0;
  if constexpr (PHIMODE) {
    dp = hess<IN>(Fphi, G, i, j);
    if constexpr (SHIFT) adv(Fphi, D1{dp.r, dp.z}, DAr_phi, DAz_phi);
  } else {
    const D2 dc = hess<IN>(Fchi, G, i, j);
    const double ic = 1.0 / chi;
    const double qr = dc.r * ic, qz = dc.z * ic;
    dp.r = -0.5 * qr;
    dp.z = -0.5 * qz;
    dp.rr = -0.5 * (dc.rr * ic - qr * qr);
    dp.zz = -0.5 * (dc.zz * ic - qz * qz);
    dp.rz = -0.5 * (dc.rz * ic - qr * qz);
    if constexpr (SHIFT) {
      adv(Fchi, D1{dc.r, dc.z}, DAr_chi, DAz_chi);
      DAr_phi = -0.5 * DAr_chi * ic;
      DAz_phi = -0.5 * DAz_chi * ic;
    }
  }

  // ---- conformal metric and its inverse (:325-507); hdet = 1
  const D2 dA = hess<IN>(FA, G, i, j), dB = hess<IN>(FB, G, i, j), dH = hess<IN>(FH, G, i, j), dC = hess<IN>(FC, G, i, j);
  const double gA = B * H, gB = A * H, gH = 1.0 / H, gC = -C * H;
  auto fgA = [&](int ii, int jj) { return FB(ii, jj) * FH(ii, jj); };
  auto fgB = [&](int ii, int jj) { return FA(ii, jj) * FH(ii, jj); };
  auto fgH = [&](int ii, int jj) { return 1.0 / FH(ii, jj); };
  auto fgC = [&](int ii, int jj) { return -FC(ii, jj) * FH(ii, jj); };
  const D1 dgA = grad<IN>(fgA, G, i, j), dgB = grad<IN>(fgB, G, i, j), dgC = grad<IN>(fgC, G, i, j);
  (void)fgH;

  // ---- KT_H, regularisation (:747-821)
  const double KTH = -(gA * KTA + gB * KTB + 2.0 * r2 * gC * KTC) * H;
  const double lambda = (A - H) * ir * ir;
  const double Alambda = (KTA - KTH) * ir * ir;
  const double g_lambda = C * C - B * lambda;

  // ---- extrinsic curvature derivatives (:830-906)
  const D1 dK = grad<IN>(FK, G, i, j), dKA = grad<IN>(FKA, G, i, j), dKB = grad<IN>(FKB, G, i, j), dKC = grad<IN>(FKC, G, i, j);

  // ---- contravariant KT, KT2, K2 (:970-1050)
  const double gA2 = gA * gA, gB2 = gB * gB, gH2 = gH * gH, gC2 = gC * gC;
  const double r3 = r2 * r, r4 = r3 * r, r5 = r4 * r;
  const double KTup_rr = (2.0 * gA * gC * KTC + gC2 * KTB) * r2 + gA2 * KTA;
  const double KTup_zz = (gC2 * KTA + 2.0 * gC * gB * KTC) * r2 + gB2 * KTB;
  const double KTup_rz = gC2 * r3 * KTC + (gC * gA * KTA + gB * gA * KTC + gB * gC * KTB) * r;
  const double KT2_A = (2.0 * KTA * gC * KTC + KTC * KTC * gB) * r2 + gA * KTA * KTA;
  const double KT2_B = (KTC * KTC * gA + 2.0 * KTC * gC * KTB) * r2 + gB * KTB * KTB;
  const double KT2_H = KTH * KTH * gH;
  const double KT2_C = KTC * KTC * gC * r2 + KTA * gA * KTC + KTA * gC * KTB + KTC * gB * KTB;
  const double KT2 = gA * KT2_A + gB * KT2_B + gH * KT2_H + 2.0 * gC * KT2_C * r2;
  const double K2 = KT2 + third * trK * trK;

  // ---- Delta (:1085-1175)
  const double DelDef_r = -dgA.r - r * dgC.z - r * g_lambda;
  const double DelDef_z = -dgC.r * r - dgB.z - 2.0 * gC;
  const D1 dDr = grad<IN>(FDr, G, i, j), dDz = grad<IN>(FDz, G, i, j);

  // ---- shift derivatives (:587-737)
  D2 dbr_{}, dbz_{};
  double DD_beta_rr = 0.0, DIV_beta = 0.0, Dr_DIV_beta = 0.0, Dz_DIV_beta = 0.0;
  if constexpr (SHIFT) {
    const auto Fbr = fld(shift_slot(SH_beta_r)), Fbz = fld(shift_slot(SH_beta_z));
    dbr_ = hess<IN>(Fbr, G, i, j);
    dbz_ = hess<IN>(Fbz, G, i, j);
    auto fbr_r = [&](int ii, int jj) { return Fbr(ii, jj) / ((static_cast<double>(ii) - 0.5) * G.dr); };
    DD_beta_rr = s_d1r<IN>(fbr_r, G, i, j);
    DIV_beta = dbr_.r + dbz_.z + br * ir;
    Dr_DIV_beta = dbr_.rr + dbz_.rz + DD_beta_rr;
    Dz_DIV_beta = dbr_.rz + dbz_.zz + dbr_.z * ir;
  }

  const double Dra = da.r, Dza = da.z, Drp = dp.r, Dzp = dp.z;

  // ---- geometric auxiliaries: D_i D_j alpha, the Ricci tensor, div KT.
  // STAGE 0 computes them in place, STAGE 1 computes the Ricci tensor and
  // div KT and stores them (one plane each in gaux, AUX_* order) and returns,
  // STAGE 2 reads them back.  D_i D_j alpha and its trace are formed where
  // they are used (STAGE 0 and 2) from the lapse and metric derivatives that
  // phase holds anyway, rather than passed through gaux.
  double d2A, d2B, d2C, Lapla, RA, RB, RC, RSCAL, DIV_r, DIV_z, DIVA_r, DIVA_z;
  auto lapse_hessian = [&]() {
    // D_i D_j alpha and its trace (:1268-1324)
    d2A = da.rr - Dra * gC * r2 * dC.r
                 - (Dza * gB * dC.r + 0.5 * Dza * gC * dA.r + Dra * gC * C - 0.5 * Dra * gC * dA.z) * r
                 - 0.5 * Dra * gA * dA.r - Dza * gB * C + 0.5 * Dza * gB * dA.z;
    d2B = da.zz - Dza * gC * r2 * dC.z
                 - (Dra * gA * dC.z + 0.5 * Dra * gC * dB.z - 0.5 * Dza * gC * dB.r) * r
                 + 0.5 * Dra * gA * dB.r - 0.5 * Dza * gB * dB.z;
    double d2H = (0.5 * Dra * gC * dH.z + 0.5 * Dza * gC * dH.r) * r
                 + 0.5 * Dra * gA * dH.r + Dza * gC * H + 0.5 * Dza * gB * dH.z + Dra * gA * H * ir;
    d2C = da.rz * ir - 0.5 * Dra * gC * dB.r - 0.5 * Dza * gC * dA.z
                 - (0.5 * Dra * gA * dA.z + 0.5 * Dza * gB * dB.r) * ir;
    {
      const double aux = gA * Dra * Drp + gB * Dza * Dzp + r * gC * (Dra * Dzp + Dza * Drp);
      d2A -= 2.0 * (2.0 * Dra * Drp - A * aux);
      d2B -= 2.0 * (2.0 * Dza * Dzp - B * aux);
      d2H += 2.0 * H * aux;
      d2C -= 2.0 * (Dra * ir * Dzp + Dza * Drp * ir - C * aux);
    }
    Lapla = (gA * d2A + gB * d2B + gH * d2H + 2.0 * r2 * gC * d2C) * ipsi4;
  };
  if constexpr (STAGE == 2) {
    const double* const ax = gaux + k;
    RA = ax[AUX_RA * pl]; RB = ax[AUX_RB * pl]; RC = ax[AUX_RC * pl]; RSCAL = ax[AUX_RSCAL * pl];
    DIV_r = ax[AUX_DIV_r * pl]; DIV_z = ax[AUX_DIV_z * pl]; DIVA_r = ax[AUX_DIVA_r * pl]; DIVA_z = ax[AUX_DIVA_z * pl];
    lapse_hessian();
  } else {
    if constexpr (STAGE != 12) {
      if constexpr (STAGE == 0) lapse_hessian();

      // ---- Ricci tensor (:1397-1493 and calc_conformalRicci1..4, non-angmom)
      double RH;
      {
        RA = -2.0 * (dp.rr - Drp * gC * r2 * dC.r
                     - (Dzp * gB * dC.r + 0.5 * Dzp * gC * dA.r + Drp * gC * C - 0.5 * Drp * gC * dA.z) * r
                     - 0.5 * Drp * gA * dA.r - Dzp * gB * C + 0.5 * Dzp * gB * dA.z);
        RB = -2.0 * (dp.zz - Dzp * gC * r2 * dC.z
                     - (Drp * gA * dC.z + 0.5 * Drp * gC * dB.z - 0.5 * Dzp * gC * dB.r) * r
                     + 0.5 * Drp * gA * dB.r - 0.5 * Dzp * gB * dB.z);
        RH = -2.0 * ((0.5 * Drp * gC * dH.z + 0.5 * Dzp * gC * dH.r) * r
                     + 0.5 * Drp * gA * dH.r + Dzp * gC * H + 0.5 * Dzp * gB * dH.z + Drp * gA * H * ir);
        RC = -2.0 * (dp.rz * ir - 0.5 * Drp * gC * dB.r - 0.5 * Dzp * gC * dA.z
                     - (0.5 * Drp * gA * dA.z + 0.5 * Dzp * gB * dB.r) * ir);
        double aux = gA * RA + gB * RB + gH * RH + 2.0 * r2 * gC * RC;
        RA += A * aux; RB += B * aux; RH += H * aux; RC += C * aux;
        RA += 4.0 * Drp * Drp;
        RB += 4.0 * Dzp * Dzp;
        RC += 4.0 * Drp * Dzp * ir;
        aux = -2.0 * (gA * Drp * Drp + gB * Dzp * Dzp + 2.0 * r * gC * Drp * Dzp);
        RA += 2.0 * A * aux; RB += 2.0 * B * aux; RH += 2.0 * H * aux; RC += 2.0 * C * aux;
      }
      {
        // term 1: second derivatives of the conformal metric
        const double lamDef = lambda;
        RA += -gC * dA.rz * r - 0.5 * gB * dA.zz - 0.5 * gA * dA.rr - 0.5 * gH * dA.r * ir + gH * lamDef;
        RB += -gC * dB.rz * r - 0.5 * gA * dB.rr - 0.5 * gB * dB.zz - 0.5 * gH * dB.r * ir;
        RH += -gC * dH.rz * r - 0.5 * gB * dH.zz - 0.5 * gA * dH.rr - 0.5 * gH * dH.r * ir - gH * lamDef;
        RC += -gC * dC.rz * r - 0.5 * gA * dC.rr - gC * dC.z - 0.5 * gB * dC.zz
              + (-gA * dC.r - 0.5 * gH * dC.r) * ir;
        // term 2: Delta terms
        RA += C * dDz.r * r + A * dDr.r + 0.5 * (DelDef_r * dA.r + DelDef_z * dA.z);
        RB += C * dDr.z * r + B * dDz.z + 0.5 * (DelDef_r * dB.r + DelDef_z * dB.z);
        RH += 0.5 * (DelDef_r * dH.r + DelDef_z * dH.z) + Delta_r * H * ir;
        RC += 0.5 * C * (dDr.r + dDz.z) + 0.5 * (DelDef_r * dC.r + DelDef_z * dC.z)
              + 0.5 * (C * Delta_r + A * dDr.z + B * dDz.r) * ir;
      }
      {
        // terms 3 and 4: quadratic in first derivatives of the metric
        const double Ar = dA.r, Az = dA.z, Br = dB.r, Bz = dB.z, Hr = dH.r, Hz = dH.z, Cr = dC.r, Cz = dC.z;
        RA += (2.0 * gB * gC * Cr * Cz + gC2 * Ar * Cz + Az * gC2 * Cr) * r3
              + (2.0 * gB * gC * C * Cz - gB * gC * Cr * Br + Ar * gC * gA * Cr + Az * gC2 * C) * r2
              + (Az * gB * gC * Br + Az * gB * gA * Cr + Ar * gC * gA * C - gB * gC * C * Br + Az * gC * gA * Ar
                 + gB2 * Br * Cz) * r
              - 0.5 * gH2 * Hr * Hr - 0.5 * gB2 * Br * Br + gB * gA * C * Az + 0.5 * Ar * Ar * gA2
              + (gH * Hr + gH2 * A * Hr - 2.0 * gH2 * Hr * H) * ir;
        RB += (Br * gC2 * Cz + gC2 * Bz * Cr + 2.0 * gA * gC * Cz * Cr) * r3
              + (2.0 * gA * gC * Cz * C - gA * gC * Cz * Az + gC2 * Bz * C + Bz * gC * gB * Cz) * r2
              + (Br * gA * gC * Az + gA2 * Az * Cr + Br * gA * gB * Cz + Br * gB * gC * Bz) * r
              + 0.5 * gB2 * Bz * Bz + gA2 * Az * C - 0.5 * gA2 * Az * Az + gH2 * Hz * C - 0.5 * gH2 * Hz * Hz;
        RH += -gC * Hz - r * g_lambda * Hr;
        RC += (gC2 * Br * Cr + gC * gB * Cz * Cz + Az * gC2 * Cz + gC * gA * Cr * Cr) * r2
              + (-0.5 * gC * gA * Cr * Az + 2.0 * gC * gA * Cr * C - 0.5 * gC2 * Br * Az - 0.5 * gC * gB * Cz * Br
                 + 0.5 * Bz * gB * gC * Cr + gC2 * Br * C + 0.5 * Bz * gC2 * Ar + 0.5 * Ar * gA * gC * Cz) * r
              + 0.5 * Br * gC * gA * Ar + 0.5 * Br * gB * gA * Cr + 0.5 * gB2 * Bz * Cz + 0.5 * Az * gA * gB * Cz
              + 0.5 * gA2 * Ar * Cr + 0.5 * Br * Br * gB * gC + gC * gA * C * C - 0.5 * gC * gA * C * Az
              + 0.5 * Az * gB * gC * Bz + 0.5 * Az * Az * gA * gC + 0.5 * Bz * gB * gC * C
              + (0.5 * gH2 * C * Hr + 0.5 * gB * gA * Br * C - 0.5 * gH2 * Hz * Hr + 0.5 * gA2 * Ar * C) * ir
              + 0.5 * Hz * gH2 * lambda;
        RA += Az * gC2 * Cr * r3
              + (-0.5 * Az * Az * gC2 + Az * gC2 * C + gB * gA * Cr * Cr + 0.5 * Ar * gC2 * Br + Ar * gC * gA * Cr
                 + gB * gC * Cr * Br) * r2
              + (-Az * gB * gA * Cr + 2.0 * gB * gA * Cr * C + gB * gC * C * Br + Ar * gC * gA * C) * r
              + 0.25 * gH2 * Hr * Hr - gB * gA * C * Az + 0.25 * gB2 * Br * Br + gB * gA * C * C
              + 0.25 * Ar * Ar * gA2 + 0.5 * gB * gA * Az * Az;
        RB += Br * gC2 * r3 * Cz
              + (gA * gB * Cz * Cz + 0.5 * Az * gC2 * Bz - 0.5 * gC2 * Br * Br + Bz * gC * gB * Cz + gA * gC * Cz * Az) * r2
              - Br * gA * gB * Cz * r + 0.25 * gA2 * Az * Az + 0.5 * gB * gA * Br * Br + 0.25 * gH2 * Hz * Hz
              + 0.25 * gB2 * Bz * Bz;
        RH += (-0.5 * C * gC * gH * Hr + gH * gC * Hz * Hr) * r + 0.5 * gH * Hz * Hz * gB + 1.5 * gC * Hz
              + 0.5 * gH * gA * Hr * Hr - 0.5 * C * gB * gH * Hz - 0.5 * gH * A * gC * Hz - C * gC - 0.5 * gC * Hz
              + 0.5 * (2.0 * gA - gH - gH * A * gA) * Hr * ir + gH * lambda * (1.0 - gA * H);
        RC += gC2 * Cr * Cz * r3 + (-0.5 * Az * gC2 * Cz + gC2 * C * Cz - 0.5 * gC2 * Br * Cr) * r2
              + (-0.5 * gC2 * Br * C + 0.25 * Bz * gC2 * Ar + 0.5 * Ar * gA * gC * Cz + 0.5 * gC * gB * Cz * Br
                 + 0.75 * gC2 * Br * Az + 0.5 * Bz * gB * gC * Cr + 0.5 * gC * gA * Cr * Az) * r
              + 0.5 * Az * gA * gB * Cz + 0.5 * Bz * gB * gC * C + 0.5 * gC * gA * C * Az + 0.5 * Br * gB * gA * Cr
              + (0.25 * gH2 * Hz * Hr + 0.5 * gB * gA * Br * C + 0.25 * gA2 * Ar * Az + 0.25 * Bz * gB2 * Br
                 - 0.5 * gB * gA * Az * Br) * ir;
      }
      RSCAL = (gA * RA + gB * RB + gH * RH + 2.0 * r2 * gC * RC) * ipsi4;

    }
    if constexpr (STAGE != 11) {
      // ---- divergence of KT (calc_DIVKTA.f90:573-636)
      {
        const double Ar = dA.r, Az = dA.z, Br = dB.r, Bz = dB.z, Hr = dH.r, Hz = dH.z, Cr = dC.r, Cz = dC.z;
        const double KAr = dKA.r, KAz = dKA.z, KBr = dKB.r, KBz = dKB.z, KCr = dKC.r, KCz = dKC.z;
        const double gA3 = gA2 * gA, gB3 = gB2 * gB, gC3 = gC2 * gC;
        DIV_r = -gC3 * Cz * KTC * r5 - gA * gC2 * KTC * Cr * r4
          + (gC2 * KCz - 0.5 * gC3 * Bz * KTA - gA * gC2 * KTC * C - 0.5 * gC3 * Az * KTB - gA * Cz * gC2 * KTA
             - gA * Cz * gC * gB * KTC - gC2 * Cz * gB * KTB - gC2 * Bz * gB * KTC - gA * Az * gC2 * KTC) * r3
          + (2.0 * gA * gC * KCr - Delta_r * gC * KTC - gC * Br * gB * gA * KTC - gB * gA2 * KTC * Cr
             - Ar * gA2 * gC * KTC - gC * gA2 * KTA * Cr + gC2 * KBr - gA * gB * gC * KTB * Cr
             - 0.5 * gA * Ar * gC2 * KTB - 0.5 * gC2 * Br * gA * KTA) * r2
          + (2.0 * gA * gC * KTC + gA * gC * KAz - gC * gA2 * KTA * C + gC * gH * KTC - 0.5 * gH2 * KTH * gC * Hz
             - Delta_z * gC * KTB - gB * gA2 * KTC * C - 0.5 * gA2 * Az * gC * KTA + gC * gB * KBz + gA * gB * KCz
             - Delta_z * gA * KTC - gA * gB * gC * KTB * C - gC * gH2 * KTH * C - 0.5 * gC * Bz * gB2 * KTB) * r
          - 0.5 * gA * Br * gB2 * KTB - 0.5 * gH2 * KTH * gA * Hr - Delta_r * gA * KTA + gA2 * KAr
          - 0.5 * gA3 * Ar * KTA + r * gH * (gA * Alambda + gH * KTH * C * gC);
        DIV_z = -r5 * gC3 * KTC * Cr + (-gC3 * KTC * C - Cz * gC2 * gB * KTC) * r4
          + (gC2 * KCr - gB * gC2 * KTB * Cr - gC * gB * gA * KTC * Cr - 0.5 * gC3 * Br * KTA - gC2 * gA * KTA * Cr
             - 0.5 * Ar * gC3 * KTB - Ar * gA * gC2 * KTC - gB * Br * gC2 * KTC) * r3
          + (-gB * Cz * gC * gA * KTA + gC2 * KTC - Bz * gC * gB2 * KTC - Delta_z * gC * KTC - gC * Cz * gB2 * KTB
             + gC2 * KAz + 2.0 * gC * gB * KCz - gC2 * gA * KTA * C - Cz * gB2 * gA * KTC - 0.5 * Az * gB * gC2 * KTB
             - gC * Az * gB * gA * KTC - gB * gC2 * KTB * C - 0.5 * gB * Bz * gC2 * KTA - gC * gB * gA * KTC * C) * r2
          + (gB * gC * KBr + gB * gA * KCr + gC * gA * KAr - 0.5 * gH2 * KTH * gC * Hr - Delta_r * gB * KTC
             - Delta_r * gC * KTA - 0.5 * gB2 * Br * gC * KTB - 0.5 * gC * Ar * gA2 * KTA) * r
          + gC * gH * KTA + gB * gA * KTC - 0.5 * Bz * gB3 * KTB + gB2 * KBz - Delta_z * gB * KTB
          - gC * gH2 * KTH * A - 0.5 * gB * Az * gA2 * KTA - 0.5 * gH2 * KTH * gB * Hz - gC * gH * KTH
          + gB * gH * KTC - gB * gH2 * KTH * C;
        DIVA_r = -r5 * gC3 * Cz * KTC + (-gC3 * Cr * KTB - gC3 * Br * KTC - 3.0 * gC2 * Cr * gA * KTC) * r4
          + (-gC3 * C * KTB - 2.0 * gA * Cz * gC2 * KTA - 3.0 * gC2 * C * gA * KTC + gC2 * KCz - gC2 * Az * gA * KTC
             - 3.0 * gA * Cz * gC * gB * KTC - 2.0 * gC2 * Bz * gB * KTC - gC3 * Bz * KTA - gC2 * Cz * gB * KTB) * r3
          + (-gA * Br * gC2 * KTA - gA * gB * gC * KTB * Cr + 2.0 * gA * gC * KCr - 2.0 * gC * Cr * gA2 * KTA
             - Delta_r * gC * KTC + gC2 * KBr - gC2 * Br * gB * KTB - 2.0 * gA2 * Ar * gC * KTC - gA * Ar * gC2 * KTB
             - gB * gA2 * KTC * Cr - gA * Br * gC * gB * KTC) * r2
          + (-gA * Cz * gB2 * KTB - gB * gA2 * KTC * C - gC * Bz * gB2 * KTB + 2.0 * gA * gC * KTC + gC * gB * KBz
             + gC * gH * KTC + gA * gC * KAz + gA * gB * KCz - gA * Az * gB * gC * KTB - gC * gH2 * KTH * C
             - Delta_z * gC * KTB - gA2 * Az * gB * KTC - 2.0 * gC * C * gA2 * KTA - Delta_z * gA * KTC
             - gA2 * Az * gC * KTA - gA * gB * gC * KTB * C) * r
          - Ar * gA3 * KTA - Delta_r * gA * KTA + gA2 * KAr + r * gA * gH * (Alambda - lambda * gH * KTH);
        DIVA_z = -r5 * gC3 * KTC * Cr + (-3.0 * Cz * gC2 * gB * KTC - Az * gC3 * KTC - Cz * gC3 * KTA - gC3 * KTC * C) * r4
          + (-3.0 * gC * gB * gA * KTC * Cr - 2.0 * Ar * gA * gC2 * KTC + gC2 * KCr - gB * Br * gC2 * KTC
             - Ar * gC3 * KTB - 2.0 * gB * gC2 * KTB * Cr - gC2 * gA * KTA * Cr) * r3
          + (-gC * Az * gB * gA * KTC - gC2 * gA * KTA * C - 2.0 * Bz * gC * gB2 * KTC - Az * gC2 * gA * KTA
             - Az * gB * gC2 * KTB - 3.0 * gC * gB * gA * KTC * C - 2.0 * gC * Cz * gB2 * KTB - Delta_z * gC * KTC
             - Cz * gB2 * gA * KTC + gC2 * KTC + gC2 * KAz + 2.0 * gC * gB * KCz - gB * Cz * gC * gA * KTA
             - gB * Bz * gC2 * KTA - 2.0 * gB * gC2 * KTB * C) * r2
          + (-Delta_r * gB * KTC + gB * gA * KCr + gC * gA * KAr + gB * gC * KBr - gB2 * Br * gC * KTB
             - Delta_r * gC * KTA - gB2 * Br * gA * KTC - gC * Ar * gA2 * KTA - gB * Br * gC * gA * KTA
             - gB * gA2 * KTA * Cr) * r
          - Delta_z * gB * KTB - gB * gH2 * KTH * C + gB2 * KBz - gC * gH2 * KTH * A - gB * gA2 * KTA * C
          + gB * gH * KTC + gC * gH * KTA + gB * gA * KTC - Bz * gB3 * KTB;
      }
    }
  }
  // STAGE 11 / 12: the two halves of STAGE 1 (the Ricci tensor; div KT),
  // for kernels that run them one after the other
  if constexpr (STAGE == 1 || STAGE == 11 || STAGE == 12) {
    double* const ax = gaux + k;
    if constexpr (STAGE != 12) {
      ax[AUX_RA * pl] = RA; ax[AUX_RB * pl] = RB; ax[AUX_RC * pl] = RC; ax[AUX_RSCAL * pl] = RSCAL;
    }
    if constexpr (STAGE != 11) {
      ax[AUX_DIV_r * pl] = DIV_r; ax[AUX_DIV_z * pl] = DIV_z; ax[AUX_DIVA_r * pl] = DIVA_r; ax[AUX_DIVA_z * pl] = DIVA_z;
    }
    return;
  }


  // ---- real scalar field: auxiliary_matter / potential / stressenergy
  double rho = 0.0, J_r = 0.0, J_z = 0.0, mS_A = 0.0, mS_B = 0.0, mS_H = 0.0, mS_C = 0.0, trS = 0.0;
  double s_sphi = 0.0, s_sxr = 0.0, s_sxz = 0.0, s_spi = 0.0;
  double sc_phi = 0.0, sc_xr = 0.0, sc_xz = 0.0, sc_pi = 0.0;
  constexpr double smallpi = 3.14159265358979323846;
  const double dscale_s = P.scalardiss / G.dt;
  if constexpr (SCALAR) {
    const auto Fsp = fld(scalar_slot(SHIFT, SC_phi)), Fxr = fld(scalar_slot(SHIFT, SC_xi_r));
    const auto Fxz = fld(scalar_slot(SHIFT, SC_xi_z)), Fpi = fld(scalar_slot(SHIFT, SC_pi));
    sc_phi = Fsp(i, j); sc_xr = Fxr(i, j); sc_xz = Fxz(i, j); sc_pi = Fpi(i, j);
    double V = 0.0, VP = 0.0;
    if (P.scalarpotential == OA_POT_PHI2) {
      V = 0.5 * P.scalar_mass * P.scalar_mass * sc_phi * sc_phi;
      VP = P.scalar_mass * P.scalar_mass * sc_phi;
    } else if (P.scalarpotential == OA_POT_PHI4) {
      V = 0.5 * P.scalar_mass * P.scalar_mass * sc_phi * sc_phi + 0.25 * P.scalar_lambda * sc_phi * sc_phi * sc_phi * sc_phi;
      VP = P.scalar_mass * P.scalar_mass * sc_phi + P.scalar_lambda * sc_phi * sc_phi * sc_phi;
    }
    const double xr = sc_xr, xz = sc_xz, pi = sc_pi;
    const double q2 = (gA * xr * xr + gB * xz * xz + 2.0 * r * gC * xr * xz) * ipsi4;
    rho = 0.5 * (pi * pi + q2) + V;
    J_r = -(gA * xr + r * gC * xz) * pi * ipsi4;
    J_z = -(gB * xz + r * gC * xr) * pi * ipsi4;
    const double aux = 0.5 * (pi * pi - q2) - V;
    const double psi4 = 1.0 / ipsi4;
    mS_A = xr * xr + A * psi4 * aux;
    mS_B = xz * xz + B * psi4 * aux;
    mS_C = xr * xz * ir + C * psi4 * aux;
    mS_H = H * psi4 * aux;
    trS = (gA * mS_A + gB * mS_B + gH * mS_H + 2.0 * r2 * gC * mS_C) * ipsi4;

    // sources_scalar (scalar.f90:94-251)
    const D1 dpi = grad<IN>(Fpi, G, i, j);
    s_sphi = alpha * pi;
    if constexpr (SHIFT) s_sphi += br * xr + bz * xz;
    s_sxr = alpha * dpi.r + pi * Dra;
    s_sxz = alpha * dpi.z + pi * Dza;
    if constexpr (SHIFT) {
      const D1 dxr = grad<IN>(Fxr, G, i, j), dxz = grad<IN>(Fxz, G, i, j);
      s_sxr += br * s_adr<IN>(Fxr, G, i, j, br) + xr * dbr_.r + bz * s_adz<IN>(Fxr, G, i, j, bz) + xz * dbz_.r;
      s_sxz += br * s_adr<IN>(Fxz, G, i, j, br) + xr * dbr_.z + bz * s_adz<IN>(Fxz, G, i, j, bz) + xz * dbz_.z;
      (void)dxr; (void)dxz;
    }
    double dr_, dz_, lap;
    if (P.scalarmethod == OA_SCALARMETHOD_FIRST) {
      const D1 dxr = grad<IN>(Fxr, G, i, j), dxz = grad<IN>(Fxz, G, i, j);
      dr_ = xr; dz_ = xz;
      lap = alpha * (gA * dxr.r + gB * dxz.z + r * gC * (dxz.r + dxr.z));
    } else {
      const D2 ds = hess<IN>(Fsp, G, i, j);
      dr_ = ds.r; dz_ = ds.z;
      lap = alpha * (gA * ds.rr + gB * ds.zz + 2.0 * r * gC * ds.rz);
    }
    double spi = lap;
    spi += 2.0 * alpha * (Drp * (gA * dr_ + r * gC * dz_) + Dzp * (gB * dz_ + r * gC * dr_));
    spi += Dra * (gA * dr_ + r * gC * dz_) + Dza * (gB * dz_ + r * gC * dr_);
    spi += alpha * (dr_ * (dgA.r + r * dgC.z) + dz_ * (dgB.z + r * dgC.r + gC));
    spi += alpha * (gA * dr_ * ir + gC * dz_);
    spi = spi * ipsi4;
    spi += alpha * pi * trK;
    if (P.scalarpotential != OA_POT_NONE) spi -= alpha * VP;
    if constexpr (SHIFT) spi += br * s_adr<IN>(Fpi, G, i, j, br) + bz * s_adz<IN>(Fpi, G, i, j, bz);
    if (P.scalardiss != 0.0) {
      s_sphi -= s_ko<IN>(Fsp, G, i, j, dscale_s);
      spi -= s_ko<IN>(Fpi, G, i, j, dscale_s);
      s_sxr -= s_ko<IN>(Fxr, G, i, j, dscale_s);
      s_sxz -= s_ko<IN>(Fxz, G, i, j, dscale_s);
    }
    // radiative boundary on pi at every level (scalar.f90:244-251)
    const bool bnd = (i == G.Nr) || (j == G.Nz) || (!G.eqsym && j == 1 - G.g);
    if (!IN && bnd) {
      const double rr = sqrt(r * r + z * z);
      spi = -(r * dpi.r + z * dpi.z + pi) / rr;
    }
    s_spi = spi;
  }

  // ---- sources_geometry (sources_geometry.f90:77-575), lagrangian: sigma = 1
  const double dscale_g = P.geodiss / G.dt;
  double sphi = -alpha * trK / 6.0;
  double schi = 2.0 / 6.0 * alpha * chi * trK;
  double sA = -2.0 * alpha * KTA, sB = -2.0 * alpha * KTB, sH = -2.0 * alpha * KTH, sC = -2.0 * alpha * KTC;
  // advective derivatives of the geometry (only with a shift)
  double DAr_A = 0.0, DAz_A = 0.0, DAr_B = 0.0, DAz_B = 0.0, DAr_H = 0.0, DAz_H = 0.0, DAr_C = 0.0, DAz_C = 0.0;
  if constexpr (SHIFT) {
    adv(FA, D1{dA.r, dA.z}, DAr_A, DAz_A);
    adv(FB, D1{dB.r, dB.z}, DAr_B, DAz_B);
    adv(FH, D1{dH.r, dH.z}, DAr_H, DAz_H);
    adv(FC, D1{dC.r, dC.z}, DAr_C, DAz_C);
    sphi += br * DAr_phi + bz * DAz_phi + DIV_beta / 6.0;
    schi += br * DAr_chi + bz * DAz_chi - 2.0 / 6.0 * chi * DIV_beta;
    const double t23 = 2.0 * third;
    sA += br * DAr_A + bz * DAz_A + 2.0 * (A * dbr_.r + r * C * dbz_.r) - t23 * A * DIV_beta;
    sB += br * DAr_B + bz * DAz_B + 2.0 * (B * dbz_.z + r * C * dbr_.z) - t23 * B * DIV_beta;
    sH += br * DAr_H + bz * DAz_H + 2.0 * H * br * ir - t23 * H * DIV_beta;
    sC += br * DAr_C + bz * DAz_C + C * br * ir + C * (dbr_.r + dbz_.z) + (B * dbz_.r + A * dbr_.z) * ir
          - t23 * C * DIV_beta;
  }

  // trK
  double strK;
  if (P.slicing == OA_SLICING_MAXIMAL) {
    strK = 0.0;
  } else {
    strK = -Lapla + alpha * K2;
    if constexpr (SHIFT) {
      double DAr, DAz;
      adv(FK, dK, DAr, DAz);
      strK += br * DAr + bz * DAz;
    }
    if constexpr (SCALAR) strK += 4.0 * smallpi * alpha * (rho + trS);
    if (P.geodiss != 0.0) strK -= s_ko<IN>(FK, G, i, j, dscale_g);
  }

  // traceless extrinsic curvature
  const double tr = -Lapla + alpha * RSCAL;
  double sKTA = alpha * (trK * KTA - 2.0 * KT2_A) + (-d2A + alpha * RA) * ipsi4 - third * A * tr;
  double sKTB = alpha * (trK * KTB - 2.0 * KT2_B) + (-d2B + alpha * RB) * ipsi4 - third * B * tr;
  double sKTC = alpha * (trK * KTC - 2.0 * KT2_C) + (-d2C + alpha * RC) * ipsi4 - third * C * tr;
  if constexpr (SHIFT) {
    const double t23 = 2.0 * third;
    double DAr, DAz;
    adv(FKA, dKA, DAr, DAz);
    sKTA += br * DAr + bz * DAz + 2.0 * (KTA * dbr_.r + r * KTC * dbz_.r) - t23 * KTA * DIV_beta;
    adv(FKB, dKB, DAr, DAz);
    sKTB += br * DAr + bz * DAz + 2.0 * (KTB * dbz_.z + r * KTC * dbr_.z) - t23 * KTB * DIV_beta;
    adv(FKC, dKC, DAr, DAz);
    sKTC += br * DAr + bz * DAz + KTC * br * ir + KTC * (dbr_.r + dbz_.z) + (KTB * dbz_.r + KTA * dbr_.z) * ir
            - t23 * KTC * DIV_beta;
  }
  if constexpr (SCALAR) {
    const double psi4 = 1.0 / ipsi4;
    (void)psi4;
    sKTA -= 8.0 * smallpi * alpha * (mS_A * ipsi4 - third * A * trS);
    sKTB -= 8.0 * smallpi * alpha * (mS_B * ipsi4 - third * B * trS);
    sKTC -= 8.0 * smallpi * alpha * (mS_C * ipsi4 - third * C * trS);
  }
  if (P.geodiss != 0.0) {
    sKTA -= s_ko<IN>(FKA, G, i, j, dscale_g);
    sKTB -= s_ko<IN>(FKB, G, i, j, dscale_g);
    sKTC -= s_ko<IN>(FKC, G, i, j, dscale_g);
  }

  // Delta
  const double eta = P.eta;
  double sDr = -2.0 * (KTup_rr * Dra + KTup_rz * Dza) - 2.0 * alpha * DIVA_r;
  sDr += eta * alpha * (DIV_r + 6.0 * (KTup_rr * Drp + KTup_rz * Dzp) - 2.0 * third * (gA * dK.r + r * gC * dK.z));
  double sDz = -2.0 * (KTup_rz * Dra + KTup_zz * Dza) - 2.0 * alpha * DIVA_z;
  sDz += eta * alpha * (DIV_z + 6.0 * (KTup_rz * Drp + KTup_zz * Dzp) - 2.0 * third * (r * gC * dK.r + gB * dK.z));
  if constexpr (SHIFT) {
    double DAr, DAz;
    adv(FDr, dDr, DAr, DAz);
    sDr += br * DAr + bz * DAz - Delta_r * dbr_.r - Delta_z * dbr_.z;
    sDr += 2.0 * r * gC * dbr_.rz + gA * dbr_.rr + gB * dbr_.zz + gH * DD_beta_rr;
    sDr += third * (gA * Dr_DIV_beta + r * gC * Dz_DIV_beta + 2.0 * Delta_r * DIV_beta);
    adv(FDz, dDz, DAr, DAz);
    sDz += br * DAr + bz * DAz - Delta_r * dbz_.r - Delta_z * dbz_.z;
    sDz += 2.0 * r * gC * dbz_.rz + gA * dbz_.rr + gB * dbz_.zz + gH * dbz_.r * ir;
    sDz += third * (r * gC * Dr_DIV_beta + gB * Dz_DIV_beta + 2.0 * Delta_z * DIV_beta);
  }
  if constexpr (SCALAR) {
    const double psi4 = 1.0 / ipsi4;
    sDr -= 8.0 * smallpi * eta * alpha * J_r * psi4;
    sDz -= 8.0 * smallpi * eta * alpha * J_z * psi4;
  }
  if (P.geodiss != 0.0) {
    sDr -= s_ko<IN>(FDr, G, i, j, dscale_g);
    sDz -= s_ko<IN>(FDz, G, i, j, dscale_g);
  }

  // ---- sources_lapse (first-order Bona-Masso)
  double salpha = 0.0;
  if (P.slicing != OA_SLICING_STATIC && P.slicing != OA_SLICING_MAXIMAL) {
    double falpha;
    if (P.slicing == OA_SLICING_HARMONIC) falpha = alpha * alpha * P.gauge_f;
    else if (P.slicing == OA_SLICING_1PLUSLOG) falpha = alpha * P.gauge_f;
    else falpha = alpha * alpha + (P.gauge_f - 1.0);
    salpha = -falpha * trK;
    if constexpr (SHIFT) salpha += br * DAr_alpha + bz * DAz_alpha;
    if (P.gauge_eta != 0.0) salpha -= P.gauge_eta * log(fabs(alpha));
  }

  // ---- sources_shift (Gammadriver2)
  // the planes dtbeta is differentiated from (see KStep::dtb_pre)
  const bool dpre = (S.dtb_pre != nullptr);
  const auto Fdbr = fld(SHIFT ? shift_slot(SH_dtbeta_r) : 0), Fdbz = fld(SHIFT ? shift_slot(SH_dtbeta_z) : 0);
  const auto Pdbr = V.prefld(0), Pdbz = V.prefld(1);
  auto Ddbr = [&](int ii, int jj) { return dpre ? Pdbr(ii, jj) : Fdbr(ii, jj); };
  auto Ddbz = [&](int ii, int jj) { return dpre ? Pdbz(ii, jj) : Fdbz(ii, jj); };
  double sbr = 0.0, sbz = 0.0, sdbr = 0.0, sdbz = 0.0;
  if constexpr (SHIFT) {
    sbr = dbr;
    sbz = dbz;
    // the advective derivatives of sources_shift.f90:229-280 are the DA*
    // arrays: upwinded with shiftadvect, the centred ones otherwise
    if (P.driveradv) {
      double ar, az;
      adv(fld(shift_slot(SH_beta_r)), D1{dbr_.r, dbr_.z}, ar, az);
      sbr += br * ar + bz * az;
      adv(fld(shift_slot(SH_beta_z)), D1{dbz_.r, dbz_.z}, ar, az);
      sbz += br * ar + bz * az;
    }
    sdbr = P.drivercsi * sDr;
    sdbz = P.drivercsi * sDz;
    if (P.driveradv) {
      double tr_, tz_, dr_, dz_;
      adv(Ddbr, grad<IN>(Ddbr, G, i, j), tr_, tz_);
      adv(FDr, dDr, dr_, dz_);
      sdbr += br * (tr_ - dr_) + bz * (tz_ - dz_);
      adv(Ddbz, grad<IN>(Ddbz, G, i, j), tr_, tz_);
      adv(FDz, dDz, dr_, dz_);
      sdbz += br * (tr_ - dr_) + bz * (tz_ - dz_);
    }
    sdbr -= P.drivereta * dbr;
    sdbz -= P.drivereta * dbz;
  }

  // ---- radiative boundaries on level 0 (radiative_geometry.f90:58-237)
  if (!IN && P.radiative) {
    const bool onr = (i == G.Nr), onzR = (j == G.Nz), onzL = (!G.eqsym && j == 1 - G.g);
    if (onr || onzR || onzL) {
      const double irr = 1.0 / sqrt(r * r + z * z);
      const double va = sqrt(P.gauge_f);
      if (P.slicing != OA_SLICING_MAXIMAL) strK = -va * (r * dK.r + z * dK.z + trK) * irr;
      sKTA = -(r * dKA.r + z * dKA.z + KTA) * irr;
      sKTB = -(r * dKB.r + z * dKB.z + KTB) * irr;
      const double c = third * (1.0 - 1.0 / va);
      if (onr) { sKTA += c * strK * 2.0; sKTB -= c * strK; }
      if (onzL) { sKTA -= c * strK; sKTB += c * strK * 2.0; }
      if (onzR) { sKTA -= c * strK; sKTB += c * strK * 2.0; }
      sKTC = -(r * dKC.r + z * dKC.z + KTC) * irr;
      if constexpr (SHIFT) {
        if (P.drivercsi != 0.0) {
          const double vs = sqrt(4.0 * third * P.drivercsi);
          const D1 ddr = grad<IN>(Ddbr, G, i, j);
          const D1 ddz = grad<IN>(Ddbz, G, i, j);
          sdbr = -vs * (r * ddr.r + z * ddr.z + dbr) * irr;
          sdbz = -vs * (r * ddz.r + z * ddz.z + dbz) * irr;
        }
      }
    }
  }

  // ---- RK: accumulate (prl/arrays.pl:533-599) and update (:1282-1328)
  const bool first = (S.iter == 1), last = (S.iter == S.niter);
  // y_p and the accumulator of every slot, loaded together before any store
  // (so one memory round trip serves them all); on the first iteration the
  // staged input is y_p itself
  constexpr int NF = nslots(SHIFT, SCALAR);
  double yv[NF], av[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) yv[q] = yp[q * pl + k];
  const bool use_acc = S.rk4 && !first;
  if (use_acc) {
#pragma unroll
    for (int q = 0; q < NF; ++q) av[q] = acc[q * pl + k];
  }
  auto rk = [&](int slot, double s) -> double {
    double sv = s;
    if (S.rk4) {
      double* a = acc + slot * pl;
      if (!last) a[k] = first ? S.w * s : av[slot] + S.w * s;
      else sv = av[slot] + S.w * s;
    }
    return yv[slot] + S.dtw * sv;
  };

  const bool pt0 = (S.proj_t0 != 0);
  const double alpha_n = rk(S_alpha, salpha);
  Proj q;
  if (pt0 || !P.chimethod) {
    q.phi = rk(S_phi, sphi);
    q.chi = chi;
  } else {
    q.phi = 0.0;
    q.chi = rk(S_chi, schi);
  }
  q.A = rk(S_A, sA);
  q.B = rk(S_B, sB);
  q.H = rk(S_H, sH);
  q.C = rk(S_C, sC);
  const double trK_n = rk(S_trK, strK);
  const double KTA_n = rk(S_KTA, sKTA);
  const double KTB_n = rk(S_KTB, sKTB);
  const double KTC_n = rk(S_KTC, sKTC);
  double Dr_n = 0.0, Dz_n = 0.0;
  if (!pt0) {
    Dr_n = rk(S_Delta_r, sDr);
    Dz_n = rk(S_Delta_z, sDz);
  }
  project(q, r, pt0, P.chimethod, SCALAR);

  double* const yo = yout;
  put(yo + S_alpha * pl, G, i, j, alpha_n, 1.0, 1.0);
  put(yo + S_chi * pl, G, i, j, q.chi, 1.0, 1.0);
  if (pt0 || !P.chimethod) put(yo + S_phi * pl, G, i, j, q.phi, 1.0, 1.0);
  put(yo + S_A * pl, G, i, j, q.A, 1.0, 1.0);
  put(yo + S_B * pl, G, i, j, q.B, 1.0, 1.0);
  put(yo + S_H * pl, G, i, j, q.H, 1.0, 1.0);
  put(yo + S_C * pl, G, i, j, q.C, 1.0, -1.0);
  put(yo + S_trK * pl, G, i, j, trK_n, 1.0, 1.0);
  put(yo + S_KTA * pl, G, i, j, KTA_n, 1.0, 1.0);
  put(yo + S_KTB * pl, G, i, j, KTB_n, 1.0, 1.0);
  put(yo + S_KTC * pl, G, i, j, KTC_n, 1.0, -1.0);
  if (!pt0) {
    put(yo + S_Delta_r * pl, G, i, j, Dr_n, -1.0, 1.0);
    put(yo + S_Delta_z * pl, G, i, j, Dz_n, 1.0, -1.0);
  }
  if constexpr (SHIFT) {
    put(yo + shift_slot(SH_beta_r) * pl, G, i, j, rk(shift_slot(SH_beta_r), sbr), -1.0, 1.0);
    put(yo + shift_slot(SH_beta_z) * pl, G, i, j, rk(shift_slot(SH_beta_z), sbz), 1.0, -1.0);
    put(yo + shift_slot(SH_dtbeta_r) * pl, G, i, j, rk(shift_slot(SH_dtbeta_r), sdbr), -1.0, 1.0);
    put(yo + shift_slot(SH_dtbeta_z) * pl, G, i, j, rk(shift_slot(SH_dtbeta_z), sdbz), 1.0, -1.0);
  }
  if constexpr (SCALAR) {
    put(yo + scalar_slot(SHIFT, SC_phi) * pl, G, i, j, rk(scalar_slot(SHIFT, SC_phi), s_sphi), 1.0, 1.0);
    put(yo + scalar_slot(SHIFT, SC_xi_r) * pl, G, i, j, rk(scalar_slot(SHIFT, SC_xi_r), s_sxr), -1.0, 1.0);
    put(yo + scalar_slot(SHIFT, SC_xi_z) * pl, G, i, j, rk(scalar_slot(SHIFT, SC_xi_z), s_sxz), 1.0, -1.0);
    put(yo + scalar_slot(SHIFT, SC_pi) * pl, G, i, j, rk(scalar_slot(SHIFT, SC_pi), s_spi), 1.0, 1.0);
  }
}


// auxiliary on a whole level after upload or restriction: parity images and
// the pointwise projection of every computed point.
template <bool SHIFT, bool SCALAR>
OA_HD void project_point(const KGrid& G, const KPhys& P, int time0, double* __restrict__ y, int i, int j) {
  const long pl = G.plane;
  const long k = IDX(G, i, j);
  const double r = (static_cast<double>(i) - 0.5) * G.dr;
  Proj q{y[S_chi * pl + k], y[S_phi * pl + k], y[S_A * pl + k], y[S_B * pl + k], y[S_H * pl + k], y[S_C * pl + k]};
  project(q, r, time0 != 0, P.chimethod, SCALAR);
  constexpr int ns = nslots(SHIFT, SCALAR);
  for (int s = 0; s < ns; ++s) {
    double symr = 1.0, symz = 1.0;
    slot_parity(SHIFT, SCALAR, s, symr, symz);
    double v = y[s * pl + k];
    if (s == S_chi) v = q.chi;
    if (s == S_phi) v = q.phi;
    if (s == S_A) v = q.A;
    if (s == S_B) v = q.B;
    if (s == S_H) v = q.H;
    if (s == S_C) v = q.C;
    put(y + s * pl, G, i, j, v, symr, symz);
  }
}

// Delta = DelDef while the level clock is 0 (auxiliary_geometry.f90:1085-1094).
// Reads the projected metric (with parity images), writes Delta_r/z.
OA_HD void deldef_point(const KGrid& G, double* __restrict__ y, int i, int j) {
  const long pl = G.plane;
  const Plane FA{y + S_A * pl, G}, FB{y + S_B * pl, G}, FH{y + S_H * pl, G}, FC{y + S_C * pl, G};
  const double r = (static_cast<double>(i) - 0.5) * G.dr;
  auto fgA = [&](int ii, int jj) { return FB(ii, jj) * FH(ii, jj); };
  auto fgB = [&](int ii, int jj) { return FA(ii, jj) * FH(ii, jj); };
  auto fgC = [&](int ii, int jj) { return -FC(ii, jj) * FH(ii, jj); };
  const double A = FA(i, j), B = FB(i, j), H = FH(i, j), C = FC(i, j);
  const double lambda = (A - H) / (r * r);
  const double g_lambda = C * C - B * lambda;
  const double gC = -C * H;
  const double Dr_ = -s_d1r(fgA, G, i, j) - r * s_d1z(fgC, G, i, j) - r * g_lambda;
  const double Dz_ = -s_d1r(fgC, G, i, j) * r - s_d1z(fgB, G, i, j) - 2.0 * gC;
  put(y + S_Delta_r * pl, G, i, j, Dr_, -1.0, 1.0);
  put(y + S_Delta_z * pl, G, i, j, Dz_, 1.0, -1.0);
}

// saveold for the fine-boundary strips: copies the current state into the
// ring slot that becomes time level 0 (prl/arrays.pl:759-875).  Strip
// layouts: rR (m, jj) = m*nzt + jj with m = 0..g-1 the distance from i = Nr;
// zL / zR (n, ii) = n*nrt + ii with n the distance from j = 1-g / j = Nz.
OA_HD void saveold_point(const KGrid& G, const double* __restrict__ y, double* __restrict__ rR,
                         double* __restrict__ zL, double* __restrict__ zR, int f, int q) {
  const int nrt = G.Nr + G.g, nzt = G.Nz + G.g;
  const long pl = G.plane;
  const double* yf = y + f * pl;
  const int per_f_r = G.g * nzt, per_f_z = G.g * nrt;
  if (q < per_f_r) {
    const int m = q / nzt, jj = q % nzt;
    rR[f * per_f_r + q] = yf[IDX(G, G.Nr - m, jj + 1 - G.g)];
  }
  if (q < per_f_z) {
    const int n = q / nrt, ii = q % nrt;
    zL[f * per_f_z + q] = yf[IDX(G, ii + 1 - G.g, 1 - G.g + n)];
    zR[f * per_f_z + q] = yf[IDX(G, ii + 1 - G.g, G.Nz - n)];
  }
}

// ------------------------------------------------------ constraints
// constraints(box, level) (src/geometry/constraints.f90:36-206), hdet = 1,
// no angmom: Hamiltonian, momentum and Delta constraints at (i, j), from the
// geometric auxiliaries of the current state in gaux (RSCAL, DIV_KT) and
// first derivatives read from the state planes.  Output planes, in order:
// ham, mom_r, mom_z, CDelta_r, CDelta_z; the parity images are written with
// the point.  The routine's smallpi = acos(-1.0) (:41) has a default-kind
// (single precision) argument, so pi there is the single-precision value.
enum ConsSlot { CN_ham = 0, CN_mom_r, CN_mom_z, CN_CDelta_r, CN_CDelta_z, NCONS };

template <bool PHIMODE, bool SCALAR>
OA_HD void constraints_point(const KGrid& G, const KPhys& P, int sc0, const double* __restrict__ y,
                             const double* __restrict__ gaux, double* __restrict__ out, int i, int j) {
  const long pl = G.plane;
  const long k = IDX(G, i, j);
  const GlobalView V{y, G, nullptr};
  const double r = (static_cast<double>(i) - 0.5) * G.dr;
  const double r2 = r * r;
  const double ir = 1.0 / r;
  const auto FA = V.fld(S_A), FB = V.fld(S_B), FH = V.fld(S_H), FC = V.fld(S_C);
  const double chi = V.fld(S_chi)(i, j);
  const double A = FA(i, j), B = FB(i, j), H = FH(i, j), C = FC(i, j);
  const double trK = V.fld(S_trK)(i, j), KTA = V.fld(S_KTA)(i, j), KTB = V.fld(S_KTB)(i, j);
  const double KTC = V.fld(S_KTC)(i, j);
  // conformal factor derivatives, as the auxiliary call of this state formed them
  double Drp, Dzp;
  if constexpr (PHIMODE) {
    const D1 dp = grad(V.fld(S_phi), G, i, j);
    Drp = dp.r;
    Dzp = dp.z;
  } else {
    const D1 dc = grad(V.fld(S_chi), G, i, j);
    const double ic = 1.0 / chi;
    Drp = -0.5 * (dc.r * ic);
    Dzp = -0.5 * (dc.z * ic);
  }
  const double ipsi4 = chi * chi;
  // conformal metric inverse (hdet = 1) and its derivatives
  const double gA = B * H, gB = A * H, gH = 1.0 / H, gC = -C * H;
  auto fgA = [&](int ii, int jj) { return FB(ii, jj) * FH(ii, jj); };
  auto fgB = [&](int ii, int jj) { return FA(ii, jj) * FH(ii, jj); };
  auto fgC = [&](int ii, int jj) { return -FC(ii, jj) * FH(ii, jj); };
  const D1 dgA = grad(fgA, G, i, j), dgB = grad(fgB, G, i, j), dgC = grad(fgC, G, i, j);
  const D1 dK = grad(V.fld(S_trK), G, i, j);
  const double KTH = -(gA * KTA + gB * KTB + 2.0 * r2 * gC * KTC) * H;
  const double lambda = (A - H) * ir * ir;
  const double g_lambda = C * C - B * lambda;
  const double r3 = r2 * r;
  const double gA2 = gA * gA, gB2 = gB * gB, gC2 = gC * gC;
  const double KTup_rr = (2.0 * gA * gC * KTC + gC2 * KTB) * r2 + gA2 * KTA;
  const double KTup_zz = (gC2 * KTA + 2.0 * gC * gB * KTC) * r2 + gB2 * KTB;
  const double KTup_rz = gC2 * r3 * KTC + (gC * gA * KTA + gB * gA * KTC + gB * gC * KTB) * r;
  const double KT2_A = (2.0 * KTA * gC * KTC + KTC * KTC * gB) * r2 + gA * KTA * KTA;
  const double KT2_B = (KTC * KTC * gA + 2.0 * KTC * gC * KTB) * r2 + gB * KTB * KTB;
  const double KT2_H = KTH * KTH * gH;
  const double KT2_C = KTC * KTC * gC * r2 + KTA * gA * KTC + KTA * gC * KTB + KTC * gB * KTB;
  const double KT2 = gA * KT2_A + gB * KT2_B + gH * KT2_H + 2.0 * gC * KT2_C * r2;
  const double third = 1.0 / 3.0, two = 2.0, half = 0.5;
  const double* const ax = gaux + k;
  const double RSCAL = ax[AUX_RSCAL * pl], DIV_r = ax[AUX_DIV_r * pl], DIV_z = ax[AUX_DIV_z * pl];
  double ham = RSCAL - KT2 + two * third * trK * trK;                                  // :57
  double mom_r = DIV_r + 6.0 * (KTup_rr * Drp + KTup_rz * Dzp) - two * third * (gA * dK.r + r * gC * dK.z);
  double mom_z = DIV_z + 6.0 * (KTup_rz * Drp + KTup_zz * Dzp) - two * third * (r * gC * dK.r + gB * dK.z);
  if constexpr (SCALAR) {
    const double smallpi = static_cast<double>(3.14159274101257324f);   // acos(-1.0) in single precision
    const double sphi = V.fld(sc0 + SC_phi)(i, j), xr = V.fld(sc0 + SC_xi_r)(i, j);
    const double xz = V.fld(sc0 + SC_xi_z)(i, j), pi = V.fld(sc0 + SC_pi)(i, j);
    double Vp = 0.0;
    if (P.scalarpotential == OA_POT_PHI2) {
      Vp = 0.5 * P.scalar_mass * P.scalar_mass * sphi * sphi;
    } else if (P.scalarpotential == OA_POT_PHI4) {
      Vp = 0.5 * P.scalar_mass * P.scalar_mass * sphi * sphi + 0.25 * P.scalar_lambda * sphi * sphi * sphi * sphi;
    }
    const double q2 = (gA * xr * xr + gB * xz * xz + 2.0 * r * gC * xr * xz) * ipsi4;
    const double rho = 0.5 * (pi * pi + q2) + Vp;
    const double J_r = -(gA * xr + r * gC * xz) * pi * ipsi4;
    const double J_z = -(gB * xz + r * gC * xr) * pi * ipsi4;
    const double psi4 = 1.0 / ipsi4;
    ham = ham - 16.0 * smallpi * rho;                                                   // :60
    mom_r = mom_r - 8.0 * smallpi * J_r * psi4;                                         // :120
    mom_z = mom_z - 8.0 * smallpi * J_z * psi4;                                         // :133
  }
  // hdet = 1: the Dr_hdet, Dz_hdet terms vanish
  const double CDr = V.fld(S_Delta_r)(i, j) + (dgA.r + r * dgC.z + r * g_lambda);     // :165-166
  const double CDz = V.fld(S_Delta_z)(i, j) + (dgC.r * r + dgB.z + two * gC);          // :172-173
  (void)half;
  put(out + CN_ham * pl, G, i, j, ham, 1.0, 1.0);
  put(out + CN_mom_r * pl, G, i, j, mom_r, -1.0, 1.0);
  put(out + CN_mom_z * pl, G, i, j, mom_z, 1.0, -1.0);
  put(out + CN_CDelta_r * pl, G, i, j, CDr, -1.0, 1.0);
  put(out + CN_CDelta_z * pl, G, i, j, CDz, 1.0, -1.0);
}

// ------------------------------------------------------ interpolation
// interp.f90:140-263 on a level held in z-fastest planes: the cell search on
// floating coordinates, the exact-node copy, the bilinear fallback near edges
// and the product-form Lagrange weights, evaluated as the reference does.
struct Cell {
  int i0, j0;
  int copy, lin;
  double wr[4], wz[4];
};

OA_HD double coord_r(const KGrid& G, int i) { return (static_cast<double>(i) - 0.5) * G.dr; }
OA_HD double coord_z(const KGrid& G, int j) { return (static_cast<double>(j) - G.zoff) * G.dz; }

OA_HD Cell locate(const KGrid& G, double r0, double z0, int interporder) {
  Cell c;
  const double rmin = coord_r(G, 1 - G.g), zmin = coord_z(G, 1 - G.g);
  const double rmax = coord_r(G, G.Nr), zmax = coord_z(G, G.Nz);
  c.i0 = static_cast<int>((r0 - rmin) / G.dr) + 1 - G.g;
  c.j0 = static_cast<int>((z0 - zmin) / G.dz) + 1 - G.g;
  const double deltar = r0 - coord_r(G, c.i0);
  const double deltaz = z0 - coord_z(G, c.j0);
  c.copy = (deltar == 0.0 && deltaz == 0.0);
  c.lin = ((rmax - r0) < G.dr) || ((zmax - z0) < G.dz) || ((r0 - rmin) < G.dr) || ((z0 - zmin) < G.dz) ||
          interporder == OA_INTERP_BILINEAR;
  const double auxr = deltar / G.dr, auxz = deltaz / G.dz;
  if (c.lin) {
    c.wr[0] = 0.0; c.wr[1] = 1.0 - auxr; c.wr[2] = auxr; c.wr[3] = 0.0;
    c.wz[0] = 0.0; c.wz[1] = 1.0 - auxz; c.wz[2] = auxz; c.wz[3] = 0.0;
  } else {
    for (int a = -1; a <= 2; ++a) {
      double wr = 1.0, wz = 1.0;
      for (int m = -1; m <= 2; ++m) {
        if (m == a) continue;
        wr = wr * (auxr - static_cast<double>(m)) / static_cast<double>(a - m);
        wz = wz * (auxz - static_cast<double>(m)) / static_cast<double>(a - m);
      }
      c.wr[a + 1] = wr;
      c.wz[a + 1] = wz;
    }
  }
  return c;
}

OA_HD double apply(const Cell& c, const double* __restrict__ f, const KGrid& G) {
  if (c.copy) return f[IDX(G, c.i0, c.j0)];
  double out = 0.0;
  if (c.lin) {
    for (int a = 0; a <= 1; ++a)
      for (int b = 0; b <= 1; ++b) out += c.wr[a + 1] * c.wz[b + 1] * f[IDX(G, c.i0 + a, c.j0 + b)];
    return out;
  }
  for (int a = -1; a <= 2; ++a)
    for (int b = -1; b <= 2; ++b) out += c.wr[a + 1] * c.wz[b + 1] * f[IDX(G, c.i0 + a, c.j0 + b)];
  return out;
}

// apply() for N fields of one level at the same cell: one branch on the
// cell kind, then each field with apply()'s own order of operations.
template <int N>
OA_HD void apply_many(const Cell& c, const double* __restrict__ y, long plane, const int (&slot)[N], double (&out)[N],
                      const KGrid& G) {
  if (c.copy) {
#pragma unroll
    for (int q = 0; q < N; ++q) out[q] = y[slot[q] * plane + IDX(G, c.i0, c.j0)];
    return;
  }
  if (c.lin) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      const double* f = y + slot[q] * plane;
      double o = 0.0;
      for (int a = 0; a <= 1; ++a)
        for (int b = 0; b <= 1; ++b) o += c.wr[a + 1] * c.wz[b + 1] * f[IDX(G, c.i0 + a, c.j0 + b)];
      out[q] = o;
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const double* f = y + slot[q] * plane;
    double o = 0.0;
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b) o += c.wr[a + 1] * c.wz[b + 1] * f[IDX(G, c.i0 + a, c.j0 + b)];
    out[q] = o;
  }
}

// finebound(box 0, level, dtw, .true.).  The fine boundary points and the
// coarse cells they interpolate from are fixed per level and zone depth, so
// fb_geometry resolves them once into FbPoint records; finebound_point then
// does the per-call work (one thread per (point, field group)).
// Point classes: upper-r strip rows not claimed by a z strip, then the upper
// and (without equatorial symmetry) lower z strips over the whole r range,
// which is also the reference's overwrite order at the corners.
struct FbPoint {
  Cell c;            // coarse cell and interpolation weights
  int i, j;          // fine point
  int which;         // strip: 0 upper r, 1 upper z, 2 lower z; -1: no point
  int sidx;          // index of the point in its strip
  int strip_ok;      // the point lies in the saved strips (depth < ghost)
};

// number of records for zone depth D (last layer index)
OA_HHD long fb_count(const KGrid& Gf, int D) {
  const int nz_strips = Gf.eqsym ? 1 : 2;
  return static_cast<long>(D + 1) * (Gf.Nz + Gf.g) + static_cast<long>(D + 1) * nz_strips * (Gf.Nr + Gf.g);
}

OA_HD void fb_geometry(const KGrid& Gf, const KGrid& Gc, int D, int interporder, long q, FbPoint& fp) {
  const int g = Gf.g;
  const int nrt = Gf.Nr + g, nzt = Gf.Nz + g;
  const long n_r = static_cast<long>(D + 1) * nzt;        // upper-r strip (rows filtered below)
  fp.which = -1;
  int i, j, m = 0, n = 0, which;   // which: 0 rR, 1 zR, 2 zL
  if (q < n_r) {
    m = static_cast<int>(q / nzt);
    j = static_cast<int>(q % nzt) + 1 - g;
    i = Gf.Nr - m;
    const bool in_zR = (j >= Gf.Nz - D);
    const bool in_zL = (!Gf.eqsym && j <= 1 - g + D);
    if (in_zR || in_zL) return;
    which = 0;
  } else {
    const long qq = q - n_r;
    const int row = static_cast<int>(qq / nrt);
    i = static_cast<int>(qq % nrt) + 1 - g;
    if (row <= D) { n = row; j = Gf.Nz - n; which = 1; }
    else { n = row - (D + 1); j = 1 - g + n; which = 2; }
  }
  if (i < 1) return;   // axis images are refreshed by parity
  // coordinates of the fine point as finebound.f90 forms them
  const double rminf = coord_r(Gf, 1 - g), rmaxf = coord_r(Gf, Gf.Nr);
  const double zminf = coord_z(Gf, 1 - g), zmaxf = coord_z(Gf, Gf.Nz);
  double r0, z0;
  if (which == 0) { r0 = rmaxf - static_cast<double>(m) * Gf.dr; z0 = zminf + static_cast<double>(j + g - 1) * Gf.dz; }
  else if (which == 1) { r0 = rminf + static_cast<double>(i + g - 1) * Gf.dr; z0 = zmaxf - static_cast<double>(n) * Gf.dz; }
  else { r0 = rminf + static_cast<double>(i + g - 1) * Gf.dr; z0 = zminf + static_cast<double>(n) * Gf.dz; }
  fp.c = locate(Gc, r0, z0, interporder);
  fp.i = i;
  fp.j = j;
  fp.which = which;
  fp.sidx = (which == 0) ? m * nzt + (j + g - 1) : n * nrt + (i + g - 1);
  fp.strip_ok = ((which == 0 ? m : n) < g) ? 1 : 0;
}

template <bool SHIFT, bool SCALAR>
OA_HD void finebound_point(const KGrid& Gf, const KGrid& Gc, const KPhys& P, const KTime& T, int time0,
                           const double* __restrict__ yc, double* __restrict__ yf,
                           const double* __restrict__ rR0, const double* __restrict__ rR1,
                           const double* __restrict__ zL0, const double* __restrict__ zL1,
                           const double* __restrict__ zR0, const double* __restrict__ zR1, const FbPoint& fp_in,
                           int grp) {
  const FbPoint fp = fp_in;   // the whole record in one go
  if (fp.which < 0) return;
  const int g = Gf.g, i = fp.i, j = fp.j;
  const Cell& c = fp.c;
  const long pf = Gf.plane, pc = Gc.plane;
  // the strip this point belongs to, for the time levels
  const double* s0 = (fp.which == 0) ? rR0 : (fp.which == 1) ? zR0 : zL0;
  const double* s1 = (fp.which == 0) ? rR1 : (fp.which == 1) ? zR1 : zL1;
  const long sper = static_cast<long>(g) * ((fp.which == 0) ? (Gf.Nz + g) : (Gf.Nr + g));
  const long sidx = fp.sidx;
  const bool strip_ok = fp.strip_ok != 0;

  // time interpolation of one injected value (bound_interp.inc)
  auto value = [&](int s) -> double {
    const double aux2 = apply(c, yc + s * pc, Gc);
    const double b0 = strip_ok ? s0[s * sper + sidx] : 0.0;   // state at t0 (var_p)
    if (T.border == 1) return T.c_aux * aux2 + (T.same ? 0.0 : T.c_p * b0);
    const double b1 = strip_ok ? s1[s * sper + sidx] : 0.0;
    return T.c0 * aux2 + (T.same ? 0.0 : T.c1 * b0 + T.c2 * b1);
  };
  if (grp == 0) {
    // the fields coupled by the projection (chi/phi, factordet) go together;
    // the six interpolations share one branch on the cell kind and issue
    // their loads together (value() for each field in turn would wait for
    // each field's loads before starting the next)
    constexpr int kN = 6;
    const int slot[kN] = {S_chi, S_phi, S_A, S_B, S_H, S_C};
    double v[kN];
    apply_many<kN>(c, yc, pc, slot, v, Gc);
    double b0[kN], b1[kN];
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      b0[q] = strip_ok ? s0[slot[q] * sper + sidx] : 0.0;
      b1[q] = (strip_ok && T.border != 1) ? s1[slot[q] * sper + sidx] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kN; ++q) {
      if (T.border == 1) v[q] = T.c_aux * v[q] + (T.same ? 0.0 : T.c_p * b0[q]);
      else v[q] = T.c0 * v[q] + (T.same ? 0.0 : T.c1 * b0[q] + T.c2 * b1[q]);
    }
    const double r = coord_r(Gf, i);
    Proj qp{v[0], v[1], v[2], v[3], v[4], v[5]};
    project(qp, r, time0 != 0, P.chimethod, SCALAR);
    put(yf + S_chi * pf, Gf, i, j, qp.chi, 1.0, 1.0);
    if (time0 || !P.chimethod) put(yf + S_phi * pf, Gf, i, j, qp.phi, 1.0, 1.0);   // phi is not kept after the first step
    put(yf + S_A * pf, Gf, i, j, qp.A, 1.0, 1.0);
    put(yf + S_B * pf, Gf, i, j, qp.B, 1.0, 1.0);
    put(yf + S_H * pf, Gf, i, j, qp.H, 1.0, 1.0);
    put(yf + S_C * pf, Gf, i, j, qp.C, 1.0, -1.0);
  } else {
    // every other field on its own: alpha, then the slots after S_C
    const int s = (grp == 1) ? S_alpha : S_C + grp - 1;
    double symr, symz;
    slot_parity(SHIFT, SCALAR, s, symr, symz);
    put(yf + s * pf, Gf, i, j, value(s), symr, symz);
  }
}

// groups of finebound_point: the projected set, then one per remaining slot
constexpr int finebound_groups(bool shift, bool scalar) { return 2 + nslots(shift, scalar) - (S_C + 1); }

// restrict(box 0, level, .true.): coarse points well inside the fine level
// take the bicubic interpolant of the fine level (restrict.f90:88-162);
// one call per (coarse point, field slot s).
OA_HD void restrict_point(const KGrid& Gf, const KGrid& Gc, int s, int interporder,
                          const double* __restrict__ yf, double* __restrict__ yc, int ic, int jc) {
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
  const Cell c = locate(Gf, r0, z0, interporder);
  const long k = IDX(Gc, ic, jc);
  yc[s * Gc.plane + k] = apply(c, yf + s * Gf.plane, Gf);
}

}  // namespace oax
