/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see or_internal.h).
 *
 * The apparent-horizon finder restated serially in the reference's own
 * order: horizon_finder (src/geometry/horizon_finder.f90:2-237) with its
 * moving centres, ahshoot (:246-1103), ahspectral (:1113-1972) and
 * ah_source (:1983-2752), for box 0 and no angular momentum (angmom =
 * .false.: no C1/C2 arrays in this hot path).  Interpolation is or_interp
 * (interp.f90:2-270) on the auxiliary arrays of the last auxiliary call.
 *
 * Pinned by tests/test_oracle_horizon.py against closed forms: the
 * Schwarzschild horizon at isotropic radius M/2 (mass M), the same horizon
 * seen through a non-trivial axisymmetric coordinate change (A, B, C, H and
 * their derivatives all non-zero), the Painleve-Gullstrand slice whose
 * horizon at r = 2M is fixed by the extrinsic curvature alone, equatorial
 * symmetry, off-centre horizons, and flat space (no horizon).  The spectral
 * finder by tests/test_oracle_spectral.py: the same closed forms (its fixed
 * point is the horizon whatever the flow's step control), its cosine
 * coefficients of a round sphere (only a_0 = 2 r), the agreement of its
 * surface with the shooting method's on a distorted horizon, and the
 * projection of a known surface onto coefficients.  Not pinned beyond
 * that: the step control of the flow (AA halved / grown), which decides the
 * iteration count only.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "or_internal.h"

#define AH_NTHETAMAX 50000          /* Nthetamax (:350) */

/* Step control of the integration (:657-662): relative and absolute
 * tolerance on dh/dtheta.  The reference's values; the pin tests tighten
 * them to expose the grid convergence under the controller's own floor. */
static double ah_tol_rel = 1.0e-4, ah_tol_abs = 1.0e-8;

void or_ah_set_tolerance(double rel, double abs_) {
  ah_tol_rel = rel;
  ah_tol_abs = abs_;
}

/* physical extrinsic curvature KA, KB, KC, KH of every level
 * (auxiliary_geometry.f90:1038-1041) */
typedef struct {
  double *KA, *KB, *KC, *KH;
} ah_curv;

typedef struct {
  double dradlast, area;
  int err, ntheta;
} ah_shot;

/* KA, KB, KC, KH = psi4 (KT* + X trK / 3) on every level
 * (auxiliary_geometry.f90:1038-1041), from the last auxiliary arrays */
static ah_curv* ah_curv_new(const or_handle* H) {
  ah_curv* K = (ah_curv*)calloc((size_t)H->nlev, sizeof(ah_curv));
  const double third = 1.0 / 3.0;
  for (int l = 0; l < H->nlev; ++l) {
    const or_level* L = &H->L[l];
    K[l].KA = (double*)malloc(sizeof(double) * (size_t)L->n);
    K[l].KB = (double*)malloc(sizeof(double) * (size_t)L->n);
    K[l].KC = (double*)malloc(sizeof(double) * (size_t)L->n);
    K[l].KH = (double*)malloc(sizeof(double) * (size_t)L->n);
    ALL(L, q) {
      const double trK = L->var[OF_trK][q];
      K[l].KA[q] = L->psi4[q] * (L->var[OF_KTA][q] + third * L->var[OF_A][q] * trK);
      K[l].KB[q] = L->psi4[q] * (L->var[OF_KTB][q] + third * L->var[OF_B][q] * trK);
      K[l].KH[q] = L->psi4[q] * (L->var[OF_KTH][q] + third * L->var[OF_H][q] * trK);
      K[l].KC[q] = L->psi4[q] * (L->var[OF_KTC][q] + third * L->var[OF_C][q] * trK);
    }
  }
  return K;
}

static void ah_curv_free(const or_handle* H, ah_curv* K) {
  for (int l = 0; l < H->nlev; ++l) {
    free(K[l].KA);
    free(K[l].KB);
    free(K[l].KC);
    free(K[l].KH);
  }
  free(K);
}

/* finest level whose interior (one cell in from its edges) holds (r0, z0)
 * (:440-452, :2178-2192) */
static int ah_level(const or_handle* H, double r0, double z0) {
  int level = 0;
  for (int ll = 1; ll < H->nlev; ++ll) {
    const or_level* L = &H->L[ll];
    if (r0 > L->rmin + L->dr && r0 < L->rmax - L->dr && z0 > L->zmin + L->dz && z0 < L->zmax - L->dz) level = ll;
  }
  return level;
}

/* ah_source (:1983-2752): h'' (sdh) and the area element (sarea) at
 * (h, h', theta) for a horizon centred at zcenter on the axis.  Returns the
 * error flag. */
static int ah_source(const or_handle* H, const ah_curv* K, double rad, double dh, double th, double zcenter,
                     double* sdh, double* sarea) {
  const oa_params* p = &H->p;
  const or_level* L0 = &H->L[0];
  const double zero = 0.0, half = 0.5, one = 1.0, two = 2.0;
  const double smallpi = acos(-1.0);
  const double ahrmax = (p->ahrmax > 0.9) ? 0.9 : p->ahrmax;   /* ahshoot :367-373 */

  if (rad < 0.0) {                                               /* :2140-2145 */
    *sdh = 0.0;
    return 1;
  }
  const double rc = rad * sin(th);                               /* :2151-2159 */
  const double zc = rad * cos(th);
  const double r0 = rc;
  double z0 = zc + zcenter;
  if (p->eqsym && z0 < 0.0) z0 = -z0;
  if (r0 > ahrmax * L0->rmax || fabs(z0) > ahrmax * L0->zmax) {  /* :2163-2168 */
    *sdh = 0.0;
    return 1;
  }
  /* :2172-2178 (the z test against zmax subtracts drl, as the reference does) */
  if (r0 < L0->rmin + L0->dr || r0 > L0->rmax - L0->dr || z0 < L0->zmin + L0->dz || z0 > L0->zmax - L0->dr) {
    *sdh = 0.0;
    return 1;
  }
  if (z0 * zcenter < 0.0 && p->ahmethod == OA_AH_SHOOT) {        /* :2185-2190 */
    *sdh = 0.0;
    return 1;
  }

  const int lev = ah_level(H, r0, z0);
  const or_level* L = &H->L[lev];
  const ah_curv* KL = &K[lev];
#define IP(arr) or_interp(p, L, (arr), r0, z0)

  /* conformal factor (:2215-2245) */
  const double interp_psi = IP(L->psi);
  const double interp_psi4 = (interp_psi * interp_psi) * (interp_psi * interp_psi);
  const double interp_Drphi = IP(L->Dr_phi);
  const double interp_Dzphi = IP(L->Dz_phi);
  double Dphi[3];
  Dphi[0] = (rc * interp_Drphi + zc * interp_Dzphi) / rad;
  Dphi[1] = (zc * interp_Drphi - rc * interp_Dzphi);
  Dphi[2] = zero;

  /* conformal metric (:2252-2340) */
  const double iA = IP(L->var[OF_A]), iDrA = IP(L->Dr_A), iDzA = IP(L->Dz_A);
  const double iB = IP(L->var[OF_B]), iDrB = IP(L->Dr_B), iDzB = IP(L->Dz_B);
  const double iC = IP(L->var[OF_C]), iDrC = IP(L->Dr_C), iDzC = IP(L->Dz_C);
  const double iH = IP(L->var[OF_H]), iDrH = IP(L->Dr_H), iDzH = IP(L->Dz_H);

  const double rc2 = rc * rc, zc2 = zc * zc, rad2 = rad * rad;
  double gd[3][3], gu[3][3];
  gd[0][0] = (rc2 * iA + zc2 * iB + two * rc2 * zc * iC) / rad2;                       /* :2320-2323 */
  gd[1][1] = (zc2 * iA + rc2 * iB - two * rc2 * zc * iC);
  gd[0][1] = rc * (zc * (iA - iB) + (zc2 - rc2) * iC) / rad;
  gd[2][2] = rc2 * iH;
  gd[0][2] = zero;
  gd[1][2] = zero;
  gd[1][0] = gd[0][1];
  gd[2][0] = gd[0][2];
  gd[2][1] = gd[1][2];

  if (th != zero && th != smallpi) {                                                 /* :2345-2370 */
    const double aux = gd[0][0] * gd[1][1] - gd[0][1] * gd[0][1];
    gu[0][0] = gd[1][1] / aux;
    gu[1][1] = gd[0][0] / aux;
    gu[2][2] = one / gd[2][2];
    gu[0][1] = -gd[0][1] / aux;
    gu[0][2] = zero;
    gu[1][2] = zero;
  } else {
    gu[0][0] = one / gd[0][0];
    gu[1][1] = one / gd[1][1];
    gu[2][2] = zero;
    gu[0][1] = zero;
    gu[0][2] = zero;
    gu[1][2] = zero;
  }
  gu[1][0] = gu[0][1];
  gu[2][0] = gu[0][2];
  gu[2][1] = gu[1][2];

  /* metric derivatives in spherical coordinates, Dg[d][i][j] (:2388-2459) */
  double Dg[3][3][3];
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

  Dg[0][0][2] = zero;
  Dg[1][0][2] = zero;
  Dg[0][1][2] = zero;
  Dg[1][1][2] = zero;

  for (int d = 0; d < 2; ++d) {                                                       /* :2461-2472 */
    Dg[d][1][0] = Dg[d][0][1];
    Dg[d][2][0] = Dg[d][0][2];
    Dg[d][2][1] = Dg[d][1][2];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Dg[2][i][j] = zero;

  /* extrinsic curvature (:2479-2530) */
  const double interp_trK = IP(L->var[OF_trK]);
  const double iKA = IP(KL->KA), iKB = IP(KL->KB), iKC = IP(KL->KC), iKH = IP(KL->KH);
  double Kd[3][3];
  Kd[0][0] = (rc2 * iKA + zc2 * iKB + two * rc2 * zc * iKC) / rad2;
  Kd[1][1] = (zc2 * iKA + rc2 * iKB - two * rc2 * zc * iKC);
  Kd[0][1] = rc * (zc * (iKA - iKB) + (zc2 - rc2) * iKC) / rad;
  Kd[2][2] = rc2 * iKH;
  Kd[0][2] = zero;
  Kd[1][2] = zero;
  Kd[1][0] = Kd[0][1];
  Kd[2][0] = Kd[0][2];
  Kd[2][1] = Kd[1][2];
#undef IP

  /* gradient of F = r - h(theta), its norm, the denominator (:2537-2552) */
  double DF[3];
  for (int i = 0; i < 3; ++i) DF[i] = (gu[0][i] - gu[1][i] * dh) / interp_psi4;
  const double u2 = (gu[0][0] - two * gu[0][1] * dh + gu[1][1] * dh * dh) / interp_psi4;
  const double u = sqrt(fabs(u2));
  const double denom = (gu[0][0] * gu[1][1] - gu[0][1] * gu[0][1]) / (interp_psi4 * interp_psi4);

  /* Christoffel symbols with an upper index 1 or 2 (:2559-2608) */
  double Gamma[2][3][3];
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double gm = zero;
        for (int m = 0; m < 3; ++m) gm = gm + half * gu[k][m] * (Dg[i][j][m] + Dg[j][i][m] - Dg[m][i][j]);
        Gamma[k][i][j] = gm;
      }
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double dki = (k == i) ? one : zero, dkj = (k == j) ? one : zero;
        Gamma[k][i][j] = Gamma[k][i][j] + two * (dki * Dphi[j] + dkj * Dphi[i] -
                                                 gd[i][j] * (gu[k][0] * Dphi[0] + gu[k][1] * Dphi[1]));
      }

  /* the source, term by term (:2615-2650) */
  double s = zero;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s = s + u2 * gu[i][j] * (Gamma[0][i][j] - Gamma[1][i][j] * dh) / interp_psi4;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s = s - DF[i] * DF[j] * (Gamma[0][i][j] - Gamma[1][i][j] * dh);
  s = s + u * u * u * interp_trK;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s = s - u * DF[i] * DF[j] * Kd[i][j];
  s = -s / denom;

  if (fabs(th) < 1.0e-5 || fabs(th - smallpi) < 1.0e-5) s = zero;                     /* :2663-2667 */
  if (s != s || fabs(s) > 1.0e50) {                                                    /* :2671-2676 */
    *sdh = 1.0e10;
    return 1;
  }
  *sdh = s;

  /* area element (:2708-2716) */
  const double aux = two * smallpi * interp_psi4;
  *sarea = aux * sqrt(fabs(gd[2][2] * (dh * dh * gd[0][0] + two * dh * gd[0][1] + gd[1][1])));
  return 0;
}

/* One integration of the horizon ODE from the axis point zinit (ahshoot
 * :478-706): adaptive RK4 with step doubling.  theta/rad, when not NULL,
 * receive the accepted points (AH_NTHETAMAX + 1 room). */
static ah_shot ah_integrate(const or_handle* H, const ah_curv* K, double zinit, double zcenter, double* theta_out,
                            double* rad_out) {
  const oa_params* p = &H->p;
  const double zero = 0.0, half = 0.5, two = 2.0;
  const double smallpi = acos(-1.0);
  ah_shot out = {0.0, 0.0, 0, 0};

  double rad = fabs(zinit - zcenter);                     /* :484 */
  double drad = 0.0, area = 0.0;
  double dth = smallpi / 100.0;                           /* :513-520 */
  double theta;
  if (zcenter >= zero) {
    theta = 1.0e-3;
  } else {
    dth = -dth;
    theta = smallpi - 1.0e-3;
  }
  double thetafinal;                                      /* :526-532 */
  if (p->eqsym && zcenter == zero) thetafinal = half * smallpi;
  else if (zcenter >= zero) thetafinal = smallpi - 1.0e-3;
  else thetafinal = 1.0e-3;

  int i = 0, err = 0;
  if (theta_out) {
    theta_out[0] = theta;
    rad_out[0] = rad;
  }
  double sdh, sarea;
  const double sixth = 1.0 / 6.0;
  while (fabs(theta - thetafinal) > 1.0e-3) {             /* :536 */
    if (zcenter >= zero) {                                /* :540-544 */
      if (theta + dth > thetafinal) dth = thetafinal - theta;
    } else {
      if (theta + dth < thetafinal) dth = thetafinal - theta;
    }
    const double dth2 = half * dth;
    double k_1, k_2, k_3, k_4, l_1, l_2, l_3, l_4, m_1, m_2, m_3, m_4;

    /* one large step (:550-575) */
    if ((err = ah_source(H, K, rad, drad, theta, zcenter, &sdh, &sarea))) break;
    k_1 = dth * drad;
    l_1 = dth * sdh;
    m_1 = fabs(dth2) * sarea;
    if ((err = ah_source(H, K, rad + 0.5 * k_1, drad + 0.5 * l_1, theta + 0.5 * dth, zcenter, &sdh, &sarea))) break;
    k_2 = dth * (drad + 0.5 * l_1);
    l_2 = dth * sdh;
    if ((err = ah_source(H, K, rad + 0.5 * k_2, drad + 0.5 * l_2, theta + 0.5 * dth, zcenter, &sdh, &sarea))) break;
    k_3 = dth * (drad + 0.5 * l_2);
    l_3 = dth * sdh;
    if ((err = ah_source(H, K, rad + k_3, drad + l_3, theta + dth, zcenter, &sdh, &sarea))) break;
    k_4 = dth * (drad + l_3);
    l_4 = dth * sdh;
    /* (rad_aux, :573, only enters rad_error, which nothing reads) */
    const double drad_aux = drad + sixth * (l_1 + 2.0 * l_2 + 2.0 * l_3 + l_4);

    /* first small step, reusing the first stage (:577-618) */
    k_1 = half * k_1;
    l_1 = half * l_1;
    if ((err = ah_source(H, K, rad + 0.5 * k_1, drad + 0.5 * l_1, theta + 0.5 * dth2, zcenter, &sdh, &sarea))) break;
    k_2 = dth2 * (drad + 0.5 * l_1);
    l_2 = dth2 * sdh;
    m_2 = fabs(dth2) * sarea;
    if ((err = ah_source(H, K, rad + 0.5 * k_2, drad + 0.5 * l_2, theta + 0.5 * dth2, zcenter, &sdh, &sarea))) break;
    k_3 = dth2 * (drad + 0.5 * l_2);
    l_3 = dth2 * sdh;
    m_3 = fabs(dth2) * sarea;
    if ((err = ah_source(H, K, rad + k_3, drad + l_3, theta + dth2, zcenter, &sdh, &sarea))) break;
    k_4 = dth2 * (drad + l_3);
    l_4 = dth2 * sdh;
    m_4 = fabs(dth2) * sarea;
    double rad1 = rad + sixth * (k_1 + 2.0 * k_2 + 2.0 * k_3 + k_4);
    double drad1 = drad + sixth * (l_1 + 2.0 * l_2 + 2.0 * l_3 + l_4);
    double area1 = area + sixth * (m_1 + 2.0 * m_2 + 2.0 * m_3 + m_4);

    /* second small step (:620-650) */
    if ((err = ah_source(H, K, rad1, drad1, theta + dth, zcenter, &sdh, &sarea))) break;
    k_1 = dth2 * drad1;
    l_1 = dth2 * sdh;
    m_1 = fabs(dth2) * sarea;
    if ((err = ah_source(H, K, rad1 + 0.5 * k_1, drad1 + 0.5 * l_1, theta + 1.5 * dth2, zcenter, &sdh, &sarea))) break;
    k_2 = dth2 * (drad1 + 0.5 * l_1);
    l_2 = dth2 * sdh;
    m_2 = fabs(dth2) * sarea;
    if ((err = ah_source(H, K, rad1 + 0.5 * k_2, drad1 + 0.5 * l_2, theta + 1.5 * dth2, zcenter, &sdh, &sarea))) break;
    k_3 = dth2 * (drad1 + 0.5 * l_2);
    l_3 = dth2 * sdh;
    m_3 = fabs(dth2) * sarea;
    if ((err = ah_source(H, K, rad1 + k_3, drad1 + l_3, theta + dth, zcenter, &sdh, &sarea))) break;
    k_4 = dth2 * (drad1 + l_3);
    l_4 = dth2 * sdh;
    m_4 = fabs(dth2) * sarea;
    const double rad2 = rad1 + sixth * (k_1 + 2.0 * k_2 + 2.0 * k_3 + k_4);
    const double drad2 = drad1 + sixth * (l_1 + 2.0 * l_2 + 2.0 * l_3 + l_4);
    const double area2 = area1 + sixth * (m_1 + 2.0 * m_2 + 2.0 * m_3 + m_4);

    /* error control (:652-700).  The reference's "local extrapolation"
     * (:672-677) writes the slot after the accepted one, which the next step
     * overwrites before reading: it has no effect and is not restated. */
    const double drad_error = fabs(drad2 - drad_aux);
    const double eps = ah_tol_rel * (fabs(drad2) + fabs(drad_aux));
    if (drad_error > fmax(eps, ah_tol_abs)) {
      dth = half * dth;
    } else {
      theta = theta + dth;
      rad = rad2;
      drad = drad2;
      area = area2;
      i = i + 1;
      if (theta_out) {
        theta_out[i] = theta;
        rad_out[i] = rad;
      }
      if (i >= AH_NTHETAMAX) {
        err = 1;
        break;
      }
      if (drad_error < 0.1 * eps) dth = two * dth;
    }
  }
  out.err = err;
  out.ntheta = i;               /* :704-705 */
  out.dradlast = drad;
  out.area = area;
  return out;
}

/* level of the axis point (0, zinit) for the march (:462-476) */
static double ah_dzlocal(const or_handle* H, double zinit) {
  int level = 0;
  for (int ll = 1; ll < H->nlev; ++ll) {
    const or_level* L = &H->L[ll];
    if (0.0 > L->rmin + L->dr && 0.0 < L->rmax - L->dr && zinit > L->zmin + L->dz && zinit < L->zmax - L->dz)
      level = ll;
  }
  return H->L[level].dz;
}

/* ahshoot (:246-1103) for horizon k centred at zcenter.  *found carries
 * ahfound(k) from the previous search (it is kept in one branch, :816-845). */
static void ah_shoot(const or_handle* H, const ah_curv* K, double zcenter, int* found, oa_horizon* out,
                     double* theta_out, double* rad_out) {
  const oa_params* p = &H->p;
  const or_level* L0 = &H->L[0];
  const double zero = 0.0, half = 0.5, two = 2.0;
  const double smallpi = acos(-1.0);
  const double eps_shoot = p->ahepsilon_shoot;

  double zinit_old = zero, zinit;                         /* :377-405 */
  if (p->ahzinit == 0.0) {
    zinit = (zcenter >= zero) ? 0.8 * L0->zmax : 0.8 * L0->zmin;
  } else if (p->ahzinit > 0.0 && p->ahzinit <= 0.9 * L0->zmax) {
    zinit = (zcenter >= zero) ? p->ahzinit : -p->ahzinit;
  } else {
    zinit = (zcenter >= zero) ? 0.8 * L0->zmax : 0.8 * L0->zmin;
  }
  double dzlocal = L0->dz;                                /* :409 */
  double dzz = 0.5 * dzlocal;
  double dradlast = 0.0, dradlast_old, drada = 0.0, za, zb;
  int errorflag = 0, shots = 0;
  ah_shot last = {0.0, 0.0, 0, 0};
  (void)zb;

  for (;;) {                                              /* 50 continue (:434) */
    int bracket = 0, firstiter = 1;
    errorflag = 0;
    za = 1.0e10;
    zb = -1.0e10;
    while (fabs(zinit - zcenter) > 5.0 * dzlocal && fabs(zinit_old - zinit) > eps_shoot) {   /* :454 */
      dzlocal = ah_dzlocal(H, zinit);
      dzz = 0.5 * dzlocal;
      dradlast_old = firstiter ? 1.0e10 : dradlast;       /* :490-494 */
      last = ah_integrate(H, K, zinit, zcenter, theta_out, rad_out);
      ++shots;
      errorflag = last.err;
      dradlast = last.dradlast;
      /* :739-792 */
      if (firstiter || (!bracket && dradlast * dradlast_old >= zero)) {
        firstiter = 0;
        zinit_old = zinit;
        zinit = (zcenter >= zero) ? zinit - dzz : zinit + dzz;
      } else if (!bracket) {
        bracket = 1;
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
    /* :797-855 */
    if (fabs(zinit - zinit_old) <= eps_shoot) {
      if (!errorflag && fabs(dradlast) < 1.0e-1) {
        *found = 1;
      } else if (fabs(zinit - zcenter) > dzlocal) {
        zinit = (zcenter >= zero) ? zinit - two * dzz : zinit + two * dzz;
        continue;                                         /* goto 50 (:842) */
      }
    } else {
      *found = 0;
    }
    break;
  }

  out->found = *found;
  out->ntheta = last.ntheta;
  out->shots = shots;
  out->errorflag = errorflag;
  out->zcenter = zcenter;
  out->zinit = zinit;
  out->dradlast = dradlast;
  double ah_area = zero;                                  /* :989-997 */
  if (*found) {
    ah_area = last.area;
    if (p->eqsym && zcenter == 0.0) ah_area = two * ah_area;
  }
  out->area = ah_area;
  out->radius = half * sqrt(ah_area / smallpi);
  out->mass = *found ? half * out->radius : zero;         /* :1065-1076 */
}

#define AH_MODES 50          /* ac(0:50), ack(0:2,0:50) (:1165-1168) */

/* ahshoot's projection of the surface of its last integration onto cosine
 * coefficients (:872-910), kept in ack(k,:) for the spectral finder:
 * trapezoids over the integration's uneven points, modes 0..20 (even ones
 * only with equatorial symmetry about a centre at 0); every other
 * coefficient is cleared. */
static void ah_shoot_coefficients(const oa_params* p, double zcenter, int n, const double* theta,
                                  const double* rad, double* ack) {
  const double half = 0.5, two = 2.0;
  const double smallpi = acos(-1.0);
  const int iaux = (p->eqsym && zcenter == 0.0) ? 2 : 1;
  for (int j = 0; j <= AH_MODES; ++j) ack[j] = 0.0;
  for (int j = 0; j <= 20; j += iaux) {
    for (int i = 0; i < n; ++i)
      ack[j] = ack[j] + half * (theta[i + 1] - theta[i]) *
                            (rad[i + 1] * cos((double)j * theta[i + 1]) + rad[i] * cos((double)j * theta[i]));
    ack[j] = two * (double)iaux * ack[j] / smallpi;
  }
}

/* h(theta_i) and h'(theta_i) from the cosine coefficients (:1320-1327,
 * :1602-1609) */
static void ah_reconstruct(int N, int Nmode, int step, const double* theta, const double* ac, double* rad,
                           double* drad) {
  for (int i = 0; i <= N; ++i) {
    double r = 0.5 * ac[0], d = 0.0;
    for (int j = step; j <= Nmode; j += step) {
      r = r + ac[j] * cos((double)j * theta[i]);
      d = d - (double)j * ac[j] * sin((double)j * theta[i]);
    }
    rad[i] = r;
    drad[i] = d;
  }
}

/* ahspectral (:1113-1972) for horizon k centred at zcenter: the fast flow
 * of the cosine coefficients of h(theta) (Gundlach, gr-qc/9707050), two
 * Runge-Kutta stages per iteration, with the step AA halved when the two
 * stages and one Euler step disagree and grown again (up to ahAA) when they
 * agree closely.  *found is ahfound(k) and ack its coefficients ack(k,:),
 * both shared with ahshoot; *firstcall is the finder's own first-call flag
 * of this horizon.  The surface of the last iteration goes to theta_out /
 * rad_out (Ntheta + 1 points), which ahshootfirst's search also uses as its
 * scratch. */
static void ah_spectral(const or_handle* H, const ah_curv* K, double zcenter, int* found, double* ack,
                        int* firstcall, oa_horizon* out, double* theta_out, double* rad_out) {
  const oa_params* p = &H->p;
  const or_level* L0 = &H->L[0];
  const double zero = 0.0, one = 1.0, two = 2.0, half = 0.5;
  const double smallpi = acos(-1.0);
  const int maxiter = p->ahmaxiter;
  const int Nmode = p->ahNmodes;                          /* :1221-1232 */
  const int N = 10 * Nmode;
  const int step = (p->eqsym && zcenter == 0.0) ? 2 : 1;  /* :1250-1258 */
  const double dtheta = smallpi / (double)(step * N);
  const double ahrmax = (p->ahrmax > 0.9) ? 0.9 : p->ahrmax;   /* :1273-1280 */
  double AA = p->ahAA;
  const double BB = p->ahBB;
  double ac[AH_MODES + 1], cc[AH_MODES + 1], ac_old[AH_MODES + 1], cc_old[AH_MODES + 1];
  double* theta = malloc(sizeof(double) * (size_t)(N + 1));
  double* rad = malloc(sizeof(double) * (size_t)(N + 1));
  double* drad = malloc(sizeof(double) * (size_t)(N + 1));
  double* rad_p = malloc(sizeof(double) * (size_t)(N + 1));
  double* drad_p = malloc(sizeof(double) * (size_t)(N + 1));
  double* Ifunc = malloc(sizeof(double) * (size_t)(N + 1));
  double* area = malloc(sizeof(double) * (size_t)(N + 1));
  for (int j = 0; j <= AH_MODES; ++j) ac[j] = cc[j] = ac_old[j] = cc_old[j] = zero;
  for (int i = 0; i <= N; ++i) {
    theta[i] = (double)i * dtheta;
    area[i] = zero;
    Ifunc[i] = zero;
  }
  memset(out, 0, sizeof *out);

  if (*firstcall && !*found)                              /* :1286-1289 */
    for (int j = 0; j <= AH_MODES; ++j) ack[j] = zero;
  if (*firstcall && p->ahshootfirst) {                    /* :1292-1303 */
    ah_shoot(H, K, zcenter, found, out, theta_out, rad_out);
    if (*found) ah_shoot_coefficients(p, zcenter, out->ntheta, theta_out, rad_out, ack);
    ack[0] = 1.05 * ack[0];
  }
  const int shots = out->shots;
  const double zinit = out->zinit, dradlast = out->dradlast;

  if (*found) {                                           /* :1311-1330 */
    for (int j = 0; j <= AH_MODES; ++j) ac[j] = ack[j];
    ah_reconstruct(N, Nmode, step, theta, ac, rad, drad);
  } else {                                                /* :1340-1356 */
    if (zcenter == 0.0) ac[0] = (p->ahzinit == 0.0) ? 1.8 * ahrmax * L0->zmax : 2.0 * p->ahzinit;
    else ac[0] = 1.95 * fabs(zcenter);
    for (int i = 0; i <= N; ++i) {
      rad[i] = half * ac[0];
      drad[i] = zero;
    }
    if (zcenter != 0.0 && p->aheccentric != 0.0) {        /* :1365-1407 */
      for (int i = 0; i <= N; ++i) {
        const double e = p->aheccentric * sin(theta[i]);
        rad[i] = half * ac[0] / sqrt(one - e * e);
      }
      for (int j = 0; j <= Nmode; ++j) {
        double a = half * dtheta * rad[0];
        for (int i = 1; i < N; ++i) a = a + dtheta * rad[i] * cos((double)j * theta[i]);
        a = a + half * dtheta * rad[N] * cos((double)j * theta[N]);
        ac[j] = two * (double)step * a / smallpi;
      }
      ah_reconstruct(N, Nmode, step, theta, ac, rad, drad);
    }
  }

  double res = 1.0;                                       /* :1448-1451 */
  int iter = 0, errorflag = 0;
  while (res > p->ahepsilon_spectral && iter < maxiter) {
    ++iter;
    for (int j = 0; j <= AH_MODES; ++j) {
      ac_old[j] = ac[j];
      cc_old[j] = cc[j];
    }
    memcpy(rad_p, rad, sizeof(double) * (size_t)(N + 1));
    memcpy(drad_p, drad, sizeof(double) * (size_t)(N + 1));
    for (int q = 1; q <= 2; ++q) {
      for (int i = 0; i <= N; ++i) {                      /* :1494-1501 */
        double sdh, sarea;
        errorflag = ah_source(H, K, rad[i], drad[i], theta[i], zcenter, &sdh, &sarea);
        if (errorflag) goto found_check;
        Ifunc[i] = sdh;
        area[i] = sarea;
      }
      Ifunc[0] = Ifunc[1];                                /* :1506-1515 */
      area[0] = zero;
      if (step == 1) {
        Ifunc[N] = Ifunc[N - 1];
        area[N] = zero;
      }
      for (int j = 0; j <= Nmode; j += step) {            /* :1539-1563 */
        double c = half * dtheta * Ifunc[0];
        for (int i = 1; i < N; ++i) c = c + dtheta * Ifunc[i] * cos((double)j * theta[i]);
        c = c + half * dtheta * Ifunc[N] * cos((double)j * theta[N]);
        c = two * (double)step * c / smallpi;
        c = c + ((double)j * (double)j) * ac[j];
        cc[j] = c;
        if (q == 1) cc_old[j] = c;
      }
      for (int j = 0; j <= Nmode; j += step)              /* :1590-1592 */
        ac[j] = ac_old[j] - half * AA * (double)q / (one + BB * ((double)j * (double)j)) * cc[j];
      ah_reconstruct(N, Nmode, step, theta, ac, rad, drad);
    }
    res = zero;                                           /* :1636-1640 */
    for (int j = 0; j <= Nmode; j += step) res = fmax(res, fabs(cc[j]));
    double aux1 = zero, aux2 = zero;                      /* :1651-1682 */
    for (int j = 0; j <= Nmode; j += step) {
      const double tc = ac_old[j] - AA / (one + BB * ((double)j * (double)j)) * cc_old[j];
      aux1 = aux1 + fabs(ac[j] - tc);
      aux2 = aux2 + fabs(ac_old[j]);
    }
    if (aux1 > 0.05 * aux2) {
      AA = half * AA;
      for (int j = 0; j <= Nmode; j += step) ac[j] = ac_old[j];
      memcpy(rad, rad_p, sizeof(double) * (size_t)(N + 1));
      memcpy(drad, drad_p, sizeof(double) * (size_t)(N + 1));
    } else if (aux1 < 0.001 * aux2) {
      AA = fmin(1.5 * AA, p->ahAA);
    }
  }

found_check:                                              /* :1700-1733 */
  if (ac[0] < p->dr0 + p->dz0) errorflag = 1;
  if (iter < maxiter && !errorflag) {
    *found = 1;
    for (int j = 0; j <= AH_MODES; ++j) ack[j] = ac[j];
  } else {
    *found = 0;
  }
  *firstcall = 0;                                         /* :1748-1754 */

  double ah_area = zero;                                  /* :1826-1843 */
  if (*found) {
    ah_area = half * dtheta * area[0];
    for (int i = 1; i < N; ++i) ah_area = ah_area + dtheta * area[i];
    ah_area = ah_area + half * dtheta * area[N];
    if (p->eqsym && zcenter == 0.0) ah_area = two * ah_area;
  }
  out->found = *found;
  out->ntheta = N;
  out->shots = shots;
  out->errorflag = errorflag;
  out->zcenter = zcenter;
  out->zinit = zinit;
  out->dradlast = dradlast;
  out->area = ah_area;
  out->radius = half * sqrt(ah_area / smallpi);
  out->mass = *found ? half * out->radius : zero;         /* :1935-1945 */
  out->method = p->ahmethod;
  out->iterations = iter;
  out->residual = res;
  if (theta_out)
    for (int i = 0; i <= N; ++i) {                        /* ah_radius_SP (:1766-1790) */
      theta_out[i] = theta[i];
      rad_out[i] = *found ? rad[i] : zero;
    }
  free(theta);
  free(rad);
  free(drad);
  free(rad_p);
  free(drad_p);
  free(Ifunc);
  free(area);
}

/* horizon_finder (:2-237) for horizon k: sanity checks, the centre, then
 * the method's search (:209-227).  *found and ack (AH_MODES + 1 values)
 * are ahfound(k) and ack(k,:), *firstcall the spectral finder's first-call
 * flag; all three carry over between calls.  KA..KH are formed here from
 * the auxiliary arrays.  Returns 0 or an OA_ERR_* code. */
int or_ahfind(or_handle* H, int k, int* found, double* ack, int* firstcall, oa_horizon* out, double* theta_out,
              double* rad_out) {
  const oa_params* p = &H->p;
  if (p->N_ah > 3 || p->N_ah < 1 || k < 0 || k >= p->N_ah) return OA_ERR_ARG;   /* :48-57 */
  if (p->eqsym && p->N_ah == 3) return OA_ERR_PARAM;                              /* :62-71 */
  if (p->z_ah1 < 0.0 || p->z_ah2 > 0.0) return OA_ERR_PARAM;                      /* :76-95 */
  if (p->ahmethod != OA_AH_SHOOT && (p->ahNmodes < 1 || p->ahNmodes > AH_MODES)) return OA_ERR_PARAM;
  double zah[3] = {0.0, 0.0, 0.0};                                                 /* :104-111 */
  if (p->N_ah > 1) {
    zah[1] = p->z_ah1;
    zah[2] = p->z_ah2;
  }
  /* moving centres (:117-180): from z_ah(k), step along the axis by dz0
   * towards larger phi until Dz_phi changes sign, then place the centre at
   * the linear zero of Dz_phi.  The level is the finest one holding the
   * starting point.  A walk that leaves that level (the reference then stops
   * in interp) returns OA_ERR_STATE. */
  if (p->ahmove && p->N_ah > 1 && k > 0) {
    const double r0 = 0.0;
    double z0 = zah[k], z0_old = 0.0;
    const int level = ah_level(H, r0, z0);
    const or_level* L = &H->L[level];
    double Dzphi = 0.0, Dzphi_old = 0.0;
    int bracket = 0;
    for (int it = 0; !bracket; ++it) {
      if (it > 1000000 || z0 < L->zmin || z0 > L->zmax) return OA_ERR_STATE;
      Dzphi = or_interp(p, L, L->Dz_phi, r0, z0);
      if (Dzphi * Dzphi_old >= 0.0) {
        z0_old = z0;
        Dzphi_old = Dzphi;
        z0 = (Dzphi > 0.0) ? z0 + p->dz0 : z0 - p->dz0;
      } else {
        bracket = 1;
        zah[k] = z0_old - Dzphi_old * (z0 - z0_old) / (Dzphi - Dzphi_old);
      }
    }
  }
  double *th_scratch = NULL, *rad_scratch = NULL;
  if (!theta_out || !rad_out) {       /* the spectral paths keep the shooting surface here */
    th_scratch = malloc(sizeof(double) * (AH_NTHETAMAX + 1));
    rad_scratch = malloc(sizeof(double) * (AH_NTHETAMAX + 1));
    theta_out = th_scratch;
    rad_out = rad_scratch;
  }
  ah_curv* K = ah_curv_new(H);
  if (p->ahmethod == OA_AH_SHOOT) {
    ah_shoot(H, K, zah[k], found, out, theta_out, rad_out);
  } else if (p->ahmethod == OA_AH_SPECTRAL) {
    ah_spectral(H, K, zah[k], found, ack, firstcall, out, theta_out, rad_out);
  } else {
    oa_horizon shot;
    memset(&shot, 0, sizeof shot);
    ah_shoot(H, K, zah[k], found, &shot, theta_out, rad_out);
    if (*found) ah_shoot_coefficients(p, zah[k], shot.ntheta, theta_out, rad_out, ack);
    ah_spectral(H, K, zah[k], found, ack, firstcall, out, theta_out, rad_out);
    out->shots = shot.shots;
    out->zinit = shot.zinit;
    out->dradlast = shot.dradlast;
  }
  ah_curv_free(H, K);
  free(th_scratch);
  free(rad_scratch);
  return 0;
}

/* One integration from zinit (for the parity tests of single shots). */
int or_ah_integrate(or_handle* H, double zinit, double zcenter, double* res) {
  ah_curv* K = ah_curv_new(H);
  const ah_shot s = ah_integrate(H, K, zinit, zcenter, NULL, NULL);
  res[0] = s.dradlast;
  res[1] = s.area;
  res[2] = (double)s.err;
  res[3] = (double)s.ntheta;
  ah_curv_free(H, K);
  return 0;
}

/* or_ahfind with the result as thirteen doubles (found, ntheta, shots,
 * errorflag, zcenter, zinit, dradlast, area, radius, mass, method,
 * iterations, residual), for the ctypes tests. */
int or_ahfind_flat(or_handle* H, int k, int* found, double* ack, int* firstcall, double* res, double* theta_out,
                   double* rad_out) {
  oa_horizon o;
  memset(&o, 0, sizeof o);
  const int rc = or_ahfind(H, k, found, ack, firstcall, &o, theta_out, rad_out);
  res[0] = o.found;
  res[1] = o.ntheta;
  res[2] = o.shots;
  res[3] = o.errorflag;
  res[4] = o.zcenter;
  res[5] = o.zinit;
  res[6] = o.dradlast;
  res[7] = o.area;
  res[8] = o.radius;
  res[9] = o.mass;
  res[10] = o.method;
  res[11] = o.iterations;
  res[12] = o.residual;
  return rc;
}
