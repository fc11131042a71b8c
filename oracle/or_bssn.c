/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see or_internal.h).
 *
 * auxiliary_geometry / sources_geometry / sources_lapse / sources_shift /
 * radiative_geometry of the reference, restated for the branches this hot path
 * supports: angmom = .false., bssnflavor = lagrangian, formulation = bssn,
 * nolambda = .true., evolveH = .true., evolveKTH = .false., factorhdet = .true.
 * Consecutive pointwise reference statements are grouped in one loop; that
 * keeps the per-point evaluation order of the reference.
 */
#include <math.h>
#include <string.h>
#include "or_internal.h"

#define EV(f) L->var[OF_##f]
#define SRC(f) L->src[OF_##f]

static int has_shift(const oa_params* p) { return p->shift != OA_SHIFT_NONE; }

/* derivative helpers honouring ext_derivs (pointwise tests supply them) */
#define D1R(f, s, o) do { if (!L->ext_derivs) or_diff1r(p, L, f, s, o); } while (0)
#define D1Z(f, s, o) do { if (!L->ext_derivs) or_diff1z(p, L, f, s, o); } while (0)
#define D2R(f, s, o) do { if (!L->ext_derivs) or_diff2r(p, L, f, s, o); } while (0)
#define D2Z(f, s, o) do { if (!L->ext_derivs) or_diff2z(p, L, f, s, o); } while (0)
#define D2RZ(f, sr, sz, o) do { if (!L->ext_derivs) or_diff2rz(p, L, f, sr, sz, o); } while (0)

/* advective derivatives: shiftadvect selects upwinded (derivadvect.f90) or
 * plain copies (auxiliary_geometry.f90:64-72 and its repeats) */
static void adv_pair(const oa_params* p, or_level* L, const double* f, int symr, int symz,
                     const double* Dr, const double* Dz, double* DAr, double* DAz) {
  if (L->ext_derivs) return;
  if (p->shiftadvect) {
    or_diffadvr(p, L, f, symr, EV(beta_r), DAr);
    or_diffadvz(p, L, f, symz, EV(beta_z), DAz);
  } else {
    memcpy(DAr, Dr, sizeof(double) * L->n);
    memcpy(DAz, Dz, sizeof(double) * L->n);
  }
}

/* factordet: factordet.f90:40-64 (non-angmom) */
static void factordet(or_level* L) {
  double *A = EV(A), *B = EV(B), *C = EV(C), *H = EV(H);
  ALL(L, k) {
    const double r = L->r[k];
    double aux = (A[k] * B[k] - (r * C[k]) * (r * C[k])) * H[k];
    aux = aux / L->hdet[k];
    aux = pow(aux, 1.0 / 3.0);
    A[k] = A[k] / aux;
    B[k] = B[k] / aux;
    C[k] = C[k] / aux;
    H[k] = H[k] / aux;
  }
}

/* auxiliary_geometry: auxiliary_geometry.f90:55-1502 (non-angmom) */
void or_auxiliary_geometry(const oa_params* p, or_level* L) {
  const int time0 = (L->t == 0.0);
  const int shift = has_shift(p);
  const double half = 0.5, two = 2.0, third = 1.0 / 3.0, four = 4.0, fourth = 0.25;
  const double cp = (double)p->chipower;

  /* lapse derivatives (:55-72) */
  D1R(EV(alpha), +1, L->Dr_alpha);
  D1Z(EV(alpha), +1, L->Dz_alpha);
  D2R(EV(alpha), +1, L->Drr_alpha);
  D2Z(EV(alpha), +1, L->Dzz_alpha);
  D2RZ(EV(alpha), +1, +1, L->Drz_alpha);
  if (shift) adv_pair(p, L, EV(alpha), +1, +1, L->Dr_alpha, L->Dz_alpha, L->DAr_alpha, L->DAz_alpha);

  /* Bona-Masso function falpha = alpha^2 f(alpha) (:116-168) */
  ALL(L, k) {
    const double a = EV(alpha)[k];
    switch (p->slicing) {
      case OA_SLICING_STATIC: case OA_SLICING_MAXIMAL: L->falpha[k] = 0.0; break;
      case OA_SLICING_HARMONIC: case OA_SLICING_HARMONIC2: L->falpha[k] = a * a * p->gauge_f; break;
      case OA_SLICING_1PLUSLOG: case OA_SLICING_1PLUSLOG2: L->falpha[k] = a * p->gauge_f; break;
      default: L->falpha[k] = a * a + (p->gauge_f - 1.0); break;
    }
  }

  /* conformal factor (:178-236) */
  if (!p->chimethod || time0) {
    ALL(L, k) EV(chi)[k] = exp(-cp * EV(phi)[k]);
    D1R(EV(phi), +1, L->Dr_phi);
    D1Z(EV(phi), +1, L->Dz_phi);
    D2R(EV(phi), +1, L->Drr_phi);
    D2Z(EV(phi), +1, L->Dzz_phi);
    D2RZ(EV(phi), +1, +1, L->Drz_phi);
    if (shift) adv_pair(p, L, EV(phi), +1, +1, L->Dr_phi, L->Dz_phi, L->DAr_phi, L->DAz_phi);
  } else {
    ALL(L, k) EV(phi)[k] = -log(EV(chi)[k]) / cp;
    D1R(EV(chi), +1, L->Dr_chi);
    D1Z(EV(chi), +1, L->Dz_chi);
    D2R(EV(chi), +1, L->Drr_chi);
    D2Z(EV(chi), +1, L->Dzz_chi);
    D2RZ(EV(chi), +1, +1, L->Drz_chi);
    ALL(L, k) {
      const double chi = EV(chi)[k];
      const double dr = L->Dr_chi[k], dz = L->Dz_chi[k];
      L->Dr_phi[k] = -dr / chi / cp;
      L->Dz_phi[k] = -dz / chi / cp;
      L->Drr_phi[k] = -(L->Drr_chi[k] / chi - (dr / chi) * (dr / chi)) / cp;
      L->Dzz_phi[k] = -(L->Dzz_chi[k] / chi - (dz / chi) * (dz / chi)) / cp;
      L->Drz_phi[k] = -(L->Drz_chi[k] / chi - dr * dz / (chi * chi)) / cp;
    }
    if (shift) {
      adv_pair(p, L, EV(chi), +1, +1, L->Dr_chi, L->Dz_chi, L->DAr_chi, L->DAz_chi);
      ALL(L, k) {
        L->DAr_phi[k] = -L->DAr_chi[k] / EV(chi)[k] / cp;
        L->DAz_phi[k] = -L->DAz_chi[k] / EV(chi)[k] / cp;
      }
    }
  }
  ALL(L, k) {
    const double ps = exp(EV(phi)[k]);
    L->psi[k] = ps;
    L->psi2[k] = ps * ps;
    L->psi4[k] = ps * ps * ps * ps;
  }

  /* determinant of the conformal metric (:265-318): factorhdet at t=0 fixes
   * hdet=1; later lagrangian steps rescale with factordet. */
  if (time0) {
    ALL(L, k) { L->hdet[k] = 1.0; L->ihdet[k] = 1.0; L->Dr_hdet[k] = 0.0; L->Dz_hdet[k] = 0.0; }
  } else {
    factordet(L);
  }

  /* conformal metric derivatives (:325-399) */
  D1R(EV(A), +1, L->Dr_A); D1Z(EV(A), +1, L->Dz_A); D2R(EV(A), +1, L->Drr_A);
  D2Z(EV(A), +1, L->Dzz_A); D2RZ(EV(A), +1, +1, L->Drz_A);
  if (shift) adv_pair(p, L, EV(A), +1, +1, L->Dr_A, L->Dz_A, L->DAr_A, L->DAz_A);
  D1R(EV(B), +1, L->Dr_B); D1Z(EV(B), +1, L->Dz_B); D2R(EV(B), +1, L->Drr_B);
  D2Z(EV(B), +1, L->Dzz_B); D2RZ(EV(B), +1, +1, L->Drz_B);
  if (shift) adv_pair(p, L, EV(B), +1, +1, L->Dr_B, L->Dz_B, L->DAr_B, L->DAz_B);
  D1R(EV(H), +1, L->Dr_H); D1Z(EV(H), +1, L->Dz_H); D2R(EV(H), +1, L->Drr_H);
  D2Z(EV(H), +1, L->Dzz_H); D2RZ(EV(H), +1, +1, L->Drz_H);
  if (shift) adv_pair(p, L, EV(H), +1, +1, L->Dr_H, L->Dz_H, L->DAr_H, L->DAz_H);
  D1R(EV(C), +1, L->Dr_C); D1Z(EV(C), -1, L->Dz_C); D2R(EV(C), +1, L->Drr_C);
  D2Z(EV(C), -1, L->Dzz_C); D2RZ(EV(C), +1, -1, L->Drz_C);
  if (shift) adv_pair(p, L, EV(C), +1, -1, L->Dr_C, L->Dz_C, L->DAr_C, L->DAz_C);

  /* inverse conformal metric (:484-487) and its derivatives (:493-507) */
  ALL(L, k) {
    const double A = EV(A)[k], B = EV(B)[k], H = EV(H)[k], C = EV(C)[k], ih = L->ihdet[k];
    L->g_A[k] = B * H * ih;
    L->g_B[k] = A * H * ih;
    L->g_H[k] = 1.0 / H;
    L->g_C[k] = -C * H * ih;
  }
  D1R(L->g_A, +1, L->Dr_g_A); D1Z(L->g_A, +1, L->Dz_g_A);
  D1R(L->g_B, +1, L->Dr_g_B); D1Z(L->g_B, +1, L->Dz_g_B);
  D1R(L->g_H, +1, L->Dr_g_H); D1Z(L->g_H, +1, L->Dz_g_H);
  D1R(L->g_C, +1, L->Dr_g_C); D1Z(L->g_C, -1, L->Dz_g_C);

  /* shift (:587-737) */
  if (shift) {
    D1R(EV(beta_r), -1, L->Dr_beta_r); D1Z(EV(beta_r), +1, L->Dz_beta_r);
    D2R(EV(beta_r), -1, L->Drr_beta_r); D2Z(EV(beta_r), +1, L->Dzz_beta_r);
    D2RZ(EV(beta_r), -1, +1, L->Drz_beta_r);
    adv_pair(p, L, EV(beta_r), -1, +1, L->Dr_beta_r, L->Dz_beta_r, L->DAr_beta_r, L->DAz_beta_r);
    D1R(EV(beta_z), +1, L->Dr_beta_z); D1Z(EV(beta_z), -1, L->Dz_beta_z);
    D2R(EV(beta_z), +1, L->Drr_beta_z); D2Z(EV(beta_z), -1, L->Dzz_beta_z);
    D2RZ(EV(beta_z), +1, -1, L->Drz_beta_z);
    adv_pair(p, L, EV(beta_z), +1, -1, L->Dr_beta_z, L->Dz_beta_z, L->DAr_beta_z, L->DAz_beta_z);
    ALL(L, k) L->auxarray[k] = EV(beta_r)[k] / L->r[k];
    D1R(L->auxarray, +1, L->DD_beta_rr);
    D1R(EV(dtbeta_r), -1, L->Dr_dtbeta_r); D1Z(EV(dtbeta_r), +1, L->Dz_dtbeta_r);
    adv_pair(p, L, EV(dtbeta_r), -1, +1, L->Dr_dtbeta_r, L->Dz_dtbeta_r, L->DAr_dtbeta_r, L->DAz_dtbeta_r);
    D1R(EV(dtbeta_z), +1, L->Dr_dtbeta_z); D1Z(EV(dtbeta_z), -1, L->Dz_dtbeta_z);
    adv_pair(p, L, EV(dtbeta_z), +1, -1, L->Dr_dtbeta_z, L->Dz_dtbeta_z, L->DAr_dtbeta_z, L->DAz_dtbeta_z);
    ALL(L, k) {
      const double r = L->r[k];
      L->DIV_beta[k] = L->Dr_beta_r[k] + L->Dz_beta_z[k] + EV(beta_r)[k] / r;
      L->Dr_DIV_beta[k] = L->Drr_beta_r[k] + L->Drz_beta_z[k] + L->DD_beta_rr[k];
      L->Dz_DIV_beta[k] = L->Drz_beta_r[k] + L->Dzz_beta_z[k] + L->Dz_beta_r[k] / r;
    }
  }

  /* KTH from tracelessness (:747-756), lambdas (:767-821) */
  ALL(L, k) {
    const double r = L->r[k];
    EV(KTH)[k] = -(L->g_A[k] * EV(KTA)[k] + L->g_B[k] * EV(KTB)[k]
                   + two * r * r * L->g_C[k] * EV(KTC)[k]) / L->g_H[k];
    L->lamDef[k] = (EV(A)[k] - EV(H)[k]) / (r * r);
    L->AlamDef[k] = (EV(KTA)[k] - EV(KTH)[k]) / (r * r);
    EV(lambda)[k] = L->lamDef[k];         /* nolambda */
    EV(Alambda)[k] = L->AlamDef[k];
    L->g_lamDef[k] = (L->g_A[k] - L->g_H[k]) / (r * r);
    L->g_lambda[k] = (EV(C)[k] * EV(C)[k] - EV(B)[k] * EV(lambda)[k]) * L->ihdet[k];
  }

  /* extrinsic curvature derivatives (:830-906) */
  D1R(EV(trK), +1, L->Dr_trK); D1Z(EV(trK), +1, L->Dz_trK);
  if (shift) adv_pair(p, L, EV(trK), +1, +1, L->Dr_trK, L->Dz_trK, L->DAr_trK, L->DAz_trK);
  D1R(EV(KTA), +1, L->Dr_KTA); D1Z(EV(KTA), +1, L->Dz_KTA);
  if (shift) adv_pair(p, L, EV(KTA), +1, +1, L->Dr_KTA, L->Dz_KTA, L->DAr_KTA, L->DAz_KTA);
  D1R(EV(KTB), +1, L->Dr_KTB); D1Z(EV(KTB), +1, L->Dz_KTB);
  if (shift) adv_pair(p, L, EV(KTB), +1, +1, L->Dr_KTB, L->Dz_KTB, L->DAr_KTB, L->DAz_KTB);
  D1R(EV(KTC), +1, L->Dr_KTC); D1Z(EV(KTC), -1, L->Dz_KTC);
  if (shift) adv_pair(p, L, EV(KTC), +1, -1, L->Dr_KTC, L->Dz_KTC, L->DAr_KTC, L->DAz_KTC);

  /* contravariant KT, KT2, K2 (:970-973, 1017-1031, 1050) */
  ALL(L, k) {
    const double r = L->r[k], r2 = r * r;
    const double gA = L->g_A[k], gB = L->g_B[k], gH = L->g_H[k], gC = L->g_C[k];
    const double KTA = EV(KTA)[k], KTB = EV(KTB)[k], KTC = EV(KTC)[k], KTH = EV(KTH)[k];
    L->KT_up_rr[k] = (2.0 * gA * gC * KTC + gC * gC * KTB) * r2 + gA * gA * KTA;
    L->KT_up_zz[k] = (gC * gC * KTA + 2.0 * gC * gB * KTC) * r2 + gB * gB * KTB;
    L->KT_up_rz[k] = gC * gC * (r2 * r) * KTC + (gC * gA * KTA + gB * gA * KTC + gB * gC * KTB) * r;
    L->KT2_A[k] = (2.0 * KTA * gC * KTC + KTC * KTC * gB) * r2 + gA * KTA * KTA;
    L->KT2_B[k] = (KTC * KTC * gA + 2.0 * KTC * gC * KTB) * r2 + gB * KTB * KTB;
    L->KT2_H[k] = KTH * KTH * gH;
    L->KT2_C[k] = KTC * KTC * gC * r2 + KTA * gA * KTC + KTA * gC * KTB + KTC * gB * KTB;
    L->KT2[k] = gA * L->KT2_A[k] + gB * L->KT2_B[k] + gH * L->KT2_H[k] + two * gC * L->KT2_C[k] * r2;
    const double tk = EV(trK)[k] + two * EV(z4theta)[k];
    L->K2[k] = L->KT2[k] + third * tk * tk;
  }

  /* BSSN Delta (:1085-1114) */
  ALL(L, k) {
    const double r = L->r[k], ih = L->ihdet[k];
    L->DelDef_r[k] = -L->Dr_g_A[k] - r * L->Dz_g_C[k] - r * L->g_lambda[k]
                     - half * (L->g_A[k] * L->Dr_hdet[k] + r * L->g_C[k] * L->Dz_hdet[k]) * ih;
    L->DelDef_z[k] = -L->Dr_g_C[k] * r - L->Dz_g_B[k] - two * L->g_C[k]
                     - half * (r * L->g_C[k] * L->Dr_hdet[k] + L->g_B[k] * L->Dz_hdet[k]) * ih;
    if (p->noDelta_r || time0) EV(Delta_r)[k] = L->DelDef_r[k];
    if (p->noDelta_z || time0) EV(Delta_z)[k] = L->DelDef_z[k];
  }
  D1R(EV(Delta_r), -1, L->Dr_Delta_r); D1Z(EV(Delta_r), +1, L->Dz_Delta_r);
  if (shift) adv_pair(p, L, EV(Delta_r), -1, +1, L->Dr_Delta_r, L->Dz_Delta_r, L->DAr_Delta_r, L->DAz_Delta_r);
  D1R(EV(Delta_z), +1, L->Dr_Delta_z); D1Z(EV(Delta_z), -1, L->Dz_Delta_z);
  if (shift) adv_pair(p, L, EV(Delta_z), +1, -1, L->Dr_Delta_z, L->Dz_Delta_z, L->DAr_Delta_z, L->DAz_Delta_z);

  /* second covariant derivative of the lapse, Laplacian (:1268-1324) */
  ALL(L, k) {
    const double r = L->r[k], r2 = r * r;
    const double gA = L->g_A[k], gB = L->g_B[k], gH = L->g_H[k], gC = L->g_C[k];
    const double A = EV(A)[k], B = EV(B)[k], H = EV(H)[k], C = EV(C)[k];
    const double Dra = L->Dr_alpha[k], Dza = L->Dz_alpha[k];
    const double Dr_A = L->Dr_A[k], Dz_A = L->Dz_A[k], Dr_B = L->Dr_B[k], Dz_B = L->Dz_B[k];
    const double Dr_H = L->Dr_H[k], Dz_H = L->Dz_H[k], Dr_C = L->Dr_C[k], Dz_C = L->Dz_C[k];
    const double Drp = L->Dr_phi[k], Dzp = L->Dz_phi[k];
    double dA = L->Drr_alpha[k] - Dra * gC * r2 * Dr_C
                - (Dza * gB * Dr_C + half * Dza * gC * Dr_A + Dra * gC * C - half * Dra * gC * Dz_A) * r
                - half * Dra * gA * Dr_A - Dza * gB * C + half * Dza * gB * Dz_A;
    double dB = L->Dzz_alpha[k] - Dza * gC * r2 * Dz_C
                - (Dra * gA * Dz_C + half * Dra * gC * Dz_B - half * Dza * gC * Dr_B) * r
                + half * Dra * gA * Dr_B - half * Dza * gB * Dz_B;
    double dH = -(-half * Dra * gC * Dz_H - half * Dza * gC * Dr_H) * r
                + half * Dra * gA * Dr_H + Dza * gC * H + half * Dza * gB * Dz_H + Dra * gA * H / r;
    double dC = L->Drz_alpha[k] / r - half * Dra * gC * Dr_B - half * Dza * gC * Dz_A
                - (half * Dra * gA * Dz_A + half * Dza * gB * Dr_B) / r;
    const double aux = gA * Dra * Drp + gB * Dza * Dzp + r * gC * (Dra * Dzp + Dza * Drp);
    dA = dA - two * (two * Dra * Drp - A * aux);
    dB = dB - two * (two * Dza * Dzp - B * aux);
    dH = dH + two * H * aux;
    dC = dC - two * ((Dra / r) * Dzp + Dza * (Drp / r) - C * aux);
    L->D2cov_alpha_A[k] = dA;
    L->D2cov_alpha_B[k] = dB;
    L->D2cov_alpha_H[k] = dH;
    L->D2cov_alpha_C[k] = dC;
    double lap = gA * dA + gB * dB + gH * dH + two * r2 * gC * dC;
    L->Lapla_alpha[k] = lap / L->psi4[k];
  }

  /* Ricci tensor: conformal-factor part (:1397-1469) and the four conformal
   * parts (calc_conformalRicci1..4.f90, non-angmom branches), Ricci scalar
   * (:1487-1493) */
  ALL(L, k) {
    const double r = L->r[k], r2 = r * r, r3 = r2 * r;
    const double gA = L->g_A[k], gB = L->g_B[k], gH = L->g_H[k], gC = L->g_C[k];
    const double A = EV(A)[k], B = EV(B)[k], H = EV(H)[k], C = EV(C)[k];
    const double Dr_A = L->Dr_A[k], Dz_A = L->Dz_A[k], Dr_B = L->Dr_B[k], Dz_B = L->Dz_B[k];
    const double Dr_H = L->Dr_H[k], Dz_H = L->Dz_H[k], Dr_C = L->Dr_C[k], Dz_C = L->Dz_C[k];
    const double Drp = L->Dr_phi[k], Dzp = L->Dz_phi[k];
    const double lamDef = L->lamDef[k], lambda = EV(lambda)[k], g_lambda = L->g_lambda[k];
    const double DelDef_r = L->DelDef_r[k], DelDef_z = L->DelDef_z[k];
    const double Delta_r = EV(Delta_r)[k];
    const double Dr_Delta_r = L->Dr_Delta_r[k], Dz_Delta_r = L->Dz_Delta_r[k];
    const double Dr_Delta_z = L->Dr_Delta_z[k], Dz_Delta_z = L->Dz_Delta_z[k];
    const double gA2 = gA * gA, gB2 = gB * gB, gH2 = gH * gH, gC2 = gC * gC;

    /* I) second covariant derivatives of phi */
    double RA = -two * (L->Drr_phi[k] - Drp * gC * r2 * Dr_C
                        - (Dzp * gB * Dr_C + half * Dzp * gC * Dr_A + Drp * gC * C - half * Drp * gC * Dz_A) * r
                        - half * Drp * gA * Dr_A - Dzp * gB * C + half * Dzp * gB * Dz_A);
    double RB = -two * (L->Dzz_phi[k] - Dzp * gC * r2 * Dz_C
                        - (Drp * gA * Dz_C + half * Drp * gC * Dz_B - half * Dzp * gC * Dr_B) * r
                        + half * Drp * gA * Dr_B - half * Dzp * gB * Dz_B);
    double RH = -two * (-(-half * Drp * gC * Dz_H - half * Dzp * gC * Dr_H) * r
                        + half * Drp * gA * Dr_H + Dzp * gC * H + half * Dzp * gB * Dz_H + Drp * gA * H / r);
    double RC = -two * (L->Drz_phi[k] / r - half * Drp * gC * Dr_B - half * Dzp * gC * Dz_A
                        - (half * Drp * gA * Dz_A + half * Dzp * gB * Dr_B) / r);
    /* II) Laplacian of phi */
    double aux = gA * RA + gB * RB + gH * RH + two * r2 * gC * RC;
    RA = RA + A * aux;
    RB = RB + B * aux;
    RH = RH + H * aux;
    RC = RC + C * aux;
    /* III) quadratic terms */
    RA = RA + four * (Drp * Drp);
    RB = RB + four * (Dzp * Dzp);
    RC = RC + four * (Drp * Dzp) / r;
    /* IV) contracted quadratic terms */
    aux = -two * (gA * Drp * Drp + gB * Dzp * Dzp + two * r * gC * Drp * Dzp);
    RA = RA + two * A * aux;
    RB = RB + two * B * aux;
    RH = RH + two * H * aux;
    RC = RC + two * C * aux;

    /* calc_conformalRicci1.f90:76-86 */
    RA = RA - gC * L->Drz_A[k] * r - half * gB * L->Dzz_A[k] - half * gA * L->Drr_A[k]
         - half * gH * Dr_A / r + gH * lamDef;
    RB = RB - gC * L->Drz_B[k] * r - half * gA * L->Drr_B[k] - half * gB * L->Dzz_B[k]
         - half * gH * Dr_B / r;
    RH = RH - gC * L->Drz_H[k] * r - half * gB * L->Dzz_H[k] - half * gA * L->Drr_H[k]
         - half * gH * Dr_H / r - gH * lamDef;
    RC = RC - gC * L->Drz_C[k] * r - half * gA * L->Drr_C[k] - gC * Dz_C - half * gB * L->Dzz_C[k]
         + (-gA * Dr_C - half * gH * Dr_C) / r;

    /* calc_conformalRicci2.f90:173-183 */
    RA = RA + C * Dr_Delta_z * r + A * Dr_Delta_r + half * (DelDef_r * Dr_A + DelDef_z * Dz_A);
    RB = RB + C * Dz_Delta_r * r + B * Dz_Delta_z + half * (DelDef_r * Dr_B + DelDef_z * Dz_B);
    RH = RH + half * (DelDef_r * Dr_H + DelDef_z * Dz_H) + Delta_r * H / r;
    RC = RC + half * C * (Dr_Delta_r + Dz_Delta_z) + half * (DelDef_r * Dr_C + DelDef_z * Dz_C)
         + half * (C * Delta_r + A * Dz_Delta_r + B * Dr_Delta_z) / r;

    /* calc_conformalRicci3.f90:316-343 */
    RA = RA + (2.0 * gB * gC * Dr_C * Dz_C + gC2 * Dr_A * Dz_C + Dz_A * gC2 * Dr_C) * r3
         + (2.0 * gB * gC * C * Dz_C - gB * gC * Dr_C * Dr_B + Dr_A * gC * gA * Dr_C + Dz_A * gC2 * C) * r2
         + (Dz_A * gB * gC * Dr_B + Dz_A * gB * gA * Dr_C + Dr_A * gC * gA * C - gB * gC * C * Dr_B
            + Dz_A * gC * gA * Dr_A + gB2 * Dr_B * Dz_C) * r
         - half * gH2 * Dr_H * Dr_H - half * gB2 * Dr_B * Dr_B + gB * gA * C * Dz_A
         + half * Dr_A * Dr_A * gA2 + (gH * Dr_H + gH2 * A * Dr_H - 2.0 * gH2 * Dr_H * H) / r;
    RB = RB + (Dr_B * gC2 * Dz_C + gC2 * Dz_B * Dr_C + 2.0 * gA * gC * Dz_C * Dr_C) * r3
         + (2.0 * gA * gC * Dz_C * C - gA * gC * Dz_C * Dz_A + gC2 * Dz_B * C + Dz_B * gC * gB * Dz_C) * r2
         + (Dr_B * gA * gC * Dz_A + gA2 * Dz_A * Dr_C + Dr_B * gA * gB * Dz_C + Dr_B * gB * gC * Dz_B) * r
         + half * gB2 * Dz_B * Dz_B + gA2 * Dz_A * C - half * gA2 * Dz_A * Dz_A + gH2 * Dz_H * C
         - half * gH2 * Dz_H * Dz_H;
    RH = RH - gC * Dz_H - r * g_lambda * Dr_H;
    RC = RC + (gC2 * Dr_B * Dr_C + gC * gB * Dz_C * Dz_C + Dz_A * gC2 * Dz_C + gC * gA * Dr_C * Dr_C) * r2
         + (-half * gC * gA * Dr_C * Dz_A + 2.0 * gC * gA * Dr_C * C - half * gC2 * Dr_B * Dz_A
            - half * gC * gB * Dz_C * Dr_B + half * Dz_B * gB * gC * Dr_C + gC2 * Dr_B * C
            + half * Dz_B * gC2 * Dr_A + half * Dr_A * gA * gC * Dz_C) * r
         + half * Dr_B * gC * gA * Dr_A + half * Dr_B * gB * gA * Dr_C + half * gB2 * Dz_B * Dz_C
         + half * Dz_A * gA * gB * Dz_C + half * gA2 * Dr_A * Dr_C + half * Dr_B * Dr_B * gB * gC
         + gC * gA * C * C - half * gC * gA * C * Dz_A + half * Dz_A * gB * gC * Dz_B
         + half * Dz_A * Dz_A * gA * gC + half * Dz_B * gB * gC * C
         + (half * gH2 * C * Dr_H + half * gB * gA * Dr_B * C - half * gH2 * Dz_H * Dr_H
            + half * gA2 * Dr_A * C) / r
         + half * Dz_H * gH2 * lambda;

    /* calc_conformalRicci4.f90:272-300 */
    RA = RA + Dz_A * gC2 * Dr_C * r3
         + (-half * Dz_A * Dz_A * gC2 + Dz_A * gC2 * C + gB * gA * Dr_C * Dr_C + half * Dr_A * gC2 * Dr_B
            + Dr_A * gC * gA * Dr_C + gB * gC * Dr_C * Dr_B) * r2
         + (-Dz_A * gB * gA * Dr_C + 2.0 * gB * gA * Dr_C * C + gB * gC * C * Dr_B + Dr_A * gC * gA * C) * r
         + 1.0 * fourth * gH2 * Dr_H * Dr_H - gB * gA * C * Dz_A + fourth * gB2 * Dr_B * Dr_B
         + gB * gA * C * C + fourth * Dr_A * Dr_A * gA2 + half * gB * gA * Dz_A * Dz_A;
    RB = RB + Dr_B * gC2 * r3 * Dz_C
         + (gA * gB * Dz_C * Dz_C + half * Dz_A * gC2 * Dz_B - half * gC2 * Dr_B * Dr_B
            + Dz_B * gC * gB * Dz_C + gA * gC * Dz_C * Dz_A) * r2
         - Dr_B * gA * gB * Dz_C * r + fourth * gA2 * Dz_A * Dz_A + half * gB * gA * Dr_B * Dr_B
         + fourth * gH2 * Dz_H * Dz_H + fourth * gB2 * Dz_B * Dz_B;
    RH = RH + (-half * C * gC * gH * Dr_H + gH * gC * Dz_H * Dr_H) * r + half * gH * Dz_H * Dz_H * gB
         + 3.0 * half * gC * Dz_H + half * gH * gA * Dr_H * Dr_H - half * C * gB * gH * Dz_H
         - half * gH * A * gC * Dz_H - C * gC - half * gC * Dz_H
         + half * (2.0 * gA - gH - gH * A * gA) * Dr_H / r + gH * lambda * (1.0 - gA * H);
    RC = RC + gC2 * Dr_C * Dz_C * r3
         + (-half * Dz_A * gC2 * Dz_C + gC2 * C * Dz_C - half * gC2 * Dr_B * Dr_C) * r2
         + (-half * gC2 * Dr_B * C + fourth * Dz_B * gC2 * Dr_A + half * Dr_A * gA * gC * Dz_C
            + half * gC * gB * Dz_C * Dr_B + 3.0 * fourth * gC2 * Dr_B * Dz_A
            + half * Dz_B * gB * gC * Dr_C + half * gC * gA * Dr_C * Dz_A) * r
         + half * Dz_A * gA * gB * Dz_C + half * Dz_B * gB * gC * C + half * gC * gA * C * Dz_A
         + half * Dr_B * gB * gA * Dr_C
         + (fourth * gH2 * Dz_H * Dr_H + half * gB * gA * Dr_B * C + fourth * gA2 * Dr_A * Dz_A
            + fourth * Dz_B * gB2 * Dr_B - half * gB * gA * Dz_A * Dr_B) / r;

    L->RIC_A[k] = RA;
    L->RIC_B[k] = RB;
    L->RIC_H[k] = RH;
    L->RIC_C[k] = RC;
    L->RSCAL[k] = (gA * RA + gB * RB + gH * RH + two * r2 * gC * RC) / L->psi4[k];
  }

  /* divergence of KT: calc_DIVKTA.f90:573-636 (non-angmom) */
  ALL(L, k) {
    const double r = L->r[k], r2 = r * r, r3 = r2 * r, r4 = r3 * r, r5 = r4 * r;
    const double gA = L->g_A[k], gB = L->g_B[k], gH = L->g_H[k], gC = L->g_C[k];
    const double A = EV(A)[k], C = EV(C)[k];
    const double Dr_A = L->Dr_A[k], Dz_A = L->Dz_A[k], Dr_B = L->Dr_B[k], Dz_B = L->Dz_B[k];
    const double Dr_H = L->Dr_H[k], Dz_H = L->Dz_H[k], Dr_C = L->Dr_C[k], Dz_C = L->Dz_C[k];
    const double KTA = EV(KTA)[k], KTB = EV(KTB)[k], KTC = EV(KTC)[k], KTH = EV(KTH)[k];
    const double Dr_KTA = L->Dr_KTA[k], Dz_KTA = L->Dz_KTA[k], Dr_KTB = L->Dr_KTB[k];
    const double Dz_KTB = L->Dz_KTB[k], Dr_KTC = L->Dr_KTC[k], Dz_KTC = L->Dz_KTC[k];
    const double Delta_r = EV(Delta_r)[k], Delta_z = EV(Delta_z)[k];
    const double Alambda = EV(Alambda)[k], lambda = EV(lambda)[k];
    const double gA2 = gA * gA, gB2 = gB * gB, gH2 = gH * gH, gC2 = gC * gC;
    const double gA3 = gA2 * gA, gB3 = gB2 * gB, gC3 = gC2 * gC;

    L->DIV_KT_r[k] = -gC3 * Dz_C * KTC * r5 - gA * gC2 * KTC * Dr_C * r4
      + (gC2 * Dz_KTC - half * gC3 * Dz_B * KTA - gA * gC2 * KTC * C - half * gC3 * Dz_A * KTB
         - gA * Dz_C * gC2 * KTA - gA * Dz_C * gC * gB * KTC - gC2 * Dz_C * gB * KTB
         - gC2 * Dz_B * gB * KTC - gA * Dz_A * gC2 * KTC) * r3
      + (2.0 * gA * gC * Dr_KTC - Delta_r * gC * KTC - gC * Dr_B * gB * gA * KTC
         - gB * gA2 * KTC * Dr_C - Dr_A * gA2 * gC * KTC - gC * gA2 * KTA * Dr_C + gC2 * Dr_KTB
         - gA * gB * gC * KTB * Dr_C - half * gA * Dr_A * gC2 * KTB - half * gC2 * Dr_B * gA * KTA) * r2
      + (2.0 * gA * gC * KTC + gA * gC * Dz_KTA - gC * gA2 * KTA * C + gC * gH * KTC
         - half * gH2 * KTH * gC * Dz_H - Delta_z * gC * KTB - gB * gA2 * KTC * C
         - half * gA2 * Dz_A * gC * KTA + gC * gB * Dz_KTB + gA * gB * Dz_KTC - Delta_z * gA * KTC
         - gA * gB * gC * KTB * C - gC * gH2 * KTH * C - half * gC * Dz_B * gB2 * KTB) * r
      - half * gA * Dr_B * gB2 * KTB - half * gH2 * KTH * gA * Dr_H - Delta_r * gA * KTA
      + gA2 * Dr_KTA - half * gA3 * Dr_A * KTA + r * gH * (gA * Alambda + gH * KTH * C * gC);

    L->DIV_KT_z[k] = -r5 * gC3 * KTC * Dr_C + (-gC3 * KTC * C - Dz_C * gC2 * gB * KTC) * r4
      + (gC2 * Dr_KTC - gB * gC2 * KTB * Dr_C - gC * gB * gA * KTC * Dr_C - half * gC3 * Dr_B * KTA
         - gC2 * gA * KTA * Dr_C - half * Dr_A * gC3 * KTB - Dr_A * gA * gC2 * KTC
         - gB * Dr_B * gC2 * KTC) * r3
      + (-gB * Dz_C * gC * gA * KTA + gC2 * KTC - Dz_B * gC * gB2 * KTC - Delta_z * gC * KTC
         - gC * Dz_C * gB2 * KTB + gC2 * Dz_KTA + 2.0 * gC * gB * Dz_KTC - gC2 * gA * KTA * C
         - Dz_C * gB2 * gA * KTC - half * Dz_A * gB * gC2 * KTB - gC * Dz_A * gB * gA * KTC
         - gB * gC2 * KTB * C - half * gB * Dz_B * gC2 * KTA - gC * gB * gA * KTC * C) * r2
      + (gB * gC * Dr_KTB + gB * gA * Dr_KTC + gC * gA * Dr_KTA - half * gH2 * KTH * gC * Dr_H
         - Delta_r * gB * KTC - Delta_r * gC * KTA - half * gB2 * Dr_B * gC * KTB
         - half * gC * Dr_A * gA2 * KTA) * r
      + gC * gH * KTA + gB * gA * KTC - half * Dz_B * gB3 * KTB + gB2 * Dz_KTB - Delta_z * gB * KTB
      - gC * gH2 * KTH * A - half * gB * Dz_A * gA2 * KTA - half * gH2 * KTH * gB * Dz_H
      - gC * gH * KTH + gB * gH * KTC - gB * gH2 * KTH * C;

    L->DIVA_KT_r[k] = -r5 * gC3 * Dz_C * KTC
      + (-gC3 * Dr_C * KTB - gC3 * Dr_B * KTC - 3.0 * gC2 * Dr_C * gA * KTC) * r4
      + (-gC3 * C * KTB - 2.0 * gA * Dz_C * gC2 * KTA - 3.0 * gC2 * C * gA * KTC + gC2 * Dz_KTC
         - gC2 * Dz_A * gA * KTC - 3.0 * gA * Dz_C * gC * gB * KTC - 2.0 * gC2 * Dz_B * gB * KTC
         - gC3 * Dz_B * KTA - gC2 * Dz_C * gB * KTB) * r3
      + (-gA * Dr_B * gC2 * KTA - gA * gB * gC * KTB * Dr_C + 2.0 * gA * gC * Dr_KTC
         - 2.0 * gC * Dr_C * gA2 * KTA - Delta_r * gC * KTC + gC2 * Dr_KTB - gC2 * Dr_B * gB * KTB
         - 2.0 * gA2 * Dr_A * gC * KTC - gA * Dr_A * gC2 * KTB - gB * gA2 * KTC * Dr_C
         - gA * Dr_B * gC * gB * KTC) * r2
      + (-gA * Dz_C * gB2 * KTB - gB * gA2 * KTC * C - gC * Dz_B * gB2 * KTB + 2.0 * gA * gC * KTC
         + gC * gB * Dz_KTB + gC * gH * KTC + gA * gC * Dz_KTA + gA * gB * Dz_KTC
         - gA * Dz_A * gB * gC * KTB - gC * gH2 * KTH * C - Delta_z * gC * KTB - gA2 * Dz_A * gB * KTC
         - 2.0 * gC * C * gA2 * KTA - Delta_z * gA * KTC - gA2 * Dz_A * gC * KTA
         - gA * gB * gC * KTB * C) * r
      - Dr_A * gA3 * KTA - Delta_r * gA * KTA + gA2 * Dr_KTA + r * gA * gH * (Alambda - lambda * gH * KTH);

    L->DIVA_KT_z[k] = -r5 * gC3 * KTC * Dr_C
      + (-3.0 * Dz_C * gC2 * gB * KTC - Dz_A * gC3 * KTC - Dz_C * gC3 * KTA - gC3 * KTC * C) * r4
      + (-3.0 * gC * gB * gA * KTC * Dr_C - 2.0 * Dr_A * gA * gC2 * KTC + gC2 * Dr_KTC
         - gB * Dr_B * gC2 * KTC - Dr_A * gC3 * KTB - 2.0 * gB * gC2 * KTB * Dr_C
         - gC2 * gA * KTA * Dr_C) * r3
      + (-gC * Dz_A * gB * gA * KTC - gC2 * gA * KTA * C - 2.0 * Dz_B * gC * gB2 * KTC
         - Dz_A * gC2 * gA * KTA - Dz_A * gB * gC2 * KTB - 3.0 * gC * gB * gA * KTC * C
         - 2.0 * gC * Dz_C * gB2 * KTB - Delta_z * gC * KTC - Dz_C * gB2 * gA * KTC + gC2 * KTC
         + gC2 * Dz_KTA + 2.0 * gC * gB * Dz_KTC - gB * Dz_C * gC * gA * KTA - gB * Dz_B * gC2 * KTA
         - 2.0 * gB * gC2 * KTB * C) * r2
      + (-Delta_r * gB * KTC + gB * gA * Dr_KTC + gC * gA * Dr_KTA + gB * gC * Dr_KTB
         - gB2 * Dr_B * gC * KTB - Delta_r * gC * KTA - gB2 * Dr_B * gA * KTC - gC * Dr_A * gA2 * KTA
         - gB * Dr_B * gC * gA * KTA - gB * gA2 * KTA * Dr_C) * r
      - Delta_z * gB * KTB - gB * gH2 * KTH * C + gB2 * Dz_KTB - gC * gH2 * KTH * A
      - gB * gA2 * KTA * C + gB * gH * KTC + gC * gH * KTA + gB * gA * KTC - Dz_B * gB3 * KTB;
  }
}

/* sources_geometry: sources_geometry.f90:77-575 (bssn, nolambda, !evolveKTH) */
/* constraints(box, level) (src/geometry/constraints.f90:36-206), vacuum
 * or real scalar field, hdet = 1, no angmom: the Hamiltonian, momentum and
 * Delta constraints from the arrays of the last auxiliary call.  The
 * routine's smallpi = acos(-1.0) (:41) takes a default-kind (single precision)
 * argument, so pi there is the single-precision value. */
void or_constraints(const oa_params* p, or_level* L) {
  const int matter = (p->mattertype & OA_MATTER_SCALAR) != 0;
  const double half = 0.5, two = 2.0, third = 1.0 / 3.0;
  const double smallpi = (double)acosf(-1.0f);
  for (int k = 0; k < L->n; ++k) {
    const double r = L->r[k], tk = EV(trK)[k];
    double h = L->RSCAL[k] - L->KT2[k] + two * third * tk * tk;                  /* :57 */
    if (matter) h = h - 16.0 * smallpi * L->rho[k];                            /* :60 */
    L->ham[k] = h;
    double mr = L->DIV_KT_r[k] + 6.0 * (L->KT_up_rr[k] * L->Dr_phi[k] + L->KT_up_rz[k] * L->Dz_phi[k])
                - two * third * (L->g_A[k] * L->Dr_trK[k] + r * L->g_C[k] * L->Dz_trK[k]);  /* :116-117, matter :120 */
    double mz = L->DIV_KT_z[k] + 6.0 * (L->KT_up_rz[k] * L->Dr_phi[k] + L->KT_up_zz[k] * L->Dz_phi[k])
                - two * third * (r * L->g_C[k] * L->Dr_trK[k] + L->g_B[k] * L->Dz_trK[k]);  /* :129-130, matter :133 */
    if (matter) {
      mr = mr - 8.0 * smallpi * L->J_r[k] * L->psi4[k];
      mz = mz - 8.0 * smallpi * L->J_z[k] * L->psi4[k];
    }
    L->mom_r[k] = mr;
    L->mom_z[k] = mz;
    L->CDelta_r[k] = EV(Delta_r)[k] + (L->Dr_g_A[k] + r * L->Dz_g_C[k] + r * L->g_lambda[k]
                     + half * (L->g_A[k] * L->Dr_hdet[k] + r * L->g_C[k] * L->Dz_hdet[k]) * L->ihdet[k]);  /* :165-166 */
    L->CDelta_z[k] = EV(Delta_z)[k] + (L->Dr_g_C[k] * r + L->Dz_g_B[k] + two * L->g_C[k]
                     + half * (r * L->g_C[k] * L->Dr_hdet[k] + L->g_B[k] * L->Dz_hdet[k]) * L->ihdet[k]);  /* :172-173 */
  }
  /* boundary copies of ham (:68-80), then the axis / equator images (:87, :93) */
  const int g = L->g, ni = L->ni, nj = L->nj;
#define AT(i, j) (((i) + g - 1) + ((j) + g - 1) * ni)
  for (int j = 1 - g; j <= L->Nz; ++j) L->ham[AT(1 - g, j)] = L->ham[AT(2 - g, j)];
  for (int j = 1 - g; j <= L->Nz; ++j) L->ham[AT(L->Nr, j)] = L->ham[AT(L->Nr - 1, j)];
  for (int i = 1 - g; i <= L->Nr; ++i) L->ham[AT(i, 1 - g)] = L->ham[AT(i, 2 - g)];
  for (int i = 1 - g; i <= L->Nr; ++i) L->ham[AT(i, L->Nz)] = L->ham[AT(i, L->Nz - 1)];
  for (int i = 1; i <= g; ++i)
    for (int j = 1 - g; j <= L->Nz; ++j) L->ham[AT(1 - i, j)] = L->ham[AT(i, j)];
  if (p->eqsym)
    for (int j = 1; j <= g; ++j)
      for (int i = 1 - g; i <= L->Nr; ++i) L->ham[AT(i, 1 - j)] = L->ham[AT(i, j)];
#undef AT
  (void)nj;
}

void or_sources_geometry(const oa_params* p, or_level* L) {
  const int shift = has_shift(p);
  const int matter = (p->mattertype != 0);
  const double sigma = (p->bssnflavor == OA_LAGRANGIAN) ? 1.0 : 0.0;
  const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, two = 2.0, six = 6.0;
  const double smallpi = acos(-1.0);
  const double cp = (double)p->chipower;

  ALL(L, k) {
    const double r = L->r[k], alpha = EV(alpha)[k], trK = EV(trK)[k];
    /* conformal factor (:77-101) */
    double sphi = -sixth * alpha * trK;
    if (shift) sphi = sphi + EV(beta_r)[k] * L->DAr_phi[k] + EV(beta_z)[k] * L->DAz_phi[k]
                      + sixth * sigma * L->DIV_beta[k];
    SRC(phi)[k] = sphi;
    if (p->chimethod) {
      double schi = cp * sixth * alpha * EV(chi)[k] * trK;
      if (shift) schi = schi + EV(beta_r)[k] * L->DAr_chi[k] + EV(beta_z)[k] * L->DAz_chi[k]
                        - sigma * cp * sixth * EV(chi)[k] * L->DIV_beta[k];
      SRC(chi)[k] = schi;
    }
    /* conformal metric (:108-136) */
    const double A = EV(A)[k], B = EV(B)[k], H = EV(H)[k], C = EV(C)[k];
    double sA = -two * alpha * EV(KTA)[k];
    double sB = -two * alpha * EV(KTB)[k];
    double sH = -two * alpha * EV(KTH)[k];
    double sC = -two * alpha * EV(KTC)[k];
    if (shift) {
      const double br = EV(beta_r)[k], bz = EV(beta_z)[k], dv = L->DIV_beta[k];
      sA = sA + br * L->DAr_A[k] + bz * L->DAz_A[k]
           + two * (A * L->Dr_beta_r[k] + r * C * L->Dr_beta_z[k]) - two * third * sigma * A * dv;
      sB = sB + br * L->DAr_B[k] + bz * L->DAz_B[k]
           + two * (B * L->Dz_beta_z[k] + r * C * L->Dz_beta_r[k]) - two * third * sigma * B * dv;
      sH = sH + br * L->DAr_H[k] + bz * L->DAz_H[k] + two * H * br / r - two * third * sigma * H * dv;
      sC = sC + br * L->DAr_C[k] + bz * L->DAz_C[k] + C * br / r + C * (L->Dr_beta_r[k] + L->Dz_beta_z[k])
           + (B * L->Dr_beta_z[k] + A * L->Dz_beta_r[k]) / r - two * third * sigma * C * dv;
    }
    SRC(A)[k] = sA; SRC(B)[k] = sB; SRC(H)[k] = sH; SRC(C)[k] = sC;
  }

  /* trace of K (:200-252) */
  if (p->slicing == OA_SLICING_MAXIMAL) {
    ALL(L, k) SRC(trK)[k] = 0.0;
  } else {
    ALL(L, k) {
      double s = -L->Lapla_alpha[k];
      s = s + EV(alpha)[k] * L->K2[k];
      if (shift) s = s + EV(beta_r)[k] * L->DAr_trK[k] + EV(beta_z)[k] * L->DAz_trK[k];
      if (matter) s = s + 4.0 * smallpi * EV(alpha)[k] * (L->rho[k] + L->trS[k]);
      SRC(trK)[k] = s;
    }
    if (p->geodiss != 0.0) or_dissipation(p, L, EV(trK), SRC(trK), +1, +1, p->geodiss);
  }

  /* traceless extrinsic curvature (:261-392) */
  ALL(L, k) {
    const double r = L->r[k], alpha = EV(alpha)[k], psi4 = L->psi4[k];
    const double tk = EV(trK)[k] + two * EV(z4theta)[k];
    double sKTA = alpha * (tk * EV(KTA)[k] - two * L->KT2_A[k]);
    double sKTB = alpha * (tk * EV(KTB)[k] - two * L->KT2_B[k]);
    double sKTC = alpha * (tk * EV(KTC)[k] - two * L->KT2_C[k]);
    const double tr = -L->Lapla_alpha[k] + alpha * L->RSCAL[k];
    sKTA = sKTA + (-L->D2cov_alpha_A[k] + alpha * L->RIC_A[k]) / psi4 - third * EV(A)[k] * tr;
    sKTB = sKTB + (-L->D2cov_alpha_B[k] + alpha * L->RIC_B[k]) / psi4 - third * EV(B)[k] * tr;
    sKTC = sKTC + (-L->D2cov_alpha_C[k] + alpha * L->RIC_C[k]) / psi4 - third * EV(C)[k] * tr;
    if (shift) {
      const double br = EV(beta_r)[k], bz = EV(beta_z)[k], dv = L->DIV_beta[k];
      const double KTA = EV(KTA)[k], KTB = EV(KTB)[k], KTC = EV(KTC)[k];
      sKTA = sKTA + br * L->DAr_KTA[k] + bz * L->DAz_KTA[k]
             + two * (KTA * L->Dr_beta_r[k] + r * KTC * L->Dr_beta_z[k]) - two * third * sigma * KTA * dv;
      sKTB = sKTB + br * L->DAr_KTB[k] + bz * L->DAz_KTB[k]
             + two * (KTB * L->Dz_beta_z[k] + r * KTC * L->Dz_beta_r[k]) - two * third * sigma * KTB * dv;
      sKTC = sKTC + br * L->DAr_KTC[k] + bz * L->DAz_KTC[k] + KTC * br / r
             + KTC * (L->Dr_beta_r[k] + L->Dz_beta_z[k]) + (KTB * L->Dr_beta_z[k] + KTA * L->Dz_beta_r[k]) / r
             - two * third * sigma * KTC * dv;
    }
    if (matter) {
      const double trS = L->trS[k];
      sKTA = sKTA - 8.0 * smallpi * alpha * (L->S_A[k] / psi4 - third * EV(A)[k] * trS);
      sKTB = sKTB - 8.0 * smallpi * alpha * (L->S_B[k] / psi4 - third * EV(B)[k] * trS);
      sKTC = sKTC - 8.0 * smallpi * alpha * (L->S_C[k] / psi4 - third * EV(C)[k] * trS);
    }
    SRC(KTA)[k] = sKTA; SRC(KTB)[k] = sKTB; SRC(KTC)[k] = sKTC;
  }
  if (p->geodiss != 0.0) {
    or_dissipation(p, L, EV(KTA), SRC(KTA), +1, +1, p->geodiss);
    or_dissipation(p, L, EV(KTB), SRC(KTB), +1, +1, p->geodiss);
    or_dissipation(p, L, EV(KTC), SRC(KTC), +1, -1, p->geodiss);
  }

  /* BSSN Delta (:451-575) */
  ALL(L, k) {
    const double r = L->r[k], alpha = EV(alpha)[k];
    const double gA = L->g_A[k], gB = L->g_B[k], gH = L->g_H[k], gC = L->g_C[k];
    double sr = -two * (L->KT_up_rr[k] * L->Dr_alpha[k] + L->KT_up_rz[k] * L->Dz_alpha[k])
                - two * alpha * L->DIVA_KT_r[k];
    sr = sr + p->eta * alpha * (L->DIV_KT_r[k] + six * (L->KT_up_rr[k] * L->Dr_phi[k] + L->KT_up_rz[k] * L->Dz_phi[k])
                                - two * third * (gA * L->Dr_trK[k] + r * gC * L->Dz_trK[k]));
    double sz = -two * (L->KT_up_rz[k] * L->Dr_alpha[k] + L->KT_up_zz[k] * L->Dz_alpha[k])
                - two * alpha * L->DIVA_KT_z[k];
    sz = sz + p->eta * alpha * (L->DIV_KT_z[k] + six * (L->KT_up_rz[k] * L->Dr_phi[k] + L->KT_up_zz[k] * L->Dz_phi[k])
                                - two * third * (r * gC * L->Dr_trK[k] + gB * L->Dz_trK[k]));
    if (shift) {
      const double br = EV(beta_r)[k], bz = EV(beta_z)[k];
      const double Dr_ = EV(Delta_r)[k], Dz_ = EV(Delta_z)[k], dv = L->DIV_beta[k];
      sr = sr + br * L->DAr_Delta_r[k] + bz * L->DAz_Delta_r[k] - Dr_ * L->Dr_beta_r[k] - Dz_ * L->Dz_beta_r[k];
      sr = sr + two * r * gC * L->Drz_beta_r[k] + gA * L->Drr_beta_r[k] + gB * L->Dzz_beta_r[k]
           + gH * L->DD_beta_rr[k];
      sr = sr + third * sigma * (gA * L->Dr_DIV_beta[k] + r * gC * L->Dz_DIV_beta[k] + two * Dr_ * dv);
      sz = sz + br * L->DAr_Delta_z[k] + bz * L->DAz_Delta_z[k] - Dr_ * L->Dr_beta_z[k] - Dz_ * L->Dz_beta_z[k];
      sz = sz + two * r * gC * L->Drz_beta_z[k] + gA * L->Drr_beta_z[k] + gB * L->Dzz_beta_z[k]
           + gH * L->Dr_beta_z[k] / r;
      sz = sz + third * sigma * (r * gC * L->Dr_DIV_beta[k] + gB * L->Dz_DIV_beta[k] + two * Dz_ * dv);
    }
    if (matter) {
      sr = sr - 8.0 * smallpi * p->eta * alpha * L->J_r[k] * L->psi4[k];
      sz = sz - 8.0 * smallpi * p->eta * alpha * L->J_z[k] * L->psi4[k];
    }
    SRC(Delta_r)[k] = sr;
    SRC(Delta_z)[k] = sz;
  }
  if (p->geodiss != 0.0) {
    or_dissipation(p, L, EV(Delta_r), SRC(Delta_r), -1, +1, p->geodiss);
    or_dissipation(p, L, EV(Delta_z), SRC(Delta_z), +1, -1, p->geodiss);
  }
  /* BSSN: z4theta source is zero (:679-683) */
  ALL(L, k) SRC(z4theta)[k] = 0.0;
}

/* sources_lapse: sources_lapse.f90:28-68 (first-order Bona-Masso) */
void or_sources_lapse(const oa_params* p, or_level* L) {
  if (p->slicing == OA_SLICING_STATIC || p->slicing == OA_SLICING_MAXIMAL) {
    ALL(L, k) SRC(alpha)[k] = 0.0;
    return;
  }
  ALL(L, k) {
    double s = -L->falpha[k] * EV(trK)[k];
    if (has_shift(p)) s = s + EV(beta_r)[k] * L->DAr_alpha[k] + EV(beta_z)[k] * L->DAz_alpha[k];
    if (p->gauge_eta != 0.0) s = s - p->gauge_eta * log(fabs(EV(alpha)[k]));
    SRC(alpha)[k] = s;
  }
}

/* sources_shift: sources_shift.f90:208-306 (Gammadriver2) */
void or_sources_shift(const oa_params* p, or_level* L) {
  if (p->shift == OA_SHIFT_ZERO || p->shift == OA_SHIFT_STATIC) {
    ALL(L, k) { SRC(beta_r)[k] = 0.0; SRC(beta_z)[k] = 0.0; }
    return;
  }
  ALL(L, k) {
    double sbr = EV(dtbeta_r)[k], sbz = EV(dtbeta_z)[k];
    const double br = EV(beta_r)[k], bz = EV(beta_z)[k];
    if (p->driveradv) {
      sbr = sbr + br * L->DAr_beta_r[k] + bz * L->DAz_beta_r[k];
      sbz = sbz + br * L->DAr_beta_z[k] + bz * L->DAz_beta_z[k];
    }
    double sdr = p->drivercsi * SRC(Delta_r)[k];
    double sdz = p->drivercsi * SRC(Delta_z)[k];
    if (p->driveradv) {
      sdr = sdr + br * (L->DAr_dtbeta_r[k] - L->DAr_Delta_r[k]) + bz * (L->DAz_dtbeta_r[k] - L->DAz_Delta_r[k]);
      sdz = sdz + br * (L->DAr_dtbeta_z[k] - L->DAr_Delta_z[k]) + bz * (L->DAz_dtbeta_z[k] - L->DAz_Delta_z[k]);
    }
    sdr = sdr - p->drivereta * EV(dtbeta_r)[k];
    sdz = sdz - p->drivereta * EV(dtbeta_z)[k];
    SRC(beta_r)[k] = sbr; SRC(beta_z)[k] = sbz;
    SRC(dtbeta_r)[k] = sdr; SRC(dtbeta_z)[k] = sdz;
  }
}

/* radiative_geometry: radiative_geometry.f90:58-237 (level 0 only) */
void or_radiative_geometry(const oa_params* p, or_level* L) {
  const double vl = 1.0, va = sqrt(p->gauge_f), vs = sqrt(4.0 * (1.0 / 3.0) * p->drivercsi);
  const double third = 1.0 / 3.0, one = 1.0, two = 2.0, var0 = 0.0;
  const int Nr = L->Nr, Nz = L->Nz, g = L->g;
  if (p->slicing != OA_SLICING_MAXIMAL)
    or_radbound(p, L, EV(trK), SRC(trK), L->Dr_trK, L->Dz_trK, va, var0);
  or_radbound(p, L, EV(KTA), SRC(KTA), L->Dr_KTA, L->Dz_KTA, vl, var0);
  or_radbound(p, L, EV(KTB), SRC(KTB), L->Dr_KTB, L->Dz_KTB, vl, var0);
  for (int j = 1 - g; j <= Nz; ++j) {
    int k = IX(L, Nr, j);
    SRC(KTA)[k] = SRC(KTA)[k] + third * (one - vl / va) * SRC(trK)[k] * two;
    SRC(KTB)[k] = SRC(KTB)[k] - third * (one - vl / va) * SRC(trK)[k];
  }
  if (!p->eqsym)
    for (int i = 1 - g; i <= Nr; ++i) {
      int k = IX(L, i, 1 - g);
      SRC(KTA)[k] = SRC(KTA)[k] - third * (one - vl / va) * SRC(trK)[k];
      SRC(KTB)[k] = SRC(KTB)[k] + third * (one - vl / va) * SRC(trK)[k] * two;
    }
  for (int i = 1 - g; i <= Nr; ++i) {
    int k = IX(L, i, Nz);
    SRC(KTA)[k] = SRC(KTA)[k] - third * (one - vl / va) * SRC(trK)[k];
    SRC(KTB)[k] = SRC(KTB)[k] + third * (one - vl / va) * SRC(trK)[k] * two;
  }
  or_radbound(p, L, EV(KTC), SRC(KTC), L->Dr_KTC, L->Dz_KTC, vl, var0);
  if ((p->shift == OA_SHIFT_GD2 || p->shift == OA_SHIFT_GDSHOCK2) && p->drivercsi != 0.0) {
    or_radbound(p, L, EV(dtbeta_r), SRC(dtbeta_r), L->Dr_dtbeta_r, L->Dz_dtbeta_r, vs, var0);
    or_radbound(p, L, EV(dtbeta_z), SRC(dtbeta_z), L->Dr_dtbeta_z, L->Dz_dtbeta_z, vs, var0);
  }
}
