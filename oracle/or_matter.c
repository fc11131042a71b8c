/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see or_internal.h).
 *
 * Real scalar field: auxiliary_matter (scalar part), potential, stressenergy
 * (scalar part) and sources_scalar of the reference.
 */
#include <math.h>
#include <string.h>
#include "or_internal.h"

#define EV(f) L->var[OF_##f]
#define SRC(f) L->src[OF_##f]

/* auxiliary_matter.f90:25-77 */
void or_auxiliary_matter(const oa_params* p, or_level* L) {
  if (!(p->mattertype & OA_MATTER_SCALAR) || L->ext_derivs) return;
  or_diff1r(p, L, EV(scalar_phi), +1, L->Dr_scalar_phi);
  or_diff1z(p, L, EV(scalar_phi), +1, L->Dz_scalar_phi);
  or_diff2r(p, L, EV(scalar_phi), +1, L->Drr_scalar_phi);
  or_diff2z(p, L, EV(scalar_phi), +1, L->Dzz_scalar_phi);
  or_diff2rz(p, L, EV(scalar_phi), +1, +1, L->Drz_scalar_phi);
  or_diff1r(p, L, EV(scalar_pi), +1, L->Dr_scalar_pi);
  or_diff1z(p, L, EV(scalar_pi), +1, L->Dz_scalar_pi);
  or_diff1r(p, L, EV(scalar_xi_r), -1, L->Dr_scalar_xi_r);
  or_diff1z(p, L, EV(scalar_xi_r), +1, L->Dz_scalar_xi_r);
  or_diff1r(p, L, EV(scalar_xi_z), +1, L->Dr_scalar_xi_z);
  or_diff1z(p, L, EV(scalar_xi_z), -1, L->Dz_scalar_xi_z);
  if (p->shift != OA_SHIFT_NONE) {
    /* the matter advective derivatives are always upwinded (:57-75) */
    or_diffadvr(p, L, EV(scalar_pi), +1, EV(beta_r), L->DAr_scalar_pi);
    or_diffadvz(p, L, EV(scalar_pi), +1, EV(beta_z), L->DAz_scalar_pi);
    or_diffadvr(p, L, EV(scalar_xi_r), -1, EV(beta_r), L->DAr_scalar_xi_r);
    or_diffadvz(p, L, EV(scalar_xi_r), +1, EV(beta_z), L->DAz_scalar_xi_r);
    or_diffadvr(p, L, EV(scalar_xi_z), +1, EV(beta_r), L->DAr_scalar_xi_z);
    or_diffadvz(p, L, EV(scalar_xi_z), -1, EV(beta_z), L->DAz_scalar_xi_z);
  }
}

/* potential.f90:27-73 */
void or_potential(const oa_params* p, or_level* L) {
  if (!(p->mattertype & OA_MATTER_SCALAR)) return;
  ALL(L, k) {
    const double ph = EV(scalar_phi)[k];
    double V = 0.0, VP = 0.0;
    if (p->scalarpotential == OA_POT_PHI2) {
      V = 0.5 * p->scalar_mass * p->scalar_mass * ph * ph;
      VP = p->scalar_mass * p->scalar_mass * ph;
    } else if (p->scalarpotential == OA_POT_PHI4) {
      V = 0.5 * p->scalar_mass * p->scalar_mass * ph * ph + 0.25 * p->scalar_lambda * ph * ph * ph * ph;
      VP = p->scalar_mass * p->scalar_mass * ph + p->scalar_lambda * ph * ph * ph;
    }
    L->scalar_V[k] = V;
    L->scalar_VP[k] = VP;
  }
}

/* stressenergy.f90:65-205,321-331 (vacuum + real scalar) */
void or_stressenergy(const oa_params* p, or_level* L) {
  const double half = 0.5, two = 2.0;
  ALL(L, k) {
    if (p->chimethod) EV(phi)[k] = -half * log(EV(chi)[k]);
    const double ps = exp(EV(phi)[k]);
    L->psi[k] = ps;
    L->psi4[k] = ps * ps * ps * ps;
    double rho = 0.0, Jr = 0.0, Jz = 0.0, SA = 0.0, SB = 0.0, SH = 0.0, SC = 0.0;
    if (p->mattertype & OA_MATTER_SCALAR) {
      const double r = L->r[k], psi4 = L->psi4[k];
      const double gA = L->g_A[k], gB = L->g_B[k], gC = L->g_C[k];
      const double xr = EV(scalar_xi_r)[k], xz = EV(scalar_xi_z)[k], pi = EV(scalar_pi)[k];
      const double V = L->scalar_V[k];
      rho = rho + half * (pi * pi + (gA * xr * xr + gB * xz * xz + two * r * gC * xr * xz) / psi4) + V;
      Jr = Jr - (gA * xr + r * gC * xz) * pi / psi4;
      Jz = Jz - (gB * xz + r * gC * xr) * pi / psi4;
      const double aux = half * (pi * pi - (gA * xr * xr + gB * xz * xz + two * r * gC * xr * xz) / psi4) - V;
      SA = SA + xr * xr + EV(A)[k] * psi4 * aux;
      SB = SB + xz * xz + EV(B)[k] * psi4 * aux;
      SC = SC + xr * xz / r + EV(C)[k] * psi4 * aux;
      SH = SH + EV(H)[k] * psi4 * aux;
    }
    L->rho[k] = rho; L->J_r[k] = Jr; L->J_z[k] = Jz;
    L->S_A[k] = SA; L->S_B[k] = SB; L->S_H[k] = SH; L->S_C[k] = SC;
    const double r = L->r[k];
    double trS = L->g_A[k] * SA + L->g_B[k] * SB + L->g_H[k] * SH + two * r * r * L->g_C[k] * SC;
    L->trS[k] = trS / L->psi4[k];
  }
}

/* sources_scalar: scalar.f90:94-251 */
void or_sources_scalar(const oa_params* p, or_level* L) {
  const double half = 0.5, two = 2.0;
  const int shift = (p->shift != OA_SHIFT_NONE);
  ALL(L, k) {
    const double r = L->r[k], alpha = EV(alpha)[k], pi = EV(scalar_pi)[k];
    const double xr = EV(scalar_xi_r)[k], xz = EV(scalar_xi_z)[k];
    const double gA = L->g_A[k], gB = L->g_B[k], gC = L->g_C[k];
    double sph = alpha * pi;
    if (shift) sph = sph + EV(beta_r)[k] * xr + EV(beta_z)[k] * xz;
    double sxr = alpha * L->Dr_scalar_pi[k] + pi * L->Dr_alpha[k];
    if (shift) sxr = sxr + EV(beta_r)[k] * L->DAr_scalar_xi_r[k] + xr * L->Dr_beta_r[k]
                     + EV(beta_z)[k] * L->DAz_scalar_xi_r[k] + xz * L->Dr_beta_z[k];
    double sxz = alpha * L->Dz_scalar_pi[k] + pi * L->Dz_alpha[k];
    if (shift) sxz = sxz + EV(beta_r)[k] * L->DAr_scalar_xi_z[k] + xr * L->Dz_beta_r[k]
                     + EV(beta_z)[k] * L->DAz_scalar_xi_z[k] + xz * L->Dz_beta_z[k];
    double spi;
    double dr, dz;   /* the two first derivatives the Laplacian is built from */
    if (p->scalarmethod == OA_SCALARMETHOD_FIRST) {
      dr = xr; dz = xz;
      spi = alpha * (gA * L->Dr_scalar_xi_r[k] + gB * L->Dz_scalar_xi_z[k]
                     + r * gC * (L->Dr_scalar_xi_z[k] + L->Dz_scalar_xi_r[k]));
    } else {
      dr = L->Dr_scalar_phi[k]; dz = L->Dz_scalar_phi[k];
      spi = alpha * (gA * L->Drr_scalar_phi[k] + gB * L->Dzz_scalar_phi[k] + two * r * gC * L->Drz_scalar_phi[k]);
    }
    spi = spi + two * alpha * (L->Dr_phi[k] * (gA * dr + r * gC * dz) + L->Dz_phi[k] * (gB * dz + r * gC * dr));
    spi = spi + L->Dr_alpha[k] * (gA * dr + r * gC * dz) + L->Dz_alpha[k] * (gB * dz + r * gC * dr);
    spi = spi + alpha * (dr * (L->Dr_g_A[k] + r * L->Dz_g_C[k]) + dz * (L->Dz_g_B[k] + r * L->Dr_g_C[k] + gC));
    spi = spi + alpha * (gA * dr / r + gC * dz);
    spi = spi + half * alpha * L->ihdet[k] * (L->Dr_hdet[k] * (gA * dr + r * gC * dz)
                                              + L->Dz_hdet[k] * (gB * dz + r * gC * dr));
    spi = spi / L->psi4[k];
    spi = spi + alpha * pi * EV(trK)[k];
    if (p->scalarpotential != OA_POT_NONE) spi = spi - alpha * L->scalar_VP[k];
    if (shift) spi = spi + EV(beta_r)[k] * L->DAr_scalar_pi[k] + EV(beta_z)[k] * L->DAz_scalar_pi[k];
    SRC(scalar_phi)[k] = sph;
    SRC(scalar_xi_r)[k] = sxr;
    SRC(scalar_xi_z)[k] = sxz;
    SRC(scalar_pi)[k] = spi;
  }
  if (p->scalardiss != 0.0) {
    or_dissipation(p, L, EV(scalar_phi), SRC(scalar_phi), +1, +1, p->scalardiss);
    or_dissipation(p, L, EV(scalar_pi), SRC(scalar_pi), +1, +1, p->scalardiss);
    or_dissipation(p, L, EV(scalar_xi_r), SRC(scalar_xi_r), -1, +1, p->scalardiss);
    or_dissipation(p, L, EV(scalar_xi_z), SRC(scalar_xi_z), +1, -1, p->scalardiss);
  }
  or_radbound(p, L, EV(scalar_pi), SRC(scalar_pi), L->Dr_scalar_pi, L->Dz_scalar_pi, 1.0, 0.0);
}
