/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see or_internal.h).
 *
 * Level allocation and coordinates (setupboxes.f90, evolve.f90:42-72),
 * closed-form initial data (initial.f90), the initial auxiliary pass
 * (evolve.f90:128-195), the Berger-Oliger step (onestep.f90) and a small C
 * API used by the Python tests through ctypes.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "or_internal.h"

#define EV(f) L->var[OF_##f]

/* ---------------------------------------------------------------- names */
/* name -> offset of every auxiliary array in or_level */
#define AUXLIST(X) \
  X(auxarray) X(falpha) X(psi) X(psi2) X(psi4) X(hdet) X(ihdet) X(Dr_hdet) X(Dz_hdet) \
  X(Dr_alpha) X(Dz_alpha) X(Drr_alpha) X(Dzz_alpha) X(Drz_alpha) X(DAr_alpha) X(DAz_alpha) \
  X(Dr_phi) X(Dz_phi) X(Drr_phi) X(Dzz_phi) X(Drz_phi) X(DAr_phi) X(DAz_phi) \
  X(Dr_chi) X(Dz_chi) X(Drr_chi) X(Dzz_chi) X(Drz_chi) X(DAr_chi) X(DAz_chi) \
  X(Dr_A) X(Dz_A) X(Drr_A) X(Dzz_A) X(Drz_A) X(DAr_A) X(DAz_A) \
  X(Dr_B) X(Dz_B) X(Drr_B) X(Dzz_B) X(Drz_B) X(DAr_B) X(DAz_B) \
  X(Dr_H) X(Dz_H) X(Drr_H) X(Dzz_H) X(Drz_H) X(DAr_H) X(DAz_H) \
  X(Dr_C) X(Dz_C) X(Drr_C) X(Dzz_C) X(Drz_C) X(DAr_C) X(DAz_C) \
  X(g_A) X(g_B) X(g_H) X(g_C) \
  X(Dr_g_A) X(Dz_g_A) X(Dr_g_B) X(Dz_g_B) X(Dr_g_H) X(Dz_g_H) X(Dr_g_C) X(Dz_g_C) \
  X(Dr_beta_r) X(Dz_beta_r) X(Drr_beta_r) X(Dzz_beta_r) X(Drz_beta_r) X(DAr_beta_r) X(DAz_beta_r) \
  X(Dr_beta_z) X(Dz_beta_z) X(Drr_beta_z) X(Dzz_beta_z) X(Drz_beta_z) X(DAr_beta_z) X(DAz_beta_z) \
  X(Dr_dtbeta_r) X(Dz_dtbeta_r) X(DAr_dtbeta_r) X(DAz_dtbeta_r) \
  X(Dr_dtbeta_z) X(Dz_dtbeta_z) X(DAr_dtbeta_z) X(DAz_dtbeta_z) \
  X(DD_beta_rr) X(DIV_beta) X(Dr_DIV_beta) X(Dz_DIV_beta) \
  X(lamDef) X(AlamDef) X(g_lamDef) X(g_lambda) \
  X(Dr_trK) X(Dz_trK) X(DAr_trK) X(DAz_trK) X(Dr_KTA) X(Dz_KTA) X(DAr_KTA) X(DAz_KTA) \
  X(Dr_KTB) X(Dz_KTB) X(DAr_KTB) X(DAz_KTB) X(Dr_KTC) X(Dz_KTC) X(DAr_KTC) X(DAz_KTC) \
  X(KT_up_rr) X(KT_up_zz) X(KT_up_rz) X(KT2_A) X(KT2_B) X(KT2_H) X(KT2_C) X(KT2) X(K2) \
  X(DelDef_r) X(DelDef_z) \
  X(Dr_Delta_r) X(Dz_Delta_r) X(DAr_Delta_r) X(DAz_Delta_r) \
  X(Dr_Delta_z) X(Dz_Delta_z) X(DAr_Delta_z) X(DAz_Delta_z) \
  X(D2cov_alpha_A) X(D2cov_alpha_B) X(D2cov_alpha_H) X(D2cov_alpha_C) X(Lapla_alpha) \
  X(RIC_A) X(RIC_B) X(RIC_H) X(RIC_C) X(RSCAL) \
  X(DIV_KT_r) X(DIV_KT_z) X(DIVA_KT_r) X(DIVA_KT_z) \
  X(Dr_scalar_phi) X(Dz_scalar_phi) X(Drr_scalar_phi) X(Dzz_scalar_phi) X(Drz_scalar_phi) \
  X(Dr_scalar_pi) X(Dz_scalar_pi) X(DAr_scalar_pi) X(DAz_scalar_pi) \
  X(Dr_scalar_xi_r) X(Dz_scalar_xi_r) X(DAr_scalar_xi_r) X(DAz_scalar_xi_r) \
  X(Dr_scalar_xi_z) X(Dz_scalar_xi_z) X(DAr_scalar_xi_z) X(DAz_scalar_xi_z) \
  X(scalar_V) X(scalar_VP) X(rho) X(J_r) X(J_z) X(S_A) X(S_B) X(S_H) X(S_C) X(trS) \
  X(ham) X(mom_r) X(mom_z) X(CDelta_r) X(CDelta_z) \
  X(r) X(z) X(rr)

typedef struct { const char* name; size_t off; } or_auxname;
#define AUXENTRY(nm) {#nm, offsetof(or_level, nm)},
static const or_auxname aux_names[] = { AUXLIST(AUXENTRY) };
static const int n_aux_names = (int)(sizeof(aux_names) / sizeof(aux_names[0]));

static int field_id(const char* name) {
  for (int f = 0; f < OF_NFIELDS; ++f)
    if (strcmp(or_fields[f].name, name) == 0) return f;
  return -1;
}

/* ----------------------------------------------------------- allocation */
static double* zalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

static void level_alloc(const oa_params* p, or_level* L, int level, int g) {
  memset(L, 0, sizeof(*L));
  L->g = g;
  L->Nr = p->Nrtotal;
  L->Nz = p->Nztotal;
  L->ni = L->Nr + g;
  L->nj = L->Nz + g;
  L->n = L->ni * L->nj;
  const double scale = (double)(1 << level);
  L->dr = p->dr0 / scale;
  L->dz = p->dz0 / scale;
  L->r = zalloc(L->n); L->z = zalloc(L->n); L->rr = zalloc(L->n);
  /* coordinates: evolve.f90:52-71 (box 0, single process) */
  for (int j = 1 - g; j <= L->Nz; ++j)
    for (int i = 1 - g; i <= L->Nr; ++i) {
      const int k = IX(L, i, j);
      L->r[k] = ((double)i - 0.5) * L->dr;
      if (p->eqsym) L->z[k] = ((double)j - 0.5) * L->dz;
      else L->z[k] = ((double)j - 0.5 * (double)(L->Nz - g + 1)) * L->dz;
      L->rr[k] = sqrt(L->r[k] * L->r[k] + L->z[k] * L->z[k]);
    }
  L->rmin = L->r[IX(L, 1 - g, 0)];
  L->rmax = L->r[IX(L, L->Nr, 0)];
  L->zmin = L->z[IX(L, 0, 1 - g)];
  L->zmax = L->z[IX(L, 0, L->Nz)];
  for (int f = 0; f < OF_NFIELDS; ++f) {
    int on = (f <= OF_z4theta);
    if (f >= OF_beta_r && f <= OF_fdriver) on = (p->shift != OA_SHIFT_NONE);
    if (f >= OF_scalar_phi) on = (p->mattertype & OA_MATTER_SCALAR) != 0;
    L->present[f] = on;
    if (!on) continue;
    L->var[f] = zalloc(L->n); L->var_p[f] = zalloc(L->n);
    L->var_a[f] = zalloc(L->n); L->src[f] = zalloc(L->n);
    L->b_rL[f] = zalloc((size_t)g * L->nj * 3); L->b_rR[f] = zalloc((size_t)g * L->nj * 3);
    L->b_zL[f] = zalloc((size_t)L->ni * g * 3); L->b_zR[f] = zalloc((size_t)L->ni * g * 3);
  }
  for (int a = 0; a < n_aux_names; ++a) {
    double** slot = (double**)((char*)L + aux_names[a].off);
    if (*slot == NULL) *slot = zalloc(L->n);
  }
}

static void level_free(or_level* L) {
  for (int f = 0; f < OF_NFIELDS; ++f) {
    free(L->var[f]); free(L->var_p[f]); free(L->var_a[f]); free(L->src[f]);
    free(L->b_rL[f]); free(L->b_rR[f]); free(L->b_zL[f]); free(L->b_zR[f]);
  }
  for (int a = 0; a < n_aux_names; ++a) free(*(double**)((char*)L + aux_names[a].off));
}

or_handle* or_create(const oa_params* pin) {
  or_handle* H = (or_handle*)calloc(1, sizeof(or_handle));
  H->p = *pin;
  oa_params* p = &H->p;
  /* ghost zones: main.f90:286-317 */
  const int gmin = (p->order == OA_ORDER_TWO) ? 2 : 3;
  int g = p->ghost == 0 ? gmin : (p->ghost < gmin ? gmin : p->ghost);
  p->ghost = g;
  /* setupboxes.f90:104-141 */
  if (!p->eqsym && ((p->Nztotal + g) % 2 == 0)) p->Nztotal += 1;
  H->dt0 = p->dtfac * (p->dr0 < p->dz0 ? p->dr0 : p->dz0);
  H->nlev = p->Nl0 + 1;
  H->L = (or_level*)calloc((size_t)H->nlev, sizeof(or_level));
  for (int l = 0; l < H->nlev; ++l) {
    level_alloc(p, &H->L[l], l, g);
    H->L[l].dt = H->dt0 / (double)(1 << l);
  }
  return H;
}

void or_destroy(or_handle* H) {
  if (!H) return;
  or_ell_free(H);
  for (int l = 0; l < H->nlev; ++l) level_free(&H->L[l]);
  free(H->L);
  free(H);
}

/* ---------------------------------------------------------- auxiliary */
void or_auxiliary(const oa_params* p, or_level* L) {
  or_auxiliary_geometry(p, L);
  if (p->mattertype != 0) {
    or_auxiliary_matter(p, L);
    or_potential(p, L);
    or_stressenergy(p, L);
  }
}

/* simpleboundary (generated, prl/arrays.pl FILE_SIMPLEBOUNDARY): sources of
 * the evolving arrays without NOBOUND */
static void simpleboundary(const oa_params* p, or_level* L) {
  static const int bound_fields[] = {OF_dtalpha, OF_trK, OF_KTA, OF_KTB, OF_KTH, OF_KTC,
                                     OF_Delta_r, OF_Delta_z, OF_Alambda, OF_z4theta,
                                     OF_dtbeta_r, OF_dtbeta_z, OF_fdriver, OF_scalar_pi};
  const int g = L->g;
  for (size_t q = 0; q < sizeof(bound_fields) / sizeof(int); ++q) {
    const int f = bound_fields[q];
    if (!L->present[f]) continue;
    double* s = L->src[f];
    if (p->boundtype == OA_BOUND_STATIC) {
      for (int j = 1 - g; j <= L->Nz; ++j) s[IX(L, L->Nr, j)] = 0.0;
      for (int i = 1 - g; i <= L->Nr; ++i) s[IX(L, i, L->Nz)] = 0.0;
      if (!p->eqsym) for (int i = 1 - g; i <= L->Nr; ++i) s[IX(L, i, 1 - g)] = 0.0;
    } else if (p->boundtype == OA_BOUND_FLAT) {
      for (int j = 1 - g; j <= L->Nz; ++j) s[IX(L, L->Nr, j)] = s[IX(L, L->Nr - 1, j)];
      for (int i = 1 - g; i <= L->Nr; ++i) s[IX(L, i, L->Nz)] = s[IX(L, i, L->Nz - 1)];
      if (!p->eqsym) for (int i = 1 - g; i <= L->Nr; ++i) s[IX(L, i, 1 - g)] = s[IX(L, i, 2 - g)];
    }
  }
}

/* sources for one level (onestep.f90:184-255) */
void or_level_sources(const oa_params* p, or_level* L, int level) {
  if (p->spacetime == OA_DYNAMIC) {
    or_sources_geometry(p, L);
    or_sources_lapse(p, L);
    if (p->shift != OA_SHIFT_NONE) or_sources_shift(p, L);
  }
  if (p->mattertype & OA_MATTER_SCALAR) or_sources_scalar(p, L);
  if (level == 0) {
    if (p->boundtype == OA_BOUND_RADIATIVE) or_radiative_geometry(p, L);
    else if (p->boundtype == OA_BOUND_STATIC || p->boundtype == OA_BOUND_FLAT) simpleboundary(p, L);
  }
}

/* ----------------------------------------------------------- onestep */
int or_onestep(or_handle* H, int level) {
  const oa_params* p = &H->p;
  or_level* L = &H->L[level];
  or_saveold(p, L);
  const int rk4 = (p->integrator == OA_RK4);
  const int niter = rk4 ? 4 : p->icniter;
  for (int iter = 1; iter <= niter; ++iter) {
    double dtw = 0.0, w = 0.0;
    if (rk4) {
      switch (iter) {
        case 1: dtw = 0.5 * L->dt; w = 1.0 / 6.0; break;
        case 2: dtw = 0.5 * L->dt; w = 1.0 / 3.0; break;
        case 3: dtw = L->dt; w = 1.0 / 3.0; break;
        default: dtw = L->dt; w = 1.0 / 6.0; break;
      }
    } else {
      dtw = (iter < niter) ? 0.5 * L->dt : L->dt;
    }
    or_level_sources(p, L, level);
    if (rk4) or_accumulate(p, L, iter, niter, w);
    or_update(p, L, dtw);
    if (level > 0) or_finebound(H, level, dtw);
    or_symmetries_r(p, L);
    if (p->eqsym) or_symmetries_z(p, L);
    or_auxiliary(p, L);
  }
  L->t2 = L->t1;
  L->t1 = L->t;
  L->s = L->s + 1;
  L->t = L->t + L->dt;
  if (level < H->nlev - 1) {
    or_onestep(H, level + 1);
    or_onestep(H, level + 1);
  }
  if (level > 0 && (L->s % 2) == 0) {
    or_restrict(H, level);
    or_level* Lc = &H->L[level - 1];
    or_symmetries_r(p, Lc);
    if (p->eqsym) or_symmetries_z(p, Lc);
    or_auxiliary(p, Lc);
  }
  return 0;
}

int or_coarse_step(or_handle* H) { return or_onestep(H, 0); }

int or_evolve(or_handle* H, int nsteps) {
  for (int s = 0; s < nsteps; ++s) or_onestep(H, 0);
  return 0;
}

/* ------------------------------------------------------ initial data */
/* gaussian shell / torus profile of initial.f90:613-660 and
 * idata_scalarpulse.f90:60-90 */
static double pulse_shell(double a0, double r0_, double z0_, double sr, double sz,
                          double r, double z, double rr) {
  if (r0_ == 0.0 && z0_ == 0.0)
    return a0 * exp(-(rr * rr * rr * rr / (r * r * sr * sr + z * z * sz * sz)));
  const double c = (r * r * r0_ + z * z * z0_) / (rr * rr);
  const double w = rr * rr / (r * r * sr * sr + z * z * sz * sz);
  return a0 * (exp(-(rr - c) * (rr - c) * w) + exp(-(rr + c) * (rr + c) * w));
}

int or_initial_data(or_handle* H) {
  const oa_params* p = &H->p;
  const double third = 1.0 / 3.0, one = 1.0, half = 0.5, two = 2.0;
  for (int l = 0; l < H->nlev; ++l) {
    or_level* L = &H->L[l];
    /* Minkowski defaults (initial.f90:53-150) */
    ALL(L, k) {
      EV(alpha)[k] = one; EV(dtalpha)[k] = 0.0;
      L->psi[k] = one; EV(phi)[k] = 0.0; EV(chi)[k] = one; L->psi2[k] = one; L->psi4[k] = one;
      EV(A)[k] = one; EV(B)[k] = one; EV(H)[k] = one; EV(C)[k] = 0.0;
      L->g_A[k] = one; L->g_B[k] = one; L->g_H[k] = one; L->g_C[k] = 0.0;
      EV(trK)[k] = 0.0; EV(KTA)[k] = 0.0; EV(KTB)[k] = 0.0; EV(KTC)[k] = 0.0; EV(KTH)[k] = 0.0;
      EV(Delta_r)[k] = 0.0; EV(Delta_z)[k] = 0.0; EV(lambda)[k] = 0.0; EV(Alambda)[k] = 0.0;
      EV(z4theta)[k] = 0.0;
      if (p->shift != OA_SHIFT_NONE) {
        EV(beta_r)[k] = 0.0; EV(beta_z)[k] = 0.0; EV(dtbeta_r)[k] = 0.0; EV(dtbeta_z)[k] = 0.0;
        EV(fdriver)[k] = one;
      }
    }
    if (p->idata == OA_IDATA_SCHWARZSCHILD) {   /* :189-225 */
      ALL(L, k) {
        const double rr = fabs(L->rr[k]);
        const double ps = one + half * p->BH1mass / rr;
        L->psi[k] = ps;
        EV(phi)[k] = log(ps);
        EV(chi)[k] = one / pow(ps, (double)p->chipower);
        L->psi2[k] = ps * ps;
        L->psi4[k] = ps * ps * ps * ps;
        if (p->ilapse == OA_ILAPSE_ISOTROPIC)
          EV(alpha)[k] = (one - half * p->BH1mass / rr) / (one + half * p->BH1mass / rr);
      }
    } else if (p->idata == OA_IDATA_BRILLLINDQUIST) {   /* :284-330 */
      ALL(L, k) {
        const double r = L->r[k], z = L->z[k];
        double ps;
        if (p->N_BH == 1) {
          ps = one + half * p->BH1mass / sqrt(r * r + (z - p->BH1Z) * (z - p->BH1Z));
          if (p->eqsym && p->duplicate_bhs) ps = ps + half * p->BH1mass / sqrt(r * r + (z + p->BH1Z) * (z + p->BH1Z));
        } else {
          ps = one + half * (p->BH1mass / sqrt(r * r + (z - p->BH1Z) * (z - p->BH1Z))
                             + p->BH2mass / sqrt(r * r + (z - p->BH2Z) * (z - p->BH2Z)));
          if (p->eqsym && p->duplicate_bhs)
            ps = ps + half * (p->BH1mass / sqrt(r * r + (z + p->BH1Z) * (z + p->BH1Z))
                              + p->BH2mass / sqrt(r * r + (z + p->BH2Z) * (z + p->BH2Z)));
        }
        L->psi[k] = ps;
        EV(phi)[k] = log(ps);
        EV(chi)[k] = one / pow(ps, (double)p->chipower);
        L->psi2[k] = ps * ps;
        L->psi4[k] = ps * ps * ps * ps;
      }
    } else if (p->idata == OA_IDATA_SCALARPULSE) {   /* idata_scalarpulse.f90:60-103 */
      ALL(L, k) {
        const double r = L->r[k], z = L->z[k], rr = L->rr[k];
        double v;
        if (p->scalarpulse_type == OA_PULSE_SHELL) {
          v = pulse_shell(p->scalar_a0, p->scalar_r0, p->scalar_z0, p->scalar_sr0, p->scalar_sz0, r, z, rr);
          if (p->scalar_r0 == 0.0 && p->scalar_z0 == 0.0)
            v = p->scalar_a0 * exp(-(r * r / (p->scalar_sr0 * p->scalar_sr0) + z * z / (p->scalar_sz0 * p->scalar_sz0)));
        } else if (p->scalarpulse_type == OA_PULSE_TORUS) {
          const double sr = p->scalar_sr0, sz = p->scalar_sz0, r0_ = p->scalar_r0, z0_ = p->scalar_z0;
          if (r0_ == 0.0 && z0_ == 0.0)
            v = p->scalar_a0 * exp(-(r / sr) * (r / sr)) * exp(-(z / sz) * (z / sz));
          else if (p->eqsym)
            v = p->scalar_a0 * (exp(-((r - r0_) / sr) * ((r - r0_) / sr)) + exp(-((r + r0_) / sr) * ((r + r0_) / sr)))
                * (exp(-((z - z0_) / sz) * ((z - z0_) / sz)) + exp(-((z + z0_) / sz) * ((z + z0_) / sz)));
          else
            v = p->scalar_a0 * exp(-((z - z0_) / sz) * ((z - z0_) / sz))
                * (exp(-((r - r0_) / sr) * ((r - r0_) / sr)) + exp(-((r + r0_) / sr) * ((r + r0_) / sr)));
        } else {
          v = p->scalar_a0 * exp(-(r * r + (one - p->scalar_sz0) * z * z));
        }
        EV(scalar_phi)[k] = v;
        EV(scalar_pi)[k] = 0.0;
      }
      or_diff1r(p, L, EV(scalar_phi), +1, EV(scalar_xi_r));
      or_diff1z(p, L, EV(scalar_phi), +1, EV(scalar_xi_z));
      or_potential(p, L);
    }
  }
  /* Brill waves and the scalar pulse solve the Hamiltonian constraint for
   * psi on every level (idata_brillwave.f90:117-135,
   * idata_scalarpulse.f90:131-178) */
  const int solve = (p->idata == OA_IDATA_BRILLWAVE || p->idata == OA_IDATA_SCALARPULSE);
  if (solve) or_solve_hamiltonian(H);
  for (int l = 0; l < H->nlev; ++l) {
    or_level* L = &H->L[l];
    if (p->idata == OA_IDATA_BRILLWAVE) {
      /* idata_brillwave.f90:183-208: psi from the solve, Holz q function */
      ALL(L, k) {
        const double r = L->r[k], z = L->z[k];
        const double q = p->brill_a0 * r * r * exp(-(r / p->brill_sr0) * (r / p->brill_sr0)
                                                   - (z / p->brill_sz0) * (z / p->brill_sz0));
        const double ps = L->psi[k];
        L->psi2[k] = ps * ps;
        L->psi4[k] = ps * ps * ps * ps;
        EV(phi)[k] = log(ps);
        EV(chi)[k] = one / pow(ps, (double)p->chipower);
        EV(A)[k] = exp(two * q);
        EV(B)[k] = EV(A)[k];
        EV(C)[k] = 0.0;
        EV(H)[k] = one;
        EV(lambda)[k] = (EV(A)[k] - EV(H)[k]) / (r * r);
        EV(Delta_r)[k] = (one - one / EV(A)[k]) / r;
        EV(Delta_z)[k] = 0.0;
      }
    } else if (p->idata == OA_IDATA_SCALARPULSE) {
      /* idata_scalarpulse.f90:185-206: psi from the solve */
      ALL(L, k) {
        const double ps = L->psi[k];
        L->psi2[k] = ps * ps;
        L->psi4[k] = ps * ps * ps * ps;
        EV(phi)[k] = log(ps);
        EV(chi)[k] = one / pow(ps, (double)p->chipower);
      }
    }
    /* factor the determinant (initial.f90:528-580) */
    if (p->factorhdet) {
      ALL(L, k) {
        const double r = L->r[k];
        double hdet = (EV(A)[k] * EV(B)[k] - (r * EV(C)[k]) * (r * EV(C)[k])) * EV(H)[k];
        EV(phi)[k] = EV(phi)[k] + log(hdet) / 12.0;
        const double aux = pow(hdet, third);
        EV(A)[k] /= aux; EV(B)[k] /= aux; EV(C)[k] /= aux; EV(H)[k] /= aux;
        L->hdet[k] = 1.0; L->ihdet[k] = 1.0;
        const double ps = exp(EV(phi)[k]);
        L->psi[k] = ps; L->psi2[k] = ps * ps; L->psi4[k] = ps * ps * ps * ps;
      }
    }
    /* remove the trace of KT (initial.f90:585-620) */
    ALL(L, k) {
      const double r = L->r[k];
      const double aux = L->g_A[k] * EV(KTA)[k] + L->g_B[k] * EV(KTB)[k] + L->g_H[k] * EV(KTH)[k]
                         + two * r * r * L->g_C[k] * EV(KTC)[k];
      EV(KTA)[k] -= third * EV(A)[k] * aux;
      EV(KTB)[k] -= third * EV(B)[k] * aux;
      EV(KTH)[k] -= third * EV(H)[k] * aux;
      EV(KTC)[k] -= third * EV(C)[k] * aux;
    }
    /* precollapsed lapse (initial.f90:627-645) */
    if (p->ilapse == OA_ILAPSE_PSIMINUS2) ALL(L, k) EV(alpha)[k] = one / (L->psi[k] * L->psi[k]);
    if (p->ilapse == OA_ILAPSE_PSIMINUS4) ALL(L, k) EV(alpha)[k] = one / (L->psi4[k]);
    /* lapse perturbation (initial.f90:650-700) */
    if (p->lapsepert == OA_PERT_GAUSS) {
      ALL(L, k) {
        const double r = L->r[k], z = L->z[k], rr = L->rr[k];
        if (p->ilapse == OA_ILAPSE_SHELL) {
          EV(alpha)[k] += pulse_shell(p->lapse_a0, p->lapse_r0, p->lapse_z0, p->lapse_sr0, p->lapse_sz0, r, z, rr);
        } else if (p->ilapse == OA_ILAPSE_TORUS) {
          const double sr = p->lapse_sr0, sz = p->lapse_sz0, r0_ = p->lapse_r0, z0_ = p->lapse_z0;
          if (r0_ == 0.0 && z0_ == 0.0)
            EV(alpha)[k] += p->lapse_a0 * exp(-(r / sr) * (r / sr)) * exp(-(z / sz) * (z / sz));
          else if (p->eqsym)
            EV(alpha)[k] += p->lapse_a0 * (exp(-((r - r0_) / sr) * ((r - r0_) / sr)) + exp(-((r + r0_) / sr) * ((r + r0_) / sr)))
                            * (exp(-((z - z0_) / sz) * ((z - z0_) / sz)) + exp(-((z + z0_) / sz) * ((z + z0_) / sz)));
          else
            EV(alpha)[k] += p->lapse_a0 * exp(-((z - z0_) / sz) * ((z - z0_) / sz))
                            * (exp(-((r - r0_) / sr) * ((r - r0_) / sr)) + exp(-((r + r0_) / sr) * ((r + r0_) / sr)));
        }
      }
    }
    /* shift perturbation, shell type (initial.f90:705-745) */
    if (p->shiftpert == OA_PERT_GAUSS && p->shift != OA_SHIFT_NONE && p->ishift == OA_ISHIFT_SHELL) {
      ALL(L, k) {
        const double r = L->r[k], z = L->z[k], rr = L->rr[k];
        const double sr = p->shift_sr0, sz = p->shift_sz0;
        double prof;
        if (p->shift_r0 == 0.0 && p->shift_z0 == 0.0) {
          prof = p->shift_a0 * exp(-(rr * rr * rr * rr / (r * r * sr * sr + z * z * sz * sz)));
        } else {
          const double c = (r * r * p->shift_r0 + z * z * p->shift_z0) / (rr * rr);
          const double w = rr * rr / (r * r * sr * sr + z * z * sz * sz);
          prof = p->shift_a0 * (exp(-(rr - c) * (rr - c) * w) - exp(-(rr + c) * (rr + c) * w));
        }
        EV(beta_r)[k] += (r / rr) * prof;
        EV(beta_z)[k] += (z / rr) * prof;
      }
    }
  }
  return 0;
}

/* evolve.f90:128-195 */
int or_initial_aux(or_handle* H) {
  const oa_params* p = &H->p;
  for (int l = 0; l < H->nlev; ++l) {
    or_symmetries_r(p, &H->L[l]);
    if (p->eqsym) or_symmetries_z(p, &H->L[l]);
  }
  for (int l = 0; l < H->nlev; ++l) or_auxiliary(p, &H->L[l]);
  if ((p->shift == OA_SHIFT_GD2 || p->shift == OA_SHIFT_GDSHOCK2) && !p->driverD0) {
    for (int l = 0; l < H->nlev; ++l) {
      or_level* L = &H->L[l];
      ALL(L, k) {
        EV(dtbeta_r)[k] = p->drivercsi * EV(Delta_r)[k];
        EV(dtbeta_z)[k] = p->drivercsi * EV(Delta_z)[k];
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------- ctypes API */
int or_nlevels(or_handle* H) { return H->nlev; }

int or_level_dims(or_handle* H, int lev, int* Nr, int* Nz, int* g, double* dr, double* dz, double* dt) {
  if (lev < 0 || lev >= H->nlev) return 1;
  const or_level* L = &H->L[lev];
  *Nr = L->Nr; *Nz = L->Nz; *g = L->g; *dr = L->dr; *dz = L->dz; *dt = L->dt;
  return 0;
}

double* or_field(or_handle* H, int lev, const char* name) {
  if (lev < 0 || lev >= H->nlev) return NULL;
  or_level* L = &H->L[lev];
  const int f = field_id(name);
  if (f >= 0) return L->present[f] ? L->var[f] : NULL;
  for (int a = 0; a < n_aux_names; ++a)
    if (strcmp(aux_names[a].name, name) == 0) return *(double**)((char*)L + aux_names[a].off);
  return NULL;
}

double* or_source(or_handle* H, int lev, const char* name) {
  if (lev < 0 || lev >= H->nlev) return NULL;
  const int f = field_id(name);
  return (f >= 0 && H->L[lev].present[f]) ? H->L[lev].src[f] : NULL;
}

void or_get_time(or_handle* H, int lev, double* t, double* t1, double* t2, int* s) {
  const or_level* L = &H->L[lev];
  *t = L->t; *t1 = L->t1; *t2 = L->t2; *s = L->s;
}

void or_set_time(or_handle* H, int lev, double t, double t1, double t2, int s) {
  or_level* L = &H->L[lev];
  L->t = t; L->t1 = t1; L->t2 = t2; L->s = s;
}

void or_set_ext_derivs(or_handle* H, int lev, int flag) { H->L[lev].ext_derivs = flag; }

/* kind: 0 diff1r, 1 diff1z, 2 diff2r, 3 diff2z, 4 diff2rz, 5 diffadvr, 6 diffadvz
 * (for 5/6 "beta" is the shift component that selects the upwind side) */
int or_apply_stencil(or_handle* H, int lev, int kind, const double* f, int symr, int symz,
                     const double* beta, double* out) {
  const oa_params* p = &H->p;
  const or_level* L = &H->L[lev];
  switch (kind) {
    case 0: or_diff1r(p, L, f, symr, out); break;
    case 1: or_diff1z(p, L, f, symz, out); break;
    case 2: or_diff2r(p, L, f, symr, out); break;
    case 3: or_diff2z(p, L, f, symz, out); break;
    case 4: or_diff2rz(p, L, f, symr, symz, out); break;
    case 5: or_diffadvr(p, L, f, symr, beta, out); break;
    case 6: or_diffadvz(p, L, f, symz, beta, out); break;
    default: return 1;
  }
  return 0;
}

int or_apply_dissipation(or_handle* H, int lev, const double* ev, double* src, int symr, int symz,
                         double diss) {
  or_dissipation(&H->p, &H->L[lev], ev, src, symr, symz, diss);
  return 0;
}

double or_interp_at(or_handle* H, int lev, const double* f, double r0, double z0) {
  return or_interp(&H->p, &H->L[lev], f, r0, z0);
}

int or_run_auxiliary(or_handle* H, int lev) { or_auxiliary(&H->p, &H->L[lev]); return 0; }
int or_run_constraints(or_handle* H, int lev) { or_constraints(&H->p, &H->L[lev]); return 0; }
int or_run_sources(or_handle* H, int lev) { or_level_sources(&H->p, &H->L[lev], lev); return 0; }
int or_run_saveold(or_handle* H, int lev) { or_saveold(&H->p, &H->L[lev]); return 0; }
int or_run_finebound(or_handle* H, int lev, double dtw) { or_finebound(H, lev, dtw); return 0; }
int or_run_restrict(or_handle* H, int lev) { or_restrict(H, lev); return 0; }
int or_run_update(or_handle* H, int lev, int iter, int niter, double dtw, double w) {
  const oa_params* p = &H->p;
  if (p->integrator == OA_RK4) or_accumulate(p, &H->L[lev], iter, niter, w);
  or_update(p, &H->L[lev], dtw);
  return 0;
}
int or_run_symmetries(or_handle* H, int lev) {
  or_symmetries_r(&H->p, &H->L[lev]);
  if (H->p.eqsym) or_symmetries_z(&H->p, &H->L[lev]);
  return 0;
}
