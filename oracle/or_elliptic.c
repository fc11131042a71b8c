/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see or_internal.h).
 *
 * The elliptic solvers of the Hamiltonian constraint used by Brill-wave
 * and scalar-pulse initial data (ELL_solver = wave or sor).  WaveElliptic:
 * wave_elliptic('flat', 'one')
 * (src/elliptic/wave_elliptic.f90:2-532) and its recursive step wavestep
 * (:542-1164), for box 0.  The elliptic equation
 *     lap u + S1 u + S5 u^5 = S0      (flat Laplacian in (rho, z))
 * is relaxed as the damped wave equation du/dt = v,
 * dv/dt = lap u + S1 u + S5 u^5 - S0 - WE_eta v, with Kreiss-Oliger
 * dissipation and radiative boundaries on v, until the RMS of dv/dt on the
 * base level falls below WE_epsilon (after at least 100 steps).  With
 * refinement levels the base level is solved alone first; its solution is
 * then interpolated onto every finer level and all levels are relaxed
 * together with Berger-Oliger subcycling (single-variable finebound and
 * restrict, finebound.f90 / restrict.f90 with all = .false.).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "or_internal.h"

/* one level's elliptic arrays (arrays.config ELLIPTIC / AUXILIARY ell_*) */
typedef struct or_ell_level {
  double *u, *v, *u_p, *v_p, *u_a, *v_a, *su, *sv;
  double *Dr_u, *Dz_u, *Drr_u, *Dzz_u, *Drz_u, *Dr_v, *Dz_v;
  double *S0, *S1, *S5;
  /* fine-boundary time levels, (0:g-1, 1-g:Nz, 0:2) and (1-g:Nr, 0:g-1, 0:2) */
  double *bu_rR, *bu_zL, *bu_zR, *bv_rR, *bv_zL, *bv_zR;
} or_ell_level;

static double* zalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

static double pow5(double x) {
  const double x2 = x * x;
  return x2 * x2 * x;
}

static void ell_alloc(or_handle* H) {
  if (H->ell) return;
  H->ell = (or_ell_level*)calloc((size_t)H->nlev, sizeof(or_ell_level));
  for (int l = 0; l < H->nlev; ++l) {
    const or_level* L = &H->L[l];
    or_ell_level* E = &H->ell[l];
    double** planes[] = {&E->u, &E->v, &E->u_p, &E->v_p, &E->u_a, &E->v_a, &E->su, &E->sv,
                         &E->Dr_u, &E->Dz_u, &E->Drr_u, &E->Dzz_u, &E->Drz_u, &E->Dr_v, &E->Dz_v,
                         &E->S0, &E->S1, &E->S5};
    for (size_t q = 0; q < sizeof(planes) / sizeof(planes[0]); ++q) *planes[q] = zalloc((size_t)L->n);
    E->bu_rR = zalloc((size_t)L->g * L->nj * 3);
    E->bv_rR = zalloc((size_t)L->g * L->nj * 3);
    E->bu_zL = zalloc((size_t)L->ni * L->g * 3);
    E->bv_zL = zalloc((size_t)L->ni * L->g * 3);
    E->bu_zR = zalloc((size_t)L->ni * L->g * 3);
    E->bv_zR = zalloc((size_t)L->ni * L->g * 3);
  }
}

void or_ell_free(or_handle* H) {
  if (!H->ell) return;
  for (int l = 0; l < H->nlev; ++l) {
    or_ell_level* E = &H->ell[l];
    double* planes[] = {E->u, E->v, E->u_p, E->v_p, E->u_a, E->v_a, E->su, E->sv, E->Dr_u, E->Dz_u,
                        E->Drr_u, E->Dzz_u, E->Drz_u, E->Dr_v, E->Dz_v, E->S0, E->S1, E->S5,
                        E->bu_rR, E->bu_zL, E->bu_zR, E->bv_rR, E->bv_zL, E->bv_zR};
    for (size_t q = 0; q < sizeof(planes) / sizeof(planes[0]); ++q) free(planes[q]);
  }
  free(H->ell);
  H->ell = NULL;
}

/* time levels of the fine boundary (wavestep :639-678): t-2 <- t-1 <- t0 <- now */
static void ell_save_strips(const or_level* L, const double* f, double* b_rR, double* b_zL, double* b_zR) {
  const int g = L->g, ni = L->ni, nj = L->nj;
  for (int m = 0; m < g; ++m)
    for (int j = 1 - g; j <= L->Nz; ++j) {
      const int k = m + g * (j + g - 1);
      b_rR[k + 2 * g * nj] = b_rR[k + g * nj];
      b_rR[k + g * nj] = b_rR[k];
      b_rR[k] = f[IX(L, L->Nr - m, j)];
    }
  for (int n = 0; n < g; ++n)
    for (int i = 1 - g; i <= L->Nr; ++i) {
      const int k = (i + g - 1) + ni * n;
      b_zL[k + 2 * ni * g] = b_zL[k + ni * g];
      b_zL[k + ni * g] = b_zL[k];
      b_zL[k] = f[IX(L, i, 1 - g + n)];
      b_zR[k + 2 * ni * g] = b_zR[k + ni * g];
      b_zR[k + ni * g] = b_zR[k];
      b_zR[k] = f[IX(L, i, L->Nz - n)];
    }
}

/* finebound(box, level, dtw, .false.) for one variable (finebound.f90:70-400,
 * the single-variable branches) */
static void ell_finebound(const oa_params* p, const or_level* L, const or_level* Lc, double dtw,
                          double* fine, const double* fine_p, const double* cvar,
                          const double* b_rR, const double* b_zL, const double* b_zR) {
  const int g = L->g, ni = L->ni, nj = L->nj;
  const double tp = Lc->t, t0 = L->t, tm1 = L->t1, tl = t0 + dtw;
  const int imax = (tp == tl) ? g + 2 : g - 1;
  const int jmax = imax;
  const int border = (tm1 == 0.0) ? 1 : 2;
  /* strip entries beyond g - 1 are read by the reference with weight zero */
#define BORDER_VALUE(strip, k0, k1, inside)                                                          \
  ((border == 1) ? (tl - t0) / (tp - t0) * aux2 + (tl - tp) / (t0 - tp) * fine_p[IX(L, i, j)]      \
                 : (tl - t0) * (tl - tm1) / ((tp - t0) * (tp - tm1)) * aux2                          \
                       + (tl - tp) * (tl - tm1) / ((t0 - tp) * (t0 - tm1)) * ((inside) ? strip[k0] : 0.0) \
                       + (tl - tp) * (tl - t0) / ((tm1 - tp) * (tm1 - t0)) * ((inside) ? strip[k1] : 0.0))
  for (int m = 0; m <= imax; ++m)
    for (int n = 0; n <= L->Nz + g - 1; ++n) {
      const double r0 = L->rmax - (double)m * L->dr;
      const double z0 = L->zmin + (double)n * L->dz;
      const int i = (int)lround((r0 - L->rmin) / L->dr) + 1 - g;
      const int j = (int)lround((z0 - L->zmin) / L->dz) + 1 - g;
      const double aux2 = or_interp(p, Lc, cvar, r0, z0);
      if (!(i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz)) continue;
      const int bk = m + g * (j + g - 1);
      fine[IX(L, i, j)] = BORDER_VALUE(b_rR, bk, bk + g * nj, m < g);
    }
  for (int n = 0; n <= jmax; ++n)
    for (int m = 0; m <= L->Nr + g - 1; ++m) {
      const double r0 = L->rmin + (double)m * L->dr;
      const double z0 = L->zmax - (double)n * L->dz;
      const int i = (int)lround((r0 - L->rmin) / L->dr) + 1 - g;
      const int j = (int)lround((z0 - L->zmin) / L->dz) + 1 - g;
      const double aux2 = or_interp(p, Lc, cvar, r0, z0);
      if (!(i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz)) continue;
      const int bk = (i + g - 1) + ni * n;
      fine[IX(L, i, j)] = BORDER_VALUE(b_zR, bk, bk + ni * g, n < g);
    }
  if (!p->eqsym)
    for (int n = 0; n <= jmax; ++n)
      for (int m = 0; m <= L->Nr + g - 1; ++m) {
        const double r0 = L->rmin + (double)m * L->dr;
        const double z0 = L->zmin + (double)n * L->dz;
        const int i = (int)lround((r0 - L->rmin) / L->dr) + 1 - g;
        const int j = (int)lround((z0 - L->zmin) / L->dz) + 1 - g;
        const double aux2 = or_interp(p, Lc, cvar, r0, z0);
        if (!(i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz)) continue;
        const int bk = (i + g - 1) + ni * n;
        fine[IX(L, i, j)] = BORDER_VALUE(b_zL, bk, bk + ni * g, n < g);
      }
#undef BORDER_VALUE
}

/* restrict(box, level, .false.) for one variable (restrict.f90:88-162) */
static void ell_restrict(const oa_params* p, const or_level* L, const or_level* Lc, const double* fvar,
                         double* cvar) {
  const int g = L->g;
  const double deltar = (double)(2 * g) * L->dr + 1.0e-10;
  const double deltaz = (double)(2 * g) * L->dz + 1.0e-10;
  for (int n = 0; n <= Lc->Nz + g - 1; ++n)
    for (int m = 0; m <= Lc->Nr + g - 1; ++m) {
      const double r0 = Lc->rmin + (double)m * Lc->dr;
      const double z0 = Lc->zmin + (double)n * Lc->dz;
      const double rmin = (L->rmin < 0.0) ? 0.0 : L->rmin + deltar;
      const double zmin = (p->eqsym && L->zmin < 0.0) ? 0.0 : L->zmin + deltaz;
      const double rmax = L->rmax - deltar, zmax = L->zmax - deltaz;
      if (r0 > rmin && r0 < rmax && z0 > zmin && z0 < zmax) {
        const int i = (int)lround((r0 - Lc->rmin) / Lc->dr) + 1 - g;
        const int j = (int)lround((z0 - Lc->zmin) / Lc->dz) + 1 - g;
        const double aux2 = or_interp(p, L, fvar, r0, z0);
        if (i >= 1 - g && i <= Lc->Nr && j >= 1 - g && j <= Lc->Nz) cvar[IX(Lc, i, j)] = aux2;
      }
    }
}

/* ell_u, ell_v images on the axis and (eqsym) the equator, both even */
static void ell_symmetries(const oa_params* p, const or_level* L, double* f) {
  const int g = L->g;
  for (int j = 1 - g; j <= L->Nz; ++j)
    for (int i = 1; i <= g; ++i) f[IX(L, 1 - i, j)] = f[IX(L, i, j)];
  if (p->eqsym)
    for (int i = 1 - g; i <= L->Nr; ++i)
      for (int j = 1; j <= g; ++j) f[IX(L, i, 1 - j)] = f[IX(L, i, j)];
}

/* wavestep(level, 'flat', method): wave_elliptic.f90:542-1164, box 0 */
static void wavestep(or_handle* H, int level, int nlmax) {
  const oa_params* p = &H->p;
  or_level* L = &H->L[level];
  or_ell_level* E = &H->ell[level];
  const int g = L->g;
  const int rk4 = (p->WE_method == OA_RK4);
  /* save the old time step and the fine-boundary time levels (:631-679) */
  memcpy(E->u_p, E->u, sizeof(double) * (size_t)L->n);
  memcpy(E->v_p, E->v, sizeof(double) * (size_t)L->n);
  if (level > 0) {
    ell_save_strips(L, E->u, E->bu_rR, E->bu_zL, E->bu_zR);
    ell_save_strips(L, E->v, E->bv_rR, E->bv_zL, E->bv_zR);
  }
  const int niter = rk4 ? 4 : p->icniter;
  for (int iter = 1; iter <= niter; ++iter) {
    double dtw = 0.0, weight = 0.0;
    if (rk4) {
      switch (iter) {
        case 1: dtw = 0.5 * L->dt; weight = 1.0 / 6.0; break;
        case 2: dtw = 0.5 * L->dt; weight = 1.0 / 3.0; break;
        case 3: dtw = L->dt; weight = 1.0 / 3.0; break;
        default: dtw = L->dt; weight = 1.0 / 6.0; break;
      }
    } else {
      dtw = (iter < niter) ? 0.5 * L->dt : L->dt;
    }
    /* derivatives of u and v (:773-787) */
    or_diff1r(p, L, E->u, +1, E->Dr_u);
    or_diff1z(p, L, E->u, +1, E->Dz_u);
    or_diff2r(p, L, E->u, +1, E->Drr_u);
    or_diff2z(p, L, E->u, +1, E->Dzz_u);
    or_diff2rz(p, L, E->u, +1, +1, E->Drz_u);
    or_diff1r(p, L, E->v, +1, E->Dr_v);
    or_diff1z(p, L, E->v, +1, E->Dz_v);
    /* sources: flat Laplacian, the S terms, damping (:791-851) */
    ALL(L, k) {
      E->su[k] = E->v[k];
      E->sv[k] = E->Drr_u[k] + E->Dzz_u[k] + E->Dr_u[k] / L->r[k];
    }
    ALL(L, k) E->sv[k] = E->sv[k] + E->S1[k] * E->u[k] + E->S5[k] * pow5(E->u[k]) - E->S0[k];
    ALL(L, k) E->sv[k] = E->sv[k] - p->WE_eta * E->v[k];
    /* dissipation on v (:855-857), then the symmetries of its source (:861-873) */
    or_dissipation(p, L, E->v, E->sv, +1, +1, p->WE_diss);
    for (int j = 1 - g; j <= L->Nz; ++j)
      for (int i = 1; i <= g; ++i) E->sv[IX(L, 1 - i, j)] = E->sv[IX(L, i, j)];
    if (p->eqsym)
      for (int i = 1 - g; i <= L->Nr; ++i)
        for (int j = 1; j <= g; ++j) E->sv[IX(L, i, 1 - j)] = E->sv[IX(L, i, j)];
    /* radiative boundary on v (:884-888) */
    or_radbound(p, L, E->v, E->sv, E->Dr_v, E->Dz_v, 1.0, 0.0);
    /* Runge-Kutta accumulators (:895-907) */
    if (rk4) {
      if (iter == 1) {
        ALL(L, k) { E->u_a[k] = weight * E->su[k]; E->v_a[k] = weight * E->sv[k]; }
      } else if (iter < niter) {
        ALL(L, k) { E->u_a[k] = E->u_a[k] + weight * E->su[k]; E->v_a[k] = E->v_a[k] + weight * E->sv[k]; }
      } else {
        ALL(L, k) { E->su[k] = E->u_a[k] + weight * E->su[k]; E->sv[k] = E->v_a[k] + weight * E->sv[k]; }
      }
    }
    /* update (:914-915) */
    ALL(L, k) {
      E->u[k] = E->u_p[k] + dtw * E->su[k];
      E->v[k] = E->v_p[k] + dtw * E->sv[k];
    }
    /* fine-level boundaries from the new coarse time level (:925-968) */
    if (level > 0) {
      const or_level* Lc = &H->L[level - 1];
      const or_ell_level* Ec = &H->ell[level - 1];
      ell_finebound(p, L, Lc, dtw, E->u, E->u_p, Ec->u, E->bu_rR, E->bu_zL, E->bu_zR);
      ell_finebound(p, L, Lc, dtw, E->v, E->v_p, Ec->v, E->bv_rR, E->bv_zL, E->bv_zR);
    }
  }
  /* clocks (:997-1004) */
  L->t2 = L->t1;
  L->t1 = L->t;
  L->s = L->s + 1;
  L->t = L->t + L->dt;
  if (level < nlmax) {
    wavestep(H, level + 1, nlmax);
    wavestep(H, level + 1, nlmax);
  }
  if (nlmax > 0) {   /* :1031-1061 */
    ell_symmetries(p, L, E->u);
    ell_symmetries(p, L, E->v);
  }
  if (level > 0 && (L->s % 2) == 0) {   /* :1106-1157 */
    or_level* Lc = &H->L[level - 1];
    or_ell_level* Ec = &H->ell[level - 1];
    ell_restrict(p, L, Lc, E->u, Ec->u);
    ell_restrict(p, L, Lc, E->v, Ec->v);
    ell_symmetries(p, Lc, Ec->u);
    ell_symmetries(p, Lc, Ec->v);
  }
}

static void reset_clocks(or_handle* H) {
  for (int l = 0; l < H->nlev; ++l) {
    H->L[l].t = 0.0; H->L[l].t1 = 0.0; H->L[l].t2 = 0.0; H->L[l].s = 0;
  }
}

/* Stationary exit (not in the reference; off unless or_set_ell_stationary
 * turns it on).  With refinement levels the base-level residual stalls at
 * the mismatch between the coarse and the restricted fine solution, so the
 * reference's loop runs to WE_maxiter while the state stops changing after
 * a few thousand steps.  This rule ends the all-levels solve once the
 * largest per-step change of u on the base level (interior range of the
 * residual), relative to max|u|, has stayed below 1e-12 over a window of
 * 100 steps and has stopped falling (window maximum above half the previous
 * window's): the remaining steps can only move u by rounding. */
static int g_ell_stationary = 0;
void or_set_ell_stationary(int on) { g_ell_stationary = on; }

typedef struct {
  double wmax, wprev;
} stationary_state;

static int stationary_done(stationary_state* st, int step, double delta) {
  if (delta > st->wmax) st->wmax = delta;
  if (step % 100 != 0) return 0;
  const int done = (st->wmax <= 1.0e-12 && st->wmax > 0.5 * st->wprev);
  st->wprev = st->wmax;
  st->wmax = 0.0;
  return done;
}

/* wave_elliptic('flat', 'one') (:2-532).  Solves into ell->u of every level
 * with the coefficients already in S0, S1, S5.  steps[0] / steps[1]: the
 * iterations of the base-level solve and (with refinement) of the solve on
 * all levels; *res: the last residual. */
static void wave_elliptic(or_handle* H, int* steps, double* res) {
  const oa_params* p = &H->p;
  /* keep the external clocks and time steps (:78-95) */
  double* save = (double*)malloc(sizeof(double) * 4 * (size_t)H->nlev);
  int* ssave = (int*)malloc(sizeof(int) * (size_t)H->nlev);
  for (int l = 0; l < H->nlev; ++l) {
    save[4 * l] = H->L[l].t; save[4 * l + 1] = H->L[l].t1; save[4 * l + 2] = H->L[l].t2;
    save[4 * l + 3] = H->L[l].dt; ssave[l] = H->L[l].s;
  }
  reset_clocks(H);
  /* initial guess u = 1, v = 0 (:107-141) */
  for (int l = 0; l < H->nlev; ++l) {
    or_ell_level* E = &H->ell[l];
    ALL((&H->L[l]), k) { E->u[k] = 1.0; E->v[k] = 0.0; }
  }
  /* flat Laplacian: dt0 = WE_dtfac min(dr0, dz0), halved per level (:178-254) */
  const double dt0 = p->WE_dtfac * (p->dr0 < p->dz0 ? p->dr0 : p->dz0);
  for (int l = 0; l < H->nlev; ++l) H->L[l].dt = dt0 / (double)(1 << l);
  const int nlmax_old = H->nlev - 1;
  int nlmax = nlmax_old > 0 ? 0 : nlmax_old;
  steps[0] = steps[1] = 0;
  double gres = 1.0;
  for (int phase = 0;; ++phase) {
    int step = 0;
    gres = 1.0;
    const int miniter = 100;
    const int stationary = g_ell_stationary && nlmax > 0;
    stationary_state st = {0.0, 0.0};
    const or_level* Lb = &H->L[0];
    while (step < miniter || (gres > p->WE_epsilon && step < p->WE_maxiter)) {
      step = step + 1;
      wavestep(H, 0, nlmax);
      if (stationary) {
        const double* u = H->ell[0].u;
        const double* up = H->ell[0].u_p;
        double dmax = 0.0, umax = 0.0;
        for (int j = 1; j <= Lb->Nz - Lb->g; ++j)
          for (int i = 1; i <= Lb->Nr - Lb->g; ++i) {
            const int k = IX(Lb, i, j);
            dmax = fmax(dmax, fabs(u[k] - up[k]));
            umax = fmax(umax, fabs(u[k]));
          }
        if (stationary_done(&st, step, dmax / umax)) {
          /* the residual of this step still decides nothing: leave the loop */
          const double* sv0 = H->ell[0].sv;
          double l2 = 0.0;
          for (int j = 1; j <= Lb->Nz - Lb->g; ++j)
            for (int i = 1; i <= Lb->Nr - Lb->g; ++i) l2 = l2 + sv0[IX(Lb, i, j)] * sv0[IX(Lb, i, j)];
          gres = sqrt(l2 / ((double)Lb->Nr * (double)Lb->Nz));
          break;
        }
      }
      /* residual: RMS of the source of v on the base level (:321-341) */
      const or_level* L0 = &H->L[0];
      const double* sv = H->ell[0].sv;
      double lres = 0.0;
      for (int j = 1; j <= L0->Nz - L0->g; ++j)
        for (int i = 1; i <= L0->Nr - L0->g; ++i) lres = lres + sv[IX(L0, i, j)] * sv[IX(L0, i, j)];
      const double nntot = (double)L0->Nr * (double)L0->Nz;
      gres = sqrt(lres / nntot);
    }
    steps[phase] = step;
    /* inject the base-level solution into the finer levels and solve again
     * on all of them (:426-500) */
    if (step != p->ELL_maxiter && nlmax != nlmax_old) {
      nlmax = nlmax_old;
      const or_level* L0 = &H->L[0];
      ALL(L0, k) H->ell[0].v[k] = 0.0;
      for (int l = 1; l <= nlmax; ++l) {
        const or_level* L = &H->L[l];
        double* u = H->ell[l].u;
        const int g = L->g;
        for (int m = 0; m <= L->Nr + g - 1; ++m)
          for (int n = 0; n <= L->Nz + g - 1; ++n) {
            const double r0 = L->rmin + (double)m * L->dr;
            const double z0 = L->zmin + (double)n * L->dz;
            const int i = (int)lround((r0 - L->rmin) / L->dr) + 1 - g;
            const int j = (int)lround((z0 - L->zmin) / L->dz) + 1 - g;
            const double aux2 = or_interp(p, L0, H->ell[0].u, r0, z0);
            if (i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz) u[IX(L, i, j)] = aux2;
          }
      }
      reset_clocks(H);
      continue;
    }
    break;
  }
  *res = gres;
  for (int l = 0; l < H->nlev; ++l) {
    H->L[l].t = save[4 * l]; H->L[l].t1 = save[4 * l + 1]; H->L[l].t2 = save[4 * l + 2];
    H->L[l].dt = save[4 * l + 3]; H->L[l].s = ssave[l];
  }
  free(save);
  free(ssave);
}

/* ------------------------------------------------------------------ SOR
 * sor(a, .., j, u, imax, jmax, rjac, uinf, bound) (sor_elliptic.f90:535-802)
 * for the flat Laplacian, on one level: red-black successive
 * over-relaxation with Chebyshev acceleration (the mixed-derivative
 * coefficients of the flat case are zero, so their products add exact
 * zeros and are left out).  The reference's arrays run over 1 .. N+ghost;
 * here the level's own indices i = 1-g .. Nr, j = 1-g .. Nz are used, so
 * its sweep over 2 .. imax-1 is i = 2-g .. Nr-1 (j alike), and its colour
 * (i+j) mod 2 is the same.  robin: the outer and (without equatorial
 * symmetry) lower boundaries follow the Robin condition with uinf = 1 (base
 * level); otherwise every point outside the sweep keeps its value
 * (Dirichlet, fine levels).  Returns the iterations taken (ELL_maxiter when
 * the residual never fell below ELL_epsilon). */
static int sor_level(const oa_params* p, const or_level* L, or_ell_level* E, double rjac, int robin, double* res) {
  const int g = L->g, Nr = L->Nr, Nz = L->Nz;
  const double one = 1.0, half = 0.5, two = 2.0, uinf = 1.0;
  const double dr = L->dr, dz = L->dz;
  double* u = E->u;
  const double nn = (double)(Nr + g) * (double)(Nz + g);
  double omega = 1.0, ganorm = 0.0;
  for (int n = 1; n <= p->ELL_maxiter; ++n) {
    double anorm = 0.0;
    /* sor_e = sor_aux + S5 u^4 at the start of the iteration (:606-611) */
    for (int ipass = 1; ipass <= 2; ++ipass) {
      for (int i = 2 - g; i <= Nr - 1; ++i)
        for (int j = 2 - g; j <= Nz - 1; ++j) {
          if (((i + j) & 1) != (ipass - 1)) continue;
          const int k = IX(L, i, j);
          const double r = L->r[k];
          const double a = (one / dr + half / r) / dr;
          const double b = (one / dr - half / r) / dr;
          const double c = one / (dz * dz);
          const double d = one / (dz * dz);
          const double u2 = u[k] * u[k];
          const double e = (-two * (one / (dr * dr) + one / (dz * dz)) + E->S1[k]) + E->S5[k] * (u2 * u2);
          const double resid = a * u[IX(L, i + 1, j)] + b * u[IX(L, i - 1, j)] + c * u[IX(L, i, j + 1)]
                               + d * u[IX(L, i, j - 1)] + e * u[k] - E->S0[k];
          anorm = anorm + resid * resid;
          u[k] = u[k] - omega * resid / e;
        }
      anorm = sqrt(anorm / nn);
      if (n == 1 && ipass == 1)
        omega = 1.0 / (1.0 - 0.5 * rjac * rjac);
      else
        omega = 1.0 / (1.0 - 0.25 * rjac * rjac * omega);
    }
    ganorm = anorm;
    if (robin) {
      /* outer rho boundary, then the upper and (no equatorial symmetry) the
       * lower z boundary, then the corners (:655-739) */
      double cr = L->r[IX(L, Nr, 1)] / dr, cz;
      for (int j = 2 - g; j <= Nz - 1; ++j) {
        cz = L->z[IX(L, 1, j)] / dz;
        u[IX(L, Nr, j)] = (uinf + 0.5 * cr * (4.0 * u[IX(L, Nr - 1, j)] - u[IX(L, Nr - 2, j)])
                           - 0.5 * cz * (u[IX(L, Nr, j + 1)] - u[IX(L, Nr, j - 1)])) / (1.0 + 1.5 * cr);
      }
      cz = L->z[IX(L, 1, Nz)] / dz;
      for (int i = 2 - g; i <= Nr - 1; ++i) {
        cr = L->r[IX(L, i, 1)] / dr;
        u[IX(L, i, Nz)] = (uinf + 0.5 * cz * (4.0 * u[IX(L, i, Nz - 1)] - u[IX(L, i, Nz - 2)])
                           - 0.5 * cr * (u[IX(L, i + 1, Nz)] - u[IX(L, i - 1, Nz)])) / (1.0 + 1.5 * cz);
      }
      if (!p->eqsym) {
        cz = L->z[IX(L, 1, 1 - g)] / dz;
        for (int i = 2 - g; i <= Nr - 1; ++i) {
          cr = L->r[IX(L, i, 1)] / dr;
          u[IX(L, i, 1 - g)] = (uinf - 0.5 * cz * (4.0 * u[IX(L, i, 2 - g)] - u[IX(L, i, 3 - g)])
                                - 0.5 * cr * (u[IX(L, i + 1, 1 - g)] - u[IX(L, i - 1, 1 - g)])) / (1.0 - 1.5 * cz);
        }
      }
      cr = L->r[IX(L, Nr, Nz)] / dr;
      cz = L->z[IX(L, Nr, Nz)] / dz;
      u[IX(L, Nr, Nz)] = (uinf + 0.5 * cr * (4.0 * u[IX(L, Nr - 1, Nz)] - u[IX(L, Nr - 2, Nz)])
                          + 0.5 * cz * (4.0 * u[IX(L, Nr, Nz - 1)] - u[IX(L, Nr, Nz - 2)])) / (1.0 + 1.5 * (cr + cz));
      if (!p->eqsym) {
        cr = L->r[IX(L, Nr, 1 - g)] / dr;
        cz = L->z[IX(L, Nr, 1 - g)] / dz;
        u[IX(L, Nr, 1 - g)] = (uinf + 0.5 * cr * (4.0 * u[IX(L, Nr - 1, 1 - g)] - u[IX(L, Nr - 2, 1 - g)])
                               - 0.5 * cz * (4.0 * u[IX(L, Nr, 2 - g)] - u[IX(L, Nr, 3 - g)]))
                              / (1.0 + 1.5 * (cr - cz));
      }
    }
    ell_symmetries(p, L, u);
    if (ganorm < p->ELL_epsilon) {
      *res = ganorm;
      return n;
    }
  }
  *res = ganorm;
  return p->ELL_maxiter;
}

/* sor_elliptic('flat', 'one') (sor_elliptic.f90:2-530): u = 1, the base
 * level with Robin boundaries, then each finer level in turn from the
 * interpolation of its parent's solution (initial guess and Dirichlet
 * boundary), then restriction of every level onto its parent, level 1 first
 * (restrict.f90 with all = .false.; no symmetries are applied after it).
 * steps[0]: iterations on the base level, steps[1]: their sum over the finer
 * levels. */
static void sor_elliptic(or_handle* H, int* steps, double* res) {
  const oa_params* p = &H->p;
  for (int l = 0; l < H->nlev; ++l) {
    or_ell_level* E = &H->ell[l];
    ALL((&H->L[l]), k) E->u[k] = 1.0;
  }
  /* Jacobi spectral radius of the base level's flat Laplacian (:140-143) */
  const or_level* L0 = &H->L[0];
  const double smallpi = acos(-1.0);
  const double rjac = (L0->dz / L0->dr * cos(smallpi / (double)(L0->Nr + L0->g))
                       + L0->dr / L0->dz * cos(smallpi / (double)(L0->Nz + L0->g)))
                      / (L0->dz / L0->dr + L0->dr / L0->dz);
  steps[0] = sor_level(p, L0, &H->ell[0], rjac, 1, res);
  steps[1] = 0;
  for (int l = 1; l < H->nlev; ++l) {
    const or_level* L = &H->L[l];
    const or_level* Lc = &H->L[l - 1];
    double* u = H->ell[l].u;
    const int g = L->g;
    for (int m = 0; m <= L->Nr + g - 1; ++m)
      for (int n = 0; n <= L->Nz + g - 1; ++n) {
        const double r0 = L->rmin + (double)m * L->dr;
        const double z0 = L->zmin + (double)n * L->dz;
        const int i = (int)lround((r0 - L->rmin) / L->dr) + 1 - g;
        const int j = (int)lround((z0 - L->zmin) / L->dz) + 1 - g;
        const double aux2 = or_interp(p, Lc, H->ell[l - 1].u, r0, z0);
        if (i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz) u[IX(L, i, j)] = aux2;
      }
    double lres = 0.0;
    steps[1] += sor_level(p, L, &H->ell[l], rjac, 0, &lres);
  }
  for (int l = 1; l < H->nlev; ++l) ell_restrict(p, &H->L[l], &H->L[l - 1], H->ell[l].u, H->ell[l - 1].u);
}

/* Coefficients of the Hamiltonian-constraint equation and its solution:
 * Brill waves (idata_brillwave.f90:88-105, :128-135) and the scalar pulse
 * (idata_scalarpulse.f90:143-178, time-symmetric, smallpi = acos(-1d0)).
 * Needs the scalar field, xi and V already on every level for the pulse.
 * Leaves psi = ell_u in L->psi. */
int or_solve_hamiltonian(or_handle* H) {
  const oa_params* p = &H->p;
  ell_alloc(H);
  const double half = 0.5, one = 1.0, two = 2.0, smallpi = acos(-1.0);
  for (int l = 0; l < H->nlev; ++l) {
    or_level* L = &H->L[l];
    or_ell_level* E = &H->ell[l];
    if (p->idata == OA_IDATA_BRILLWAVE) {
      const double sr = p->brill_sr0, sz = p->brill_sz0;
      ALL(L, k) {
        const double r = L->r[k], z = L->z[k];
        const double q = p->brill_a0 * r * r * exp(-(r / sr) * (r / sr) - (z / sz) * (z / sz));
        E->S1[k] = half * q * (one / (r * r) + two * ((r / (sr * sr)) * (r / (sr * sr)) + (z / (sz * sz)) * (z / (sz * sz)))
                               - 5.0 / (sr * sr) - one / (sz * sz));
        E->S0[k] = 0.0;
        E->S5[k] = 0.0;
      }
    } else {
      const double* xr = L->var[OF_scalar_xi_r];
      const double* xz = L->var[OF_scalar_xi_z];
      ALL(L, k) {
        E->S0[k] = 0.0;
        E->S1[k] = smallpi * (xr[k] * xr[k] + xz[k] * xz[k]);
        E->S5[k] = (p->scalarpotential == OA_POT_NONE) ? 0.0 : two * smallpi * L->scalar_V[k];
      }
    }
  }
  if (p->ELL_solver == OA_ELL_SOR)
    sor_elliptic(H, H->ell_steps, &H->ell_res);
  else
    wave_elliptic(H, H->ell_steps, &H->ell_res);
  for (int l = 0; l < H->nlev; ++l) {
    or_level* L = &H->L[l];
    memcpy(L->psi, H->ell[l].u, sizeof(double) * (size_t)L->n);
  }
  return 0;
}

/* ctypes: iterations of the two solves and the last residual */
void or_ell_stats(or_handle* H, int* steps0, int* steps1, double* res) {
  *steps0 = H->ell_steps[0];
  *steps1 = H->ell_steps[1];
  *res = H->ell_res;
}

/* ctypes: an elliptic array of a level (u, v, S0, S1, S5), NULL if the solve has not run */
double* or_ell_field(or_handle* H, int lev, const char* name) {
  if (!H->ell || lev < 0 || lev >= H->nlev) return NULL;
  or_ell_level* E = &H->ell[lev];
  if (!strcmp(name, "ell_u")) return E->u;
  if (!strcmp(name, "ell_v")) return E->v;
  if (!strcmp(name, "ell_S0")) return E->S0;
  if (!strcmp(name, "ell_S1")) return E->S1;
  if (!strcmp(name, "ell_S5")) return E->S5;
  return NULL;
}
