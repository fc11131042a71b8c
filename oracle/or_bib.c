/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see or_internal.h).
 *
 * Box-in-box transfer for box 0: interp, finebound (all evolving arrays) and
 * restrict of the reference.
 */
#include <math.h>
#include <stdio.h>
#include "or_internal.h"

static inline double R_(const or_level* L, int i) { return L->r[IX(L, i, 0)]; }
static inline double Z_(const or_level* L, int j) { return L->z[IX(L, 0, j)]; }

/* interp: interp.f90:2-270 (single process) */
double or_interp(const oa_params* p, const or_level* L, const double* f, double r0, double z0) {
  const double small = 1.0e-10;
  const int g = L->g;
  if (r0 < L->rmin - small || r0 > L->rmax || z0 < L->zmin - small || z0 > L->zmax) {
    fprintf(stderr, "oracle interp: point (%g,%g) outside the grid\n", r0, z0);
    return NAN;
  }
  const int i0 = (int)((r0 - R_(L, 1 - g)) / L->dr) + 1 - g;
  const double deltar = r0 - R_(L, i0);
  const int j0 = (int)((z0 - Z_(L, 1 - g)) / L->dz) + 1 - g;
  const double deltaz = z0 - Z_(L, j0);
  if (deltar == 0.0 && deltaz == 0.0) return f[IX(L, i0, j0)];
  const int forcelinear = ((L->rmax - r0) < L->dr) || ((L->zmax - z0) < L->dz) ||
                          ((r0 - L->rmin) < L->dr) || ((z0 - L->zmin) < L->dz);
  double fa[4][4];
  if (forcelinear) {
    for (int i = 0; i <= 1; ++i)
      for (int j = 0; j <= 1; ++j) fa[i + 1][j + 1] = f[IX(L, i0 + i, j0 + j)];
  } else {
    for (int i = -1; i <= 2; ++i)
      for (int j = -1; j <= 2; ++j) fa[i + 1][j + 1] = f[IX(L, i0 + i, j0 + j)];
  }
  const double auxr = deltar / L->dr, auxz = deltaz / L->dz;
  if (p->interporder == OA_INTERP_BILINEAR || forcelinear) {
    return fa[1][1] * (1.0 - auxr) * (1.0 - auxz) + fa[2][1] * auxr * (1.0 - auxz)
           + fa[1][2] * (1.0 - auxr) * auxz + fa[2][2] * auxr * auxz;
  }
  double out = 0.0;
  for (int i = -1; i <= 2; ++i)
    for (int j = -1; j <= 2; ++j) {
      double aux = 1.0;
      for (int m = -1; m <= 2; ++m) {
        if (m == i) continue;
        aux = aux * (auxr - (double)m) / (double)(i - m);
      }
      for (int n = -1; n <= 2; ++n) {
        if (n == j) continue;
        aux = aux * (auxz - (double)n) / (double)(j - n);
      }
      out = out + aux * fa[i + 1][j + 1];
    }
  return out;
}

/* time interpolation of one injected value, generated bound_interp.inc
 * (prl/arrays.pl:1330-1397).  b0/b1 are the strip values at t0 and tm1. */
static double time_interp(int border, double tl, double tp, double t0, double tm1,
                          double aux2, double var_p, double b0, double b1) {
  if (border == 1) return (tl - t0) / (tp - t0) * aux2 + (tl - tp) / (t0 - tp) * var_p;
  return (tl - t0) * (tl - tm1) / ((tp - t0) * (tp - tm1)) * aux2
         + (tl - tp) * (tl - tm1) / ((t0 - tp) * (t0 - tm1)) * b0
         + (tl - tp) * (tl - t0) / ((tm1 - tp) * (tm1 - t0)) * b1;
}

/* finebound(0, level, dtw, .true.): finebound.f90:70-400 */
void or_finebound(or_handle* Hd, int level, double dtw) {
  const oa_params* p = &Hd->p;
  or_level* L = &Hd->L[level];
  const or_level* Lc = &Hd->L[level - 1];
  const int g = L->g, ni = L->ni, nj = L->nj;
  const double tp = Lc->t, t0 = L->t, tm1 = L->t1, tl = t0 + dtw;
  const int imax = (tp == tl) ? g + 2 : g - 1;
  const int jmax = imax;
  const int border = (tm1 == 0.0) ? 1 : 2;
  const int same = (tp == tl);   /* strip coefficients are exactly zero */

  /* upper r boundary */
  for (int m = 0; m <= imax; ++m)
    for (int n = 0; n <= L->Nz + g - 1; ++n) {
      const double r0 = L->rmax - (double)m * L->dr;
      const double z0 = L->zmin + (double)n * L->dz;
      const int i = (int)lround((r0 - R_(L, 1 - g)) / L->dr) + 1 - g;
      const int j = (int)lround((z0 - Z_(L, 1 - g)) / L->dz) + 1 - g;
      const int flag1 = (i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz);
      for (int f = 0; f < OF_NFIELDS; ++f) {
        if (!L->present[f]) continue;
        const double aux2 = or_interp(p, Lc, Lc->var[f], r0, z0);
        if (!flag1) continue;
        const int bk = m + g * (j + g - 1);
        const double b0 = (same || m >= g) ? 0.0 : L->b_rR[f][bk];
        const double b1 = (same || m >= g) ? 0.0 : L->b_rR[f][bk + g * nj];
        L->var[f][IX(L, i, j)] = time_interp(border, tl, tp, t0, tm1, aux2,
                                             L->var_p[f][IX(L, i, j)], b0, b1);
      }
    }
  /* upper z boundary */
  for (int n = 0; n <= jmax; ++n)
    for (int m = 0; m <= L->Nr + g - 1; ++m) {
      const double r0 = L->rmin + (double)m * L->dr;
      const double z0 = L->zmax - (double)n * L->dz;
      const int i = (int)lround((r0 - R_(L, 1 - g)) / L->dr) + 1 - g;
      const int j = (int)lround((z0 - Z_(L, 1 - g)) / L->dz) + 1 - g;
      const int flag1 = (i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz);
      for (int f = 0; f < OF_NFIELDS; ++f) {
        if (!L->present[f]) continue;
        const double aux2 = or_interp(p, Lc, Lc->var[f], r0, z0);
        if (!flag1) continue;
        const int bk = (i + g - 1) + ni * n;
        const double b0 = (same || n >= g) ? 0.0 : L->b_zR[f][bk];
        const double b1 = (same || n >= g) ? 0.0 : L->b_zR[f][bk + ni * g];
        L->var[f][IX(L, i, j)] = time_interp(border, tl, tp, t0, tm1, aux2,
                                             L->var_p[f][IX(L, i, j)], b0, b1);
      }
    }
  /* lower z boundary (no equatorial symmetry) */
  if (!p->eqsym)
    for (int n = 0; n <= jmax; ++n)
      for (int m = 0; m <= L->Nr + g - 1; ++m) {
        const double r0 = L->rmin + (double)m * L->dr;
        const double z0 = L->zmin + (double)n * L->dz;
        const int i = (int)lround((r0 - R_(L, 1 - g)) / L->dr) + 1 - g;
        const int j = (int)lround((z0 - Z_(L, 1 - g)) / L->dz) + 1 - g;
        const int flag1 = (i >= 1 - g && i <= L->Nr && j >= 1 - g && j <= L->Nz);
        for (int f = 0; f < OF_NFIELDS; ++f) {
          if (!L->present[f]) continue;
          const double aux2 = or_interp(p, Lc, Lc->var[f], r0, z0);
          if (!flag1) continue;
          const int bk = (i + g - 1) + ni * n;
          const double b0 = (same || n >= g) ? 0.0 : L->b_zL[f][bk];
          const double b1 = (same || n >= g) ? 0.0 : L->b_zL[f][bk + ni * g];
          L->var[f][IX(L, i, j)] = time_interp(border, tl, tp, t0, tm1, aux2,
                                               L->var_p[f][IX(L, i, j)], b0, b1);
        }
      }
}

/* restrict(0, level, .true.): restrict.f90:88-162 */
void or_restrict(or_handle* Hd, int level) {
  const oa_params* p = &Hd->p;
  const or_level* L = &Hd->L[level];
  or_level* Lc = &Hd->L[level - 1];
  const int g = L->g;
  const double deltar = (double)(2 * g) * L->dr + 1.0e-10;
  const double deltaz = (double)(2 * g) * L->dz + 1.0e-10;
  for (int n = 0; n <= Lc->Nz + g - 1; ++n)
    for (int m = 0; m <= Lc->Nr + g - 1; ++m) {
      const double r0 = Lc->rmin + (double)m * Lc->dr;
      const double z0 = Lc->zmin + (double)n * Lc->dz;
      const double rmin = (L->rmin < 0.0) ? 0.0 : L->rmin + deltar;
      const double zmin = (p->eqsym && L->zmin < 0.0) ? 0.0 : L->zmin + deltaz;
      const double rmax = L->rmax - deltar;
      const double zmax = L->zmax - deltaz;
      if (r0 > rmin && r0 < rmax && z0 > zmin && z0 < zmax) {
        const int i = (int)lround((r0 - R_(Lc, 1 - g)) / Lc->dr) + 1 - g;
        const int j = (int)lround((z0 - Z_(Lc, 1 - g)) / Lc->dz) + 1 - g;
        const int flag1 = (i >= 1 - g && i <= Lc->Nr && j >= 1 - g && j <= Lc->Nz);
        for (int f = 0; f < OF_NFIELDS; ++f) {
          if (!L->present[f]) continue;
          const double aux2 = or_interp(p, L, L->var[f], r0, z0);
          if (flag1) Lc->var[f][IX(Lc, i, j)] = aux2;
        }
      }
    }
}
