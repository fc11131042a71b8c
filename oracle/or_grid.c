/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY (see or_internal.h).
 *
 * Finite differences, Kreiss-Oliger dissipation, radiative boundaries and the
 * generated per-array routines (symmetries, saveold, accumulate, update) of
 * the reference, restated whole-array in C.
 */
#include <math.h>
#include <string.h>
#include "or_internal.h"

const or_fielddesc or_fields[OF_NFIELDS] = {
  /* arrays.config:154,165,311,329,336,355-358,568,585-588,716-717,768,782,850,
   * 212-213,234-235,263,1029,1040-1041,1045 */
  {"alpha", +1, +1}, {"dtalpha", +1, +1}, {"phi", +1, +1}, {"dtphi", +1, +1},
  {"chi", +1, +1},
  {"A", +1, +1}, {"B", +1, +1}, {"H", +1, +1}, {"C", +1, -1},
  {"trK", +1, +1}, {"KTA", +1, +1}, {"KTB", +1, +1}, {"KTH", +1, +1}, {"KTC", +1, -1},
  {"Delta_r", -1, +1}, {"Delta_z", +1, -1}, {"lambda", +1, +1}, {"Alambda", +1, +1},
  {"z4theta", +1, +1},
  {"beta_r", -1, +1}, {"beta_z", +1, -1}, {"dtbeta_r", -1, +1}, {"dtbeta_z", +1, -1},
  {"fdriver", +1, +1},
  {"scalar_phi", +1, +1}, {"scalar_xi_r", -1, +1}, {"scalar_xi_z", +1, -1},
  {"scalar_pi", +1, +1},
};

static inline double sgn(int s) { return s >= 0 ? 1.0 : -1.0; }

/* diff1r: derivatives.f90:15-117 */
void or_diff1r(const oa_params* p, const or_level* L, const double* f, int symr, double* out) {
  const int Nr = L->Nr, g = L->g;
  const double idr = 1.0 / L->dr, hidr = 0.5 * idr;
  for (int j = 1 - g; j <= L->Nz; ++j) {
#define F(ii) f[IX(L, ii, j)]
    if (p->order == OA_ORDER_TWO) {
      for (int i = 2 - g; i <= Nr - 1; ++i) out[IX(L, i, j)] = hidr * (F(i + 1) - F(i - 1));
      int i = Nr;
      out[IX(L, i, j)] = hidr * (3.0 * F(Nr) - 4.0 * F(Nr - 1) + F(Nr - 2));
    } else {
      for (int i = 1; i <= Nr - 2; ++i)
        out[IX(L, i, j)] = idr * (8.0 * (F(i + 1) - F(i - 1)) - (F(i + 2) - F(i - 2))) / 12.0;
      int i = Nr - 1;
      out[IX(L, i, j)] = idr * (3.0 * F(i + 1) + 10.0 * F(i) - 18.0 * F(i - 1)
                                + 6.0 * F(i - 2) - F(i - 3)) / 12.0;
      i = Nr;
      out[IX(L, i, j)] = idr * (25.0 * F(i) - 48.0 * F(i - 1) + 36.0 * F(i - 2)
                                - 16.0 * F(i - 3) + 3.0 * F(i - 4)) / 12.0;
    }
#undef F
    for (int i = 1; i <= g; ++i) out[IX(L, 1 - i, j)] = -sgn(symr) * out[IX(L, i, j)];
  }
}

/* diff1z: derivatives.f90:123-241 */
void or_diff1z(const oa_params* p, const or_level* L, const double* f, int symz, double* out) {
  const int Nz = L->Nz, g = L->g;
  const double idz = 1.0 / L->dz, hidz = 0.5 * idz;
  for (int i = 1 - g; i <= L->Nr; ++i) {
#define F(jj) f[IX(L, i, jj)]
#define O(jj) out[IX(L, i, jj)]
    if (p->order == OA_ORDER_TWO) {
      for (int j = 2 - g; j <= Nz - 1; ++j) O(j) = hidz * (F(j + 1) - F(j - 1));
      int j = 1 - g;
      O(j) = -hidz * (3.0 * F(j) - 4.0 * F(j + 1) + F(j + 2));
      j = Nz;
      O(j) = +hidz * (3.0 * F(j) - 4.0 * F(j - 1) + F(j - 2));
    } else {
      for (int j = 3 - g; j <= Nz - 2; ++j)
        O(j) = idz * (8.0 * (F(j + 1) - F(j - 1)) - (F(j + 2) - F(j - 2))) / 12.0;
      int j = 2 - g;
      O(j) = -idz * (3.0 * F(j - 1) + 10.0 * F(j) - 18.0 * F(j + 1) + 6.0 * F(j + 2) - F(j + 3)) / 12.0;
      j = 1 - g;
      O(j) = -idz * (25.0 * F(j) - 48.0 * F(j + 1) + 36.0 * F(j + 2) - 16.0 * F(j + 3) + 3.0 * F(j + 4)) / 12.0;
      j = Nz - 1;
      O(j) = +idz * (3.0 * F(j + 1) + 10.0 * F(j) - 18.0 * F(j - 1) + 6.0 * F(j - 2) - F(j - 3)) / 12.0;
      j = Nz;
      O(j) = +idz * (25.0 * F(j) - 48.0 * F(j - 1) + 36.0 * F(j - 2) - 16.0 * F(j - 3) + 3.0 * F(j - 4)) / 12.0;
    }
    if (p->eqsym)
      for (int j = 1; j <= g; ++j) O(1 - j) = -sgn(symz) * O(j);
#undef F
#undef O
  }
}

/* diff2r: derivatives.f90:250-359 */
void or_diff2r(const oa_params* p, const or_level* L, const double* f, int symr, double* out) {
  const int Nr = L->Nr, g = L->g;
  const double idr = 1.0 / L->dr, idr2 = idr * idr;
  for (int j = 1 - g; j <= L->Nz; ++j) {
#define F(ii) f[IX(L, ii, j)]
#define O(ii) out[IX(L, ii, j)]
    if (p->order == OA_ORDER_TWO) {
      for (int i = 2 - g; i <= Nr - 1; ++i) O(i) = idr2 * (F(i + 1) - 2.0 * F(i) + F(i - 1));
      int i = 1 - g;
      O(i) = idr2 * (2.0 * F(i) - 5.0 * F(i + 1) + 4.0 * F(i + 2) - F(i + 3));
      i = Nr;
      O(i) = idr2 * (2.0 * F(i) - 5.0 * F(i - 1) + 4.0 * F(i - 2) - F(i - 3));
    } else {
      for (int i = 3 - g; i <= Nr - 2; ++i)
        O(i) = -idr2 * (30.0 * F(i) - 16.0 * (F(i + 1) + F(i - 1)) + F(i + 2) + F(i - 2)) / 12.0;
      int i = Nr - 1;
      O(i) = idr2 * (10.0 * F(i + 1) - 15.0 * F(i) - 4.0 * F(i - 1) + 14.0 * F(i - 2)
                     - 6.0 * F(i - 3) + F(i - 4)) / 12.0;
      i = Nr;
      O(i) = idr2 * (45.0 * F(i) - 154.0 * F(i - 1) + 214.0 * F(i - 2) - 156.0 * F(i - 3)
                     + 61.0 * F(i - 4) - 10.0 * F(i - 5)) / 12.0;
    }
    for (int i = 1; i <= g; ++i) O(1 - i) = sgn(symr) * O(i);
#undef F
#undef O
  }
}

/* diff2z: derivatives.f90:367-488 */
void or_diff2z(const oa_params* p, const or_level* L, const double* f, int symz, double* out) {
  const int Nz = L->Nz, g = L->g;
  const double idz = 1.0 / L->dz, idz2 = idz * idz;
  for (int i = 1 - g; i <= L->Nr; ++i) {
#define F(jj) f[IX(L, i, jj)]
#define O(jj) out[IX(L, i, jj)]
    if (p->order == OA_ORDER_TWO) {
      for (int j = 2 - g; j <= Nz - 1; ++j) O(j) = idz2 * (F(j + 1) - 2.0 * F(j) + F(j - 1));
      int j = 1 - g;
      O(j) = idz2 * (2.0 * F(j) - 5.0 * F(j + 1) + 4.0 * F(j + 2) - F(j + 3));
      j = Nz;
      O(j) = idz2 * (2.0 * F(j) - 5.0 * F(j - 1) + 4.0 * F(j - 2) - F(j - 3));
    } else {
      for (int j = 3 - g; j <= Nz - 2; ++j)
        O(j) = -idz2 * (30.0 * F(j) - 16.0 * (F(j + 1) + F(j - 1)) + F(j + 2) + F(j - 2)) / 12.0;
      int j = 2 - g;
      O(j) = idz2 * (10.0 * F(j - 1) - 15.0 * F(j) - 4.0 * F(j + 1) + 14.0 * F(j + 2)
                     - 6.0 * F(j + 3) + F(j + 4)) / 12.0;
      j = 1 - g;
      O(j) = idz2 * (45.0 * F(j) - 154.0 * F(j + 1) + 214.0 * F(j + 2) - 156.0 * F(j + 3)
                     + 61.0 * F(j + 4) - 10.0 * F(j + 5)) / 12.0;
      j = Nz - 1;
      O(j) = idz2 * (10.0 * F(j + 1) - 15.0 * F(j) - 4.0 * F(j - 1) + 14.0 * F(j - 2)
                     - 6.0 * F(j - 3) + F(j - 4)) / 12.0;
      j = Nz;
      O(j) = idz2 * (45.0 * F(j) - 154.0 * F(j - 1) + 214.0 * F(j - 2) - 156.0 * F(j - 3)
                     + 61.0 * F(j - 4) - 10.0 * F(j - 5)) / 12.0;
    }
    if (p->eqsym)
      for (int j = 1; j <= g; ++j) O(1 - j) = sgn(symz) * O(j);
#undef F
#undef O
  }
}

/* diff2rz: derivatives.f90:496-732.  The 2nd-order edge formulas of the
 * 4th-order scheme are written out as the reference writes them, including
 * which loop overwrites which at the shared points. */
void or_diff2rz(const oa_params* p, const or_level* L, const double* f, int symr, int symz,
                double* out) {
  const int Nr = L->Nr, Nz = L->Nz, g = L->g;
  const double idrz = (1.0 / L->dr) * (1.0 / L->dz);
#define F(ii, jj) f[IX(L, ii, jj)]
#define O(ii, jj) out[IX(L, ii, jj)]
  /* one-sided edge pieces shared by both orders */
#define RLOW(i, j) (-0.25 * idrz * ((3.0 * F(i, j + 1) - 4.0 * F(i + 1, j + 1) + F(i + 2, j + 1)) \
                                  - (3.0 * F(i, j - 1) - 4.0 * F(i + 1, j - 1) + F(i + 2, j - 1))))
#define RHIGH(i, j) (+0.25 * idrz * ((3.0 * F(i, j + 1) - 4.0 * F(i - 1, j + 1) + F(i - 2, j + 1)) \
                                   - (3.0 * F(i, j - 1) - 4.0 * F(i - 1, j - 1) + F(i - 2, j - 1))))
#define ZLOW(i, j) (-0.25 * idrz * ((3.0 * F(i + 1, j) - 4.0 * F(i + 1, j + 1) + F(i + 1, j + 2)) \
                                  - (3.0 * F(i - 1, j) - 4.0 * F(i - 1, j + 1) + F(i - 1, j + 2))))
#define ZHIGH(i, j) (+0.25 * idrz * ((3.0 * F(i + 1, j) - 4.0 * F(i + 1, j - 1) + F(i + 1, j - 2)) \
                                   - (3.0 * F(i - 1, j) - 4.0 * F(i - 1, j - 1) + F(i - 1, j - 2))))
  if (p->order == OA_ORDER_TWO) {
    for (int i = 2 - g; i <= Nr - 1; ++i)
      for (int j = 2 - g; j <= Nz - 1; ++j)
        O(i, j) = 0.25 * idrz * (F(i + 1, j + 1) + F(i - 1, j - 1) - F(i + 1, j - 1) - F(i - 1, j + 1));
    for (int j = 2 - g; j <= Nz - 1; ++j) {
      O(1 - g, j) = RLOW(1 - g, j);
      O(Nr, j) = RHIGH(Nr, j);
    }
    for (int i = 2 - g; i <= Nr - 1; ++i) {
      O(i, 1 - g) = ZLOW(i, 1 - g);
      O(i, Nz) = ZHIGH(i, Nz);
    }
  } else {
    for (int j = 3 - g; j <= Nz - 2; ++j)
      for (int i = 3 - g; i <= Nr - 2; ++i)
        O(i, j) = idrz * ((F(i + 1, j + 1) + F(i - 1, j - 1) - F(i + 1, j - 1) - F(i - 1, j + 1)) / 3.0
                          - (F(i + 2, j + 2) + F(i - 2, j - 2) - F(i + 2, j - 2) - F(i - 2, j + 2)) / 48.0);
    for (int j = 2 - g; j <= Nz - 1; ++j) {
      int i = 2 - g;
      O(i, j) = +0.25 * idrz * (F(i + 1, j + 1) + F(i - 1, j - 1) - F(i + 1, j - 1) - F(i - 1, j + 1));
      O(1 - g, j) = RLOW(1 - g, j);
      i = Nr - 1;
      O(i, j) = +0.25 * idrz * (F(i + 1, j + 1) + F(i - 1, j - 1) - F(i + 1, j - 1) - F(i - 1, j + 1));
      O(Nr, j) = RHIGH(Nr, j);
    }
    for (int i = 2 - g; i <= Nr - 1; ++i) {
      int j = 2 - g;
      O(i, j) = +0.25 * idrz * (F(i + 1, j + 1) + F(i - 1, j - 1) - F(i - 1, j + 1) - F(i + 1, j - 1));
      O(i, 1 - g) = ZLOW(i, 1 - g);
      j = Nz - 1;
      O(i, j) = +0.25 * idrz * (F(i + 1, j + 1) + F(i - 1, j - 1) - F(i - 1, j + 1) - F(i + 1, j - 1));
      O(i, Nz) = ZHIGH(i, Nz);
    }
  }
  /* corners (both orders) */
  {
    int i = 1 - g, j = 1 - g;
    O(i, j) = +0.25 * idrz * (9.0 * F(i, j) + 16.0 * F(i + 1, j + 1) + F(i + 2, j + 2)
                              - 12.0 * (F(i, j + 1) + F(i + 1, j)) + 3.0 * (F(i, j + 2) + F(i + 2, j))
                              - 4.0 * (F(i + 1, j + 2) + F(i + 2, j + 1)));
    i = 1 - g; j = Nz;
    O(i, j) = -0.25 * idrz * (9.0 * F(i, j) + 16.0 * F(i + 1, j - 1) + F(i + 2, j - 2)
                              - 12.0 * (F(i + 1, j) + F(i, j - 1)) + 3.0 * (F(i + 2, j) + F(i, j - 2))
                              - 4.0 * (F(i + 2, j - 1) + F(i + 1, j - 2)));
    i = Nr; j = 1 - g;
    O(i, j) = -0.25 * idrz * (9.0 * F(i, j) + 16.0 * F(i - 1, j + 1) + F(i - 2, j + 2)
                              - 12.0 * (F(i, j + 1) + F(i - 1, j)) + 3.0 * (F(i, j + 2) + F(i - 2, j))
                              - 4.0 * (F(i - 1, j + 2) + F(i - 2, j + 1)));
    i = Nr; j = Nz;
    O(i, j) = +0.25 * idrz * (9.0 * F(i, j) + 16.0 * F(i - 1, j - 1) + F(i - 2, j - 2)
                              - 12.0 * (F(i, j - 1) + F(i - 1, j)) + 3.0 * (F(i, j - 2) + F(i - 2, j))
                              - 4.0 * (F(i - 1, j - 2) + F(i - 2, j - 1)));
  }
  /* symmetries: z first, then r (derivatives.f90:702-714) */
  if (p->eqsym)
    for (int i = 1 - g; i <= Nr; ++i)
      for (int j = 1; j <= g; ++j) O(i, 1 - j) = -sgn(symz) * O(i, j);
  for (int j = 1 - g; j <= Nz; ++j)
    for (int i = 1; i <= g; ++i) O(1 - i, j) = -sgn(symr) * O(i, j);
#undef F
#undef O
#undef RLOW
#undef RHIGH
#undef ZLOW
#undef ZHIGH
}

/* diffadvr: derivadvect.f90:15-206 (fourth and second order upwinding) */
void or_diffadvr(const oa_params* p, const or_level* L, const double* f, int symr,
                 const double* beta_r, double* out) {
  const int Nr = L->Nr, g = L->g;
  const double idr = 1.0 / L->dr, hidr = 0.5 * idr;
  for (int j = 1 - g; j <= L->Nz; ++j) {
#define F(ii) f[IX(L, ii, j)]
#define O(ii) out[IX(L, ii, j)]
#define BR(ii) beta_r[IX(L, ii, j)]
    if (p->order == OA_ORDER_TWO) {
      for (int i = 3 - g; i <= Nr - 2; ++i) {
        if (BR(i) <= 0.0) O(i) = +hidr * (3.0 * F(i) - 4.0 * F(i - 1) + F(i - 2));
        else O(i) = -hidr * (3.0 * F(i) - 4.0 * F(i + 1) + F(i + 2));
      }
      int i = 2 - g;
      if (BR(i) <= 0.0) O(i) = hidr * (F(i + 1) - F(i - 1));
      else O(i) = -hidr * (3.0 * F(i) - 4.0 * F(i + 1) + F(i + 2));
      i = 1 - g;
      O(i) = -hidr * (3.0 * F(i) - 4.0 * F(i + 1) + F(i + 2));
      i = Nr - 1;
      if (BR(i) <= 0.0) O(i) = hidr * (3.0 * F(i) - 4.0 * F(i - 1) + F(i - 2));
      else O(i) = hidr * (F(i + 1) - F(i - 1));
      i = Nr;
      O(i) = hidr * (3.0 * F(i) - 4.0 * F(i - 1) + F(i - 2));
    } else {
      for (int i = 4 - g; i <= Nr - 3; ++i) {
        if (BR(i) <= 0.0)
          O(i) = +0.25 * idr * (3.0 * F(i + 1) + 10.0 * F(i) - 18.0 * F(i - 1) + 6.0 * F(i - 2) - F(i - 3)) / 3.0;
        else
          O(i) = -0.25 * idr * (3.0 * F(i - 1) + 10.0 * F(i) - 18.0 * F(i + 1) + 6.0 * F(i + 2) - F(i + 3)) / 3.0;
      }
      int i = 3 - g;
      if (BR(i) <= 0.0)
        O(i) = 0.25 * idr * (8.0 * (F(i + 1) - F(i - 1)) - (F(i + 2) - F(i - 2))) / 3.0;
      else
        O(i) = -0.25 * idr * (3.0 * F(i - 1) + 10.0 * F(i) - 18.0 * F(i + 1) + 6.0 * F(i + 2) - F(i + 3)) / 3.0;
      i = 2 - g;
      O(i) = -0.25 * idr * (3.0 * F(i - 1) + 10.0 * F(i) - 18.0 * F(i + 1) + 6.0 * F(i + 2) - F(i + 3)) / 3.0;
      i = 1 - g;
      O(i) = -0.25 * idr * (25.0 * F(i) - 48.0 * F(i + 1) + 36.0 * F(i + 2) - 16.0 * F(i + 3) + 3.0 * F(i + 4)) / 3.0;
      i = Nr - 2;
      if (BR(i) <= 0.0)
        O(i) = 0.25 * idr * (3.0 * F(i + 1) + 10.0 * F(i) - 18.0 * F(i - 1) + 6.0 * F(i - 2) - F(i - 3)) / 3.0;
      else
        O(i) = 0.25 * idr * (8.0 * (F(i + 1) - F(i - 1)) - (F(i + 2) - F(i - 2))) / 3.0;
      i = Nr - 1;
      O(i) = 0.25 * idr * (3.0 * F(i + 1) + 10.0 * F(i) - 18.0 * F(i - 1) + 6.0 * F(i - 2) - F(i - 3)) / 3.0;
      i = Nr;
      O(i) = 0.25 * idr * (25.0 * F(i) - 48.0 * F(i - 1) + 36.0 * F(i - 2) - 16.0 * F(i - 3) + 3.0 * F(i - 4)) / 3.0;
    }
    for (int i = 1; i <= g; ++i) O(1 - i) = -(double)symr * O(i);
#undef F
#undef O
#undef BR
  }
}

/* diffadvz: derivadvect.f90:214-405.  Note the reference's branch at
 * j = 3-ghost takes the right-biased stencil for beta_z <= 0, the opposite
 * of diffadvr at i = 3-ghost; it is reproduced as written. */
void or_diffadvz(const oa_params* p, const or_level* L, const double* f, int symz,
                 const double* beta_z, double* out) {
  const int Nz = L->Nz, g = L->g;
  const double idz = 1.0 / L->dz, hidz = 0.5 * idz;
  for (int i = 1 - g; i <= L->Nr; ++i) {
#define F(jj) f[IX(L, i, jj)]
#define O(jj) out[IX(L, i, jj)]
#define BZ(jj) beta_z[IX(L, i, jj)]
    if (p->order == OA_ORDER_TWO) {
      for (int j = 3 - g; j <= Nz - 2; ++j) {
        if (BZ(j) <= 0.0) O(j) = +hidz * (3.0 * F(j) - 4.0 * F(j - 1) + F(j - 2));
        else O(j) = -hidz * (3.0 * F(j) - 4.0 * F(j + 1) + F(j + 2));
      }
      int j = 2 - g;
      if (BZ(j) <= 0.0) O(j) = hidz * (F(j + 1) - F(j - 1));
      else O(j) = -hidz * (3.0 * F(j) - 4.0 * F(j + 1) + F(j + 2));
      j = 1 - g;
      O(j) = -hidz * (3.0 * F(j) - 4.0 * F(j + 1) + F(j + 2));
      j = Nz - 1;
      if (BZ(j) <= 0.0) O(j) = hidz * (3.0 * F(j) - 4.0 * F(j - 1) + F(j - 2));
      else O(j) = hidz * (F(j + 1) - F(j - 1));
      j = Nz;
      O(j) = hidz * (3.0 * F(j) - 4.0 * F(j - 1) + F(j - 2));
    } else {
      for (int j = 4 - g; j <= Nz - 3; ++j) {
        if (BZ(j) <= 0.0)
          O(j) = +0.25 * idz * (3.0 * F(j + 1) + 10.0 * F(j) - 18.0 * F(j - 1) + 6.0 * F(j - 2) - F(j - 3)) / 3.0;
        else
          O(j) = -0.25 * idz * (3.0 * F(j - 1) + 10.0 * F(j) - 18.0 * F(j + 1) + 6.0 * F(j + 2) - F(j + 3)) / 3.0;
      }
      int j = 3 - g;
      if (BZ(j) <= 0.0)
        O(j) = -0.25 * idz * (3.0 * F(j - 1) + 10.0 * F(j) - 18.0 * F(j + 1) + 6.0 * F(j + 2) - F(j + 3)) / 3.0;
      else
        O(j) = 0.25 * idz * (8.0 * (F(j + 1) - F(j - 1)) - (F(j + 2) - F(j - 2))) / 3.0;
      j = 2 - g;
      O(j) = -0.25 * idz * (3.0 * F(j - 1) + 10.0 * F(j) - 18.0 * F(j + 1) + 6.0 * F(j + 2) - F(j + 3)) / 3.0;
      j = 1 - g;
      O(j) = -0.25 * idz * (25.0 * F(j) - 48.0 * F(j + 1) + 36.0 * F(j + 2) - 16.0 * F(j + 3) + 3.0 * F(j + 4)) / 3.0;
      j = Nz - 2;
      if (BZ(j) <= 0.0)
        O(j) = 0.25 * idz * (3.0 * F(j + 1) + 10.0 * F(j) - 18.0 * F(j - 1) + 6.0 * F(j - 2) - F(j - 3)) / 3.0;
      else
        O(j) = 0.25 * idz * (8.0 * (F(j + 1) - F(j - 1)) - (F(j + 2) - F(j - 2))) / 3.0;
      j = Nz - 1;
      O(j) = 0.25 * idz * (3.0 * F(j + 1) + 10.0 * F(j) - 18.0 * F(j - 1) + 6.0 * F(j - 2) - F(j - 3)) / 3.0;
      j = Nz;
      O(j) = 0.25 * idz * (25.0 * F(j) - 48.0 * F(j - 1) + 36.0 * F(j - 2) - 16.0 * F(j - 3) + 3.0 * F(j - 4)) / 3.0;
    }
    if (p->eqsym)
      for (int j = 1; j <= g; ++j) O(1 - j) = -(double)symz * O(j);
#undef F
#undef O
#undef BZ
  }
}

/* dissipation: dissipation.f90:2-249 (the mixed operators are commented out
 * in the reference and are absent here too). */
void or_dissipation(const oa_params* p, const or_level* L, const double* ev, double* src,
                    int symr, int symz, double diss) {
  const int Nr = L->Nr, Nz = L->Nz, g = L->g;
  const double idt = 1.0 / L->dt;
#define E(ii, jj) ev[IX(L, ii, jj)]
  if (p->order == OA_ORDER_TWO) {
    for (int i = 3 - g; i <= Nr - 2; ++i)
      for (int j = 1 - g; j <= Nz; ++j)
        src[IX(L, i, j)] = src[IX(L, i, j)] - diss * idt * (6.0 * E(i, j) - 4.0 * (E(i + 1, j) + E(i - 1, j))
                                                           + (E(i + 2, j) + E(i - 2, j))) / 6.0;
    for (int j = 3 - g; j <= Nz - 2; ++j)
      for (int i = 1 - g; i <= Nr; ++i)
        src[IX(L, i, j)] = src[IX(L, i, j)] - diss * idt * (6.0 * E(i, j) - 4.0 * (E(i, j + 1) + E(i, j - 1))
                                                           + (E(i, j + 2) + E(i, j - 2))) / 6.0;
  } else {
    for (int i = 4 - g; i <= Nr - 3; ++i)
      for (int j = 1 - g; j <= Nz; ++j)
        src[IX(L, i, j)] = src[IX(L, i, j)] - diss * idt * (20.0 * E(i, j) - 15.0 * (E(i + 1, j) + E(i - 1, j))
                                                           + 6.0 * (E(i + 2, j) + E(i - 2, j))
                                                           - (E(i + 3, j) + E(i - 3, j))) / 20.0;
    for (int j = 4 - g; j <= Nz - 3; ++j)
      for (int i = 1 - g; i <= Nr; ++i)
        src[IX(L, i, j)] = src[IX(L, i, j)] - diss * idt * (20.0 * E(i, j) - 15.0 * (E(i, j + 1) + E(i, j - 1))
                                                           + 6.0 * (E(i, j + 2) + E(i, j - 2))
                                                           - (E(i, j + 3) + E(i, j - 3))) / 20.0;
  }
#undef E
  if (p->eqsym)
    for (int i = 1 - g; i <= Nr; ++i)
      for (int j = 1; j <= g; ++j) src[IX(L, i, 1 - j)] = (double)symz * src[IX(L, i, j)];
  for (int j = 1 - g; j <= Nz; ++j)
    for (int i = 1; i <= g; ++i) src[IX(L, 1 - i, j)] = (double)symr * src[IX(L, i, j)];
}

/* radbound: radbound.f90:2-77 */
void or_radbound(const oa_params* p, const or_level* L, const double* ev, double* src,
                 const double* Drv, const double* Dzv, double v, double var0) {
  const int Nr = L->Nr, Nz = L->Nz, g = L->g;
  for (int j = 1 - g; j <= Nz; ++j) {
    int k = IX(L, Nr, j);
    src[k] = -v * (L->r[k] * Drv[k] + L->z[k] * Dzv[k] + (ev[k] - var0)) / L->rr[k];
  }
  for (int i = 1 - g; i <= Nr; ++i) {
    int k = IX(L, i, Nz);
    src[k] = -v * (L->r[k] * Drv[k] + L->z[k] * Dzv[k] + (ev[k] - var0)) / L->rr[k];
  }
  if (!p->eqsym)
    for (int i = 1 - g; i <= Nr; ++i) {
      int k = IX(L, i, 1 - g);
      src[k] = -v * (L->r[k] * Drv[k] + L->z[k] * Dzv[k] + (ev[k] - var0)) / L->rr[k];
    }
}

/* symmetries_r / symmetries_z: generated, prl/arrays.pl:925-1117 */
void or_symmetries_r(const oa_params* p, or_level* L) {
  (void)p;
  for (int f = 0; f < OF_NFIELDS; ++f) {
    if (!L->present[f]) continue;
    double* v = L->var[f];
    const double s = (double)or_fields[f].symr;
    for (int j = 1 - L->g; j <= L->Nz; ++j)
      for (int i = 1; i <= L->g; ++i) v[IX(L, 1 - i, j)] = s * v[IX(L, i, j)];
  }
}

void or_symmetries_z(const oa_params* p, or_level* L) {
  (void)p;
  for (int f = 0; f < OF_NFIELDS; ++f) {
    if (!L->present[f]) continue;
    double* v = L->var[f];
    const double s = (double)or_fields[f].symz;
    for (int i = 1 - L->g; i <= L->Nr; ++i)
      for (int j = 1; j <= L->g; ++j) v[IX(L, i, 1 - j)] = s * v[IX(L, i, j)];
  }
}

/* saveold: prl/arrays.pl:759-875.  Strip layouts:
 *   b_rL/b_rR (m, j, k): m + g*(j+g-1) + g*nj*k
 *   b_zL/b_zR (i, n, k): (i+g-1) + ni*n + ni*g*k                         */
void or_saveold(const oa_params* p, or_level* L) {
  (void)p;
  const int g = L->g, ni = L->ni, nj = L->nj;
  for (int f = 0; f < OF_NFIELDS; ++f) {
    if (!L->present[f]) continue;
    const double* v = L->var[f];
    memcpy(L->var_p[f], v, sizeof(double) * L->n);
    for (int m = 0; m < g; ++m) {
      for (int j = 1 - g; j <= L->Nz; ++j) {
        int b0 = m + g * (j + g - 1);
        double* rl = L->b_rL[f];
        double* rr = L->b_rR[f];
        rl[b0 + 2 * g * nj] = rl[b0 + g * nj];
        rl[b0 + g * nj] = rl[b0];
        rl[b0] = v[IX(L, 1 - g + m, j)];
        rr[b0 + 2 * g * nj] = rr[b0 + g * nj];
        rr[b0 + g * nj] = rr[b0];
        rr[b0] = v[IX(L, L->Nr - m, j)];
      }
      for (int i = 1 - g; i <= L->Nr; ++i) {
        int b0 = (i + g - 1) + ni * m;
        double* zl = L->b_zL[f];
        double* zr = L->b_zR[f];
        zl[b0 + 2 * ni * g] = zl[b0 + ni * g];
        zl[b0 + ni * g] = zl[b0];
        zl[b0] = v[IX(L, i, 1 - g + m)];
        zr[b0 + 2 * ni * g] = zr[b0 + ni * g];
        zr[b0 + ni * g] = zr[b0];
        zr[b0] = v[IX(L, i, L->Nz - m)];
      }
    }
  }
}

/* accumulate: prl/arrays.pl:533-599 */
void or_accumulate(const oa_params* p, or_level* L, int k, int niter, double w) {
  (void)p;
  const int first = (k == 1), last = (k == niter);
  for (int f = 0; f < OF_NFIELDS; ++f) {
    if (!L->present[f]) continue;
    double *a = L->var_a[f], *s = L->src[f];
    if (!last) {
      if (first) ALL(L, q) a[q] = w * s[q];
      else ALL(L, q) a[q] = a[q] + w * s[q];
    } else {
      ALL(L, q) s[q] = a[q] + w * s[q];
    }
  }
}

/* update: prl/arrays.pl:1282-1328 */
void or_update(const oa_params* p, or_level* L, double dtw) {
  (void)p;
  for (int f = 0; f < OF_NFIELDS; ++f) {
    if (!L->present[f]) continue;
    double *v = L->var[f], *vp = L->var_p[f], *s = L->src[f];
    ALL(L, q) v[q] = vp[q] + dtw * s[q];
  }
}
