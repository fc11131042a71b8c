/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the OllinAxis-BiB evolution hot path, written in
 * the reference's whole-array style (one loop over the level per reference
 * statement) so that every step can be compared with the Fortran source it
 * follows (file:line in each function).  It is the checker for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * legs (the throughput sample and the horizon probe's search) may load it.  The product never links, loads or calls it.
 *
 * Arrays use the reference layout: Fortran column-major, rho index fastest,
 * index range (1-ghost:Nr, 1-ghost:Nz).
 *
 * Parity: the functions below are pinned by the tests in tests/test_oracle_*.py
 * (polynomial exactness of every stencil and closure, the Kreiss-Oliger
 * symbol, RK4/ICN amplification factors, exact Minkowski stationarity,
 * 4th-order convergence of the static-Schwarzschild residual, bicubic
 * exactness of prolongation/restriction, quadratic exactness of the time
 * interpolation, an independent symbolic (sympy) derivation of the
 * covariant BSSN / Klein-Gordon right-hand sides evaluated pointwise, the
 * radiative boundary as an outgoing spherical wave, the restriction and
 * fine-boundary point sets derived from coordinates, DelDef as the sympy
 * connection difference at clock 0, the initial-data closed forms, and the
 * elliptic solve's discrete equation and Hamiltonian-constraint convergence).
 * Parity unpinned: the call-order consequences of the first steps (phi-mode
 * sources in the first substep of a level's second step, the
 * pre-initialisation dtbeta), which follow the reference's text.
 * The reference Fortran build itself is not available as a further check
 * (it needs Perl-generated sources and stand-ins for missing symbols).
 */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include "../include/oaxis.h"

/* ------------------------------------------------------------- fields */
/* Evolved arrays of the reference (src/base/arrays.config), in the subset
 * this hot path supports.  Every one gets var, var_p, var_a, s<var> and the
 * four fine-boundary time-level strips, as the generator does
 * (prl/arrays.pl:636-646). */
enum {
  OF_alpha = 0, OF_dtalpha, OF_phi, OF_dtphi, OF_chi,
  OF_A, OF_B, OF_H, OF_C,
  OF_trK, OF_KTA, OF_KTB, OF_KTH, OF_KTC,
  OF_Delta_r, OF_Delta_z, OF_lambda, OF_Alambda, OF_z4theta,
  OF_beta_r, OF_beta_z, OF_dtbeta_r, OF_dtbeta_z, OF_fdriver,
  OF_scalar_phi, OF_scalar_xi_r, OF_scalar_xi_z, OF_scalar_pi,
  OF_NFIELDS
};

typedef struct {
  const char* name;
  int symr, symz;   /* arrays.config SYMMETRYR / SYMMETRYZ */
} or_fielddesc;

extern const or_fielddesc or_fields[OF_NFIELDS];

/* ------------------------------------------------------------- level */
typedef struct {
  int Nr, Nz, g;          /* local sizes; arrays are (1-g:Nr, 1-g:Nz)      */
  int ni, nj, n;          /* Nr+g, Nz+g, ni*nj                              */
  double dr, dz, dt;
  double rmin, rmax, zmin, zmax;   /* rminl..zmaxl (evolve.f90:86-96)      */
  double *r, *z, *rr;

  int present[OF_NFIELDS];
  double *var[OF_NFIELDS], *var_p[OF_NFIELDS], *var_a[OF_NFIELDS], *src[OF_NFIELDS];
  double *b_rL[OF_NFIELDS], *b_rR[OF_NFIELDS];   /* (0:g-1, 1-g:Nz, 0:2)   */
  double *b_zL[OF_NFIELDS], *b_zR[OF_NFIELDS];   /* (1-g:Nr, 0:g-1, 0:2)   */

  /* clocks (onestep.f90:351-357) */
  double t, t1, t2;
  int s;

  /* set when derivative arrays are supplied from outside (pointwise tests) */
  int ext_derivs;

  /* auxiliary arrays (AUXILIARY intent in arrays.config) */
  double *auxarray;
  double *falpha, *psi, *psi2, *psi4, *hdet, *ihdet, *Dr_hdet, *Dz_hdet;
  double *Dr_alpha, *Dz_alpha, *Drr_alpha, *Dzz_alpha, *Drz_alpha, *DAr_alpha, *DAz_alpha;
  double *Dr_phi, *Dz_phi, *Drr_phi, *Dzz_phi, *Drz_phi, *DAr_phi, *DAz_phi;
  double *Dr_chi, *Dz_chi, *Drr_chi, *Dzz_chi, *Drz_chi, *DAr_chi, *DAz_chi;
  double *Dr_A, *Dz_A, *Drr_A, *Dzz_A, *Drz_A, *DAr_A, *DAz_A;
  double *Dr_B, *Dz_B, *Drr_B, *Dzz_B, *Drz_B, *DAr_B, *DAz_B;
  double *Dr_H, *Dz_H, *Drr_H, *Dzz_H, *Drz_H, *DAr_H, *DAz_H;
  double *Dr_C, *Dz_C, *Drr_C, *Dzz_C, *Drz_C, *DAr_C, *DAz_C;
  double *g_A, *g_B, *g_H, *g_C;
  double *Dr_g_A, *Dz_g_A, *Dr_g_B, *Dz_g_B, *Dr_g_H, *Dz_g_H, *Dr_g_C, *Dz_g_C;
  double *Dr_beta_r, *Dz_beta_r, *Drr_beta_r, *Dzz_beta_r, *Drz_beta_r, *DAr_beta_r, *DAz_beta_r;
  double *Dr_beta_z, *Dz_beta_z, *Drr_beta_z, *Dzz_beta_z, *Drz_beta_z, *DAr_beta_z, *DAz_beta_z;
  double *Dr_dtbeta_r, *Dz_dtbeta_r, *DAr_dtbeta_r, *DAz_dtbeta_r;
  double *Dr_dtbeta_z, *Dz_dtbeta_z, *DAr_dtbeta_z, *DAz_dtbeta_z;
  double *DD_beta_rr, *DIV_beta, *Dr_DIV_beta, *Dz_DIV_beta;
  double *lamDef, *AlamDef, *g_lamDef, *g_lambda;
  double *Dr_trK, *Dz_trK, *DAr_trK, *DAz_trK;
  double *Dr_KTA, *Dz_KTA, *DAr_KTA, *DAz_KTA;
  double *Dr_KTB, *Dz_KTB, *DAr_KTB, *DAz_KTB;
  double *Dr_KTC, *Dz_KTC, *DAr_KTC, *DAz_KTC;
  double *KT_up_rr, *KT_up_zz, *KT_up_rz;
  double *KT2_A, *KT2_B, *KT2_H, *KT2_C, *KT2, *K2;
  double *DelDef_r, *DelDef_z;
  double *Dr_Delta_r, *Dz_Delta_r, *DAr_Delta_r, *DAz_Delta_r;
  double *Dr_Delta_z, *Dz_Delta_z, *DAr_Delta_z, *DAz_Delta_z;
  double *D2cov_alpha_A, *D2cov_alpha_B, *D2cov_alpha_H, *D2cov_alpha_C, *Lapla_alpha;
  double *RIC_A, *RIC_B, *RIC_H, *RIC_C, *RSCAL;
  double *DIV_KT_r, *DIV_KT_z, *DIVA_KT_r, *DIVA_KT_z;
  /* scalar field */
  double *Dr_scalar_phi, *Dz_scalar_phi, *Drr_scalar_phi, *Dzz_scalar_phi, *Drz_scalar_phi;
  double *Dr_scalar_pi, *Dz_scalar_pi, *DAr_scalar_pi, *DAz_scalar_pi;
  double *Dr_scalar_xi_r, *Dz_scalar_xi_r, *DAr_scalar_xi_r, *DAz_scalar_xi_r;
  double *Dr_scalar_xi_z, *Dz_scalar_xi_z, *DAr_scalar_xi_z, *DAz_scalar_xi_z;
  double *scalar_V, *scalar_VP;
  double *rho, *J_r, *J_z, *S_A, *S_B, *S_H, *S_C, *trS;
  /* constraints */
  double *ham, *mom_r, *mom_z, *CDelta_r, *CDelta_z;
} or_level;

struct or_ell_level;        /* elliptic arrays of one level (or_elliptic.c) */

typedef struct or_handle {
  oa_params p;
  int nlev;                 /* Nl0 + 1                                     */
  double dt0;
  or_level* L;              /* [nlev]                                      */
  double* scratch;          /* one level-sized temporary                   */
  struct or_ell_level* ell; /* [nlev] while the elliptic solve has run     */
  int ell_steps[2];         /* iterations: base level alone, all levels    */
  double ell_res;           /* residual of the last iteration              */
} or_handle;

/* index of (i,j) with i in [1-g, Nr], j in [1-g, Nz] */
#define IX(L, i, j) (((i) + (L)->g - 1) + ((j) + (L)->g - 1) * (L)->ni)
/* whole-array loop over every point */
#define ALL(L, k) for (int k = 0; k < (L)->n; ++k)

/* ---------------------------------------------------------- numerics */
void or_diff1r(const oa_params* p, const or_level* L, const double* f, int symr, double* out);
void or_diff1z(const oa_params* p, const or_level* L, const double* f, int symz, double* out);
void or_diff2r(const oa_params* p, const or_level* L, const double* f, int symr, double* out);
void or_diff2z(const oa_params* p, const or_level* L, const double* f, int symz, double* out);
void or_diff2rz(const oa_params* p, const or_level* L, const double* f, int symr, int symz, double* out);
void or_diffadvr(const oa_params* p, const or_level* L, const double* f, int symr,
                 const double* beta_r, double* out);
void or_diffadvz(const oa_params* p, const or_level* L, const double* f, int symz,
                 const double* beta_z, double* out);
void or_dissipation(const oa_params* p, const or_level* L, const double* ev, double* src,
                    int symr, int symz, double diss);
void or_radbound(const oa_params* p, const or_level* L, const double* ev, double* src,
                 const double* Drv, const double* Dzv, double v, double var0);

/* ----------------------------------------------------- level pieces */
void or_symmetries_r(const oa_params* p, or_level* L);
void or_symmetries_z(const oa_params* p, or_level* L);
void or_saveold(const oa_params* p, or_level* L);
void or_accumulate(const oa_params* p, or_level* L, int k, int niter, double w);
void or_update(const oa_params* p, or_level* L, double dtw);

void or_auxiliary(const oa_params* p, or_level* L);
void or_auxiliary_geometry(const oa_params* p, or_level* L);
void or_constraints(const oa_params* p, or_level* L);
void or_auxiliary_matter(const oa_params* p, or_level* L);
void or_potential(const oa_params* p, or_level* L);
void or_stressenergy(const oa_params* p, or_level* L);
void or_sources_geometry(const oa_params* p, or_level* L);
void or_sources_lapse(const oa_params* p, or_level* L);
void or_sources_shift(const oa_params* p, or_level* L);
void or_sources_scalar(const oa_params* p, or_level* L);
void or_radiative_geometry(const oa_params* p, or_level* L);

double or_interp(const oa_params* p, const or_level* L, const double* f, double r0, double z0);
int or_solve_hamiltonian(or_handle* H);
void or_ell_free(or_handle* H);
void or_finebound(or_handle* H, int level, double dtw);
void or_restrict(or_handle* H, int level);

#endif
