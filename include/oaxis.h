/*
 * oaxis.h -- C ABI of the MI355X (gfx950) evolution hot path for OllinAxis-BiB.
 *
 * The library replaces the body of the reference's Berger-Oliger step
 * (src/base/onestep.f90:78-469) -- sources, boundary conditions, RK4/ICN
 * accumulate/update, fine-boundary fill, axis/equator symmetries, the
 * state-mutating part of "auxiliary", restriction -- with HIP kernels that keep
 * every refinement level resident in device memory.  A Fortran main loop calls
 * it through ISO_C_BINDING (see ollinaxis-bib_amd/fortran/oaxis_mod.f90).
 *
 * Conventions
 *   - Every entry point returns 0 on success and a nonzero OA_ERR_* code on
 *     failure; oa_last_error() returns a message for the last failure on the
 *     calling thread.  Nothing aborts the process.
 *   - Host arrays exchanged with the library use the reference's layout: Fortran
 *     column-major, rho index fastest, extent (1-ghost:Nr, 1-ghost:Nz) of the
 *     level, leading dimension ld >= Nr+ghost (src/base/arrays.config, generated
 *     allocation in prl/arrays.pl:636-646).  Device storage is z-fastest; the
 *     transposition happens inside upload/download.
 *   - Field names are the reference's evolved-array names (arrays.config):
 *     alpha, phi, chi, A, B, H, C, trK, KTA, KTB, KTH, KTC, Delta_r, Delta_z,
 *     lambda, Alambda, beta_r, beta_z, dtbeta_r, dtbeta_z, scalar_phi,
 *     scalar_xi_r, scalar_xi_z, scalar_pi; plus the derived psi, psi4.
 *   - A handle is not re-entrant; distinct handles are independent.
 */
#ifndef OAXIS_H
#define OAXIS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
#define OA_OK              0
#define OA_ERR_ARG         1   /* bad argument / unknown name            */
#define OA_ERR_PARAM       2   /* par-file error (syntax, range, type)   */
#define OA_ERR_UNSUPPORTED 3   /* valid reference setting outside scope  */
#define OA_ERR_HIP         4   /* HIP runtime failure                    */
#define OA_ERR_STATE       5   /* call out of order                      */
#define OA_ERR_NOMEM       6   /* device or host allocation failed       */
#define OA_ERR_IO          7   /* file could not be read                 */

/* ------------------------------------------------------- parameter values */
/* String-valued reference keys map to these codes (param.f90 ranges). */
enum { OA_ORDER_TWO = 2, OA_ORDER_FOUR = 4 };
enum { OA_ICN = 0, OA_RK4 = 1 };
enum { OA_SLICING_STATIC = 0, OA_SLICING_HARMONIC, OA_SLICING_1PLUSLOG,
       OA_SLICING_SHOCKAVOID, OA_SLICING_HARMONIC2, OA_SLICING_1PLUSLOG2,
       OA_SLICING_SHOCKAVOID2, OA_SLICING_MAXIMAL };
enum { OA_ILAPSE_ONE = 0, OA_ILAPSE_SHELL, OA_ILAPSE_TORUS, OA_ILAPSE_ISOTROPIC,
       OA_ILAPSE_QUISO_KERR, OA_ILAPSE_PSIMINUS2, OA_ILAPSE_PSIMINUS4,
       OA_ILAPSE_MAXIMAL };
enum { OA_SHIFT_NONE = 0, OA_SHIFT_ZERO, OA_SHIFT_STATIC, OA_SHIFT_GD0,
       OA_SHIFT_GD1, OA_SHIFT_GD2, OA_SHIFT_GDSHOCK1, OA_SHIFT_GDSHOCK2 };
enum { OA_IDATA_CHECKPOINT = 0, OA_IDATA_MINKOWSKI, OA_IDATA_SCHWARZSCHILD,
       OA_IDATA_KERR, OA_IDATA_BRILLLINDQUIST, OA_IDATA_BRILLWAVE,
       OA_IDATA_SCALARPULSE, OA_IDATA_COMPLEXPULSE, OA_IDATA_TESTGW,
       OA_IDATA_BOSONSTAR, OA_IDATA_ROTBOSON };
enum { OA_BOUND_NONE = 0, OA_BOUND_STATIC, OA_BOUND_FLAT, OA_BOUND_RADIATIVE };
enum { OA_INTERP_BILINEAR = 1, OA_INTERP_BICUBIC = 3 };
enum { OA_AH_SHOOT = 0, OA_AH_SPECTRAL = 1, OA_AH_BOTH = 2 };
enum { OA_EULERIAN = 0, OA_LAGRANGIAN = 1 };
enum { OA_BSSN = 0, OA_Z4C = 1 };
enum { OA_DYNAMIC = 0, OA_BACKGROUND = 1 };
enum { OA_PERT_NONE = 0, OA_PERT_GAUSS = 1 };
enum { OA_ISHIFT_ZERO = 0, OA_ISHIFT_SHELL, OA_ISHIFT_TORUS, OA_ISHIFT_QUISO_KERR };
enum { OA_POT_NONE = 0, OA_POT_PHI2, OA_POT_PHI4 };
enum { OA_PULSE_SHELL = 0, OA_PULSE_TORUS, OA_PULSE_BAUMGARTE };
enum { OA_SCALARMETHOD_FIRST = 1, OA_SCALARMETHOD_SECOND = 2 };
enum { OA_ELL_NONE = 0, OA_ELL_WAVE, OA_ELL_SOR };
/* mattertype is a "multiple" key: bit mask */
enum { OA_MATTER_SCALAR = 1, OA_MATTER_COMPLEX = 2 };

/* Parameters of the evolution hot path.  Member names are the reference's
 * par-file keys and defaults are the reference's (src/base/param.f90:95-800).
 * Keys of the reference that this library does not act on (output, horizon
 * finder tuning, complex field, ...) are still accepted by the parser so that
 * reference par files load unchanged. */
typedef struct oa_params {
  /* grid (param.f90:95-168) */
  int    eqsym;
  double dr0, dz0;
  int    Nrtotal, Nztotal;
  int    Nb;
  int    Nl0, Nl1, Nl2, Nl3;
  int    ghost;                   /* 0: chosen from "order" (main.f90:286-317) */
  int    interporder;             /* OA_INTERP_*                            */
  double ft1, ft2, ft3, ft4, ft5, ft6, ft7, ft8;
  /* time stepping (param.f90:209-221) */
  int    adjuststep;
  double dtfac;
  int    Nt;
  /* slicing (param.f90:289-303) */
  int    slicing, ilapse;
  int    maximalevery;
  double gauge_f, gauge_eta;
  int    lapsepert;
  double lapse_a0, lapse_r0, lapse_z0, lapse_sr0, lapse_sz0;
  /* shift (param.f90:330-349) */
  int    shift, ishift;
  int    driverD0, driveradv;
  double shiftafter, drivercsi, drivereta;
  int    shiftpert;
  double shift_a0, shift_r0, shift_z0, shift_sr0, shift_sz0;
  /* initial data (param.f90:359-361) */
  int    idata;
  int    factorhdet;
  /* evolution (param.f90:400-438) */
  int    spacetime, formulation, bssnflavor, integrator, order;
  int    angmom, chimethod, evolveH, evolveKTH, shiftadvect;
  int    nolambda, noDelta_r, noDelta_z, noDelta_p;
  int    icniter, chipower;
  double eta, kappa1, kappa2;
  double geodiss, scalardiss, complexdiss;
  int    boundtype;
  /* black holes (param.f90:491-502) */
  int    N_BH;
  double BH1mass, BH1spin, BH1Z, BH2mass, BH2spin, BH2Z;
  int    duplicate_bhs;
  /* horizons (param.f90:540-562) */
  int    ahfind, ahfind_every;
  double ahafter;
  int    ahmove, N_ah, ahmethod;  /* ahmethod: OA_AH_*                     */
  double z_ah1, z_ah2, ahzinit, ahrmax, ahepsilon_shoot;
  /* spectral method, fast flow (param.f90:549-567) */
  int    ahNmodes, ahmaxiter, ahshootfirst;
  double ahepsilon_spectral, ahAA, ahBB, aheccentric;
  /* Brill waves (param.f90:572-576) */
  double brill_a0, brill_sr0, brill_sz0;
  /* matter (param.f90:615-648) */
  int    mattertype;              /* OA_MATTER_* bit mask, 0 = vacuum       */
  int    scalarpotential, scalarpulse_type, scalarmethod;
  double scalar_mass, scalar_lambda;
  double scalar_a0, scalar_r0, scalar_z0, scalar_sr0, scalar_sz0;
  /* elliptic solve of the Hamiltonian constraint for Brill-wave and
   * scalar-pulse data (param.f90:454-491) */
  int    ELL_solver;              /* OA_ELL_*                               */
  int    ELL_maxiter;
  double ELL_epsilon;
  int    WE_method, WE_maxiter;   /* WE_method: OA_ICN or OA_RK4            */
  double WE_eta, WE_diss, WE_dtfac, WE_epsilon;
} oa_params;

/* Fill *p with the reference defaults. */
int    oa_params_default(oa_params* p);
/* Assign one "key = value" pair with the reference's typing and range rules
 * (generated assign.f90, prl/assign.pl:28-259).  Unknown key -> OA_ERR_PARAM. */
int    oa_params_set(oa_params* p, const char* key, const char* value);
/* Parse a whole par file ("key = value", '#' comments; parse.f90:7-373). */
int    oa_params_parse_file(oa_params* p, const char* path);
/* Parse par-file text held in memory (same rules as the file form). */
int    oa_params_parse_text(oa_params* p, const char* text);
/* Read back one key as text (for tests / drivers without struct access). */
int    oa_params_get(const oa_params* p, const char* key, char* out, int outlen);
/* Check that the settings lie inside what this library implements. */
int    oa_params_check_supported(const oa_params* p, char* msg, int msglen);
size_t oa_params_size(void);

const char* oa_last_error(void);

/* --------------------------------------------------------------- handles */
typedef struct oa_handle oa_handle;

/* Allocate every level of box 0 on the current HIP device.  Grid geometry,
 * time steps and the forced odd z size follow setupboxes.f90:104-141 and
 * evolve.f90:42-72. */
int oa_create(const oa_params* p, oa_handle** h);
int oa_destroy(oa_handle* h);

/* Level geometry: number of levels, and per level the local sizes Nr, Nz
 * (interior + boundary, ghosts excluded), ghost, dr, dz, dt. */
int oa_num_levels(const oa_handle* h, int* nlevels);
int oa_level_info(const oa_handle* h, int box, int level, int* Nr, int* Nz,
                  int* ghost, double* dr, double* dz, double* dt);

/* Host <-> device transfer of one named array, reference layout (see top). */
int oa_upload_field(oa_handle* h, int box, int level, const char* name,
                    const double* src, int ld);
int oa_download_field(oa_handle* h, int box, int level, const char* name,
                      double* dst, int ld);

/* Level clocks t, t1, t2 and step counter s (onestep.f90:351-357). */
int oa_get_time(const oa_handle* h, int box, int level, double* t, double* t1,
                double* t2, int* s);
int oa_set_time(oa_handle* h, int box, int level, double t, double t1,
                double t2, int s);

/* ---- the Berger-Oliger step body, callable piecewise in the reference order
 * (onestep.f90:97, 188-335, 351-357, 421-463). */
/* saveold: keep the previous-time state and shift the fine-boundary time
 * levels (generated saveold, prl/arrays.pl:759-875). */
int oa_saveold(oa_handle* h, int box, int level);
/* One internal iteration: sources_geometry/lapse/shift/matter, radiative
 * boundaries (level 0), accumulate (RK4), update, axis/equator symmetries and
 * the pointwise part of auxiliary (factordet, KT_H, lambda, chi/phi).
 * iter in 1..niter, dtw and weight as in onestep.f90:141-177. */
int oa_substep(oa_handle* h, int box, int level, int iter, int niter,
               double dtw, double weight);
/* finebound(box, level, dtw, .true.) for level > 0 (finebound.f90:2-407). */
int oa_finebound(oa_handle* h, int box, int level, double dtw);
/* auxiliary: completes the state projection that needs neighbours
 * (Delta = DelDef while the level clock is 0) and re-projects a level whose
 * data was changed by restriction or upload (auxiliary_geometry.f90). */
int oa_auxiliary(oa_handle* h, int box, int level);
/* restrict(box, level, .true.) into level-1 (restrict.f90:2-169) followed by
 * that level's symmetries (onestep.f90:431-451). */
int oa_restrict(oa_handle* h, int box, int level);
/* t2=t1; t1=t; s=s+1; t=t+dt (onestep.f90:351-357). */
int oa_advance_time(oa_handle* h, int box, int level);

/* ---- whole-step entry points */
/* evolve.f90:128-195: symmetries and auxiliary on every level, then the
 * Gamma-driver-2 initialisation dtbeta = drivercsi * Delta. */
int oa_initial_aux(oa_handle* h);
/* onestep(0) including the recursion into finer levels and restriction. */
int oa_coarse_step(oa_handle* h);
/* nsteps coarse steps. */
int oa_evolve(oa_handle* h, int nsteps);
/* Wait for all device work of the handle. */
int oa_sync(oa_handle* h);

/* Initial data on every level, computed on the device: minkowski (+ gauss
 * lapse/shift perturbations), schwarzschild, BrillLindquist, and BrillWave
 * (Holz) and scalarpulse with psi from the WaveElliptic solve of the
 * Hamiltonian constraint (ELL_solver = wave).  Follows
 * initial.f90:53-150,189-330,528-620, idata_brillwave.f90:88-208,
 * idata_scalarpulse.f90:60-206 and wave_elliptic.f90:2-1164. */
int oa_initial_data(oa_handle* h);
/* Stopping rule of that solve.  The reference stops when the RMS of dv/dt on
 * the base level falls below WE_epsilon (after at least 100 steps) or at
 * WE_maxiter.  With refinement levels that residual stalls at the mismatch
 * between the coarse and the restricted fine solution, so the reference runs
 * to WE_maxiter although the solution stops changing after a few thousand
 * steps.  stationary = 1 (the default; no counterpart in the reference)
 * also ends the all-levels phase once the largest change of u per step on
 * the base level, relative to max|u|, has stayed below 1e-12 over 100 steps
 * and stopped falling; the solution then differs from the one WE_maxiter
 * steps give by rounding only.  stationary = 0 keeps the reference's rule
 * alone.  oa_elliptic_stats reports the iterations of the base-level solve
 * and of the all-levels solve (0 without refinement) and the last residual. */
int oa_set_elliptic_exit(oa_handle* h, int stationary);
int oa_elliptic_stats(const oa_handle* h, int* steps0, int* steps1, double* residual);

/* Device time of the last oa_evolve / oa_coarse_step call in ms (HIP events
 * on the handle's stream), and the kernel launch counts. */
int oa_last_elapsed_ms(const oa_handle* h, double* ms);
/* Schedule of whole coarse steps (oa_coarse_step, oa_evolve).  No
 * counterpart in the reference.  A step runs either as a captured HIP graph
 * (reused while the step's sequence of launches repeats, with fresh work
 * items each step) or kernel by kernel, and the RHS blocks are either
 * renumbered per XCD or not.  Neither choice changes any arithmetic, so by
 * default an autotuner times each of the four schedules over steady-state
 * steps (one untimed and two timed steps each) and keeps the fastest.
 * oa_set_graphs fixes the first choice and turns the autotuner off.
 * oa_schedule reports the schedule in use, whether tuning has finished, and
 * the mean step time (ms) it measured for schedules 0..3 (bit 0: graph,
 * bit 1: XCD renumbering). */
int oa_set_graphs(oa_handle* h, int on);
int oa_schedule(const oa_handle* h, int* graphs, int* xcd_remap, int* tuned, double* step_ms);
/* Per-launch HIP-event timing of the dominant RHS kernel during oa_evolve
 * (off by default; adds two event records per launch on the same stream;
 * while on, steps are launched directly rather than as graphs). */
int oa_set_kernel_timing(oa_handle* h, int on);
/* Average duration (ms) of the dominant RHS kernel over the last timed
 * interval and how many launches it averaged. */
int oa_rhs_kernel_stats(const oa_handle* h, double* avg_ms, long* launches);
/* Grid-point updates (sum over levels and steps of Nrtotal*Nztotal) done
 * by this handle so far. */
int oa_point_updates(const oa_handle* h, double* updates);

/* constraints(box, level) (src/geometry/constraints.f90:36-206) on the
 * level's current state: the Hamiltonian constraint ham, the momentum
 * constraints mom_r, mom_z and the Delta constraints CDelta_r, CDelta_z,
 * computed on the device from the state's geometric auxiliaries.  Read them
 * with oa_download_field or oa_norms0d under those names, until the level's
 * state changes. */
int oa_constraints(oa_handle* h, int box, int level);
/* 0D norms of an evolved array on one level, as save0Ddata.f90:221-288
 * forms them over i in [1, Nr-ghost], j in [1, Nz-ghost]: out[0] = max,
 * out[1] = min, out[2] = nm1 = max |f|, out[3] = nm2 = sqrt(sum f^2 / N).
 * Reduced on the device (the sum in a fixed order that differs from the
 * reference's serial loop, so nm2 agrees to rounding).  A NaN in the range
 * returns OA_ERR_STATE (the reference stops the run there). */
int oa_norms0d(oa_handle* h, int box, int level, const char* name, double* out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc): the 0D norms of n arrays (1 <= n <=
 * OA_NORMS_NOALLOC_MAX) of one level, each formed as oa_norms0d forms it,
 * written to DEVICE memory out_dev: 5 doubles per array, in the order of
 * names: max, min, nm1, nm2, and 1.0 if the range holds a NaN (else 0.0).
 * Nothing is allocated: the per-block partial sums live in one fixed device
 * array of the library per GPU (OA_NORMS_NOALLOC_MAX x 256 records of 6
 * doubles, 384 KiB, held until the process ends), every record of which a
 * call rewrites before reading it.  The call returns once the work is
 * enqueued on the handle's stream (results are ready after oa_sync).  Calls
 * that overlap on one GPU (other handles' streams, other threads) share that
 * array and race. */
#define OA_NORMS_NOALLOC_MAX 32
int oa_norms0d_noalloc(oa_handle* h, int box, int level, int n, const char* const* names, double* out_dev);
/* Checkpoint and restart (checkpointsave / checkpointload,
 * src/base/checkpoint.f90:2-690).  The reference writes every evolved array,
 * the fine-boundary time levels var_bound_* with the clocks t, t1, t2 of each
 * level, and the par file, as text.  Here the same content, plus the
 * per-level state of the auxiliary pass, is one opaque byte block.  A handle
 * created with the same parameters and loaded from it continues bit for bit
 * as the handle that saved it.  Between coarse steps only (OA_ERR_STATE
 * inside a level's step); load refuses a block of other parameters. */
int oa_checkpoint_size(const oa_handle* h, size_t* bytes);
int oa_checkpoint_save(oa_handle* h, void* buf, size_t bytes);
int oa_checkpoint_load(oa_handle* h, const void* buf, size_t bytes);
/* checkpointrestart (checkpoint.f90:338-690), reached from initial.f90:157-161
 * with idata = checkpoint: read the state from the directory dir (the par
 * file's checkpointfile) that a single-process reference run wrote with
 * checkpointsave.  checkparam must agree with this handle's eqsym, angmom,
 * dr0, dz0, Nrtotal, Nztotal, Nb, Nl0..Nl3 and mattertype (the reference's
 * checks and messages, :400-530); checklist names the arrays.  For every
 * level the arrays this spacetime evolves are read from their .2D dumps
 * (9 significant digits, as the reference wrote them) and their rR, zL, zR
 * fine-boundary time levels from the _bound_ files, whose headers set the
 * level clocks t, t1, t2; the step counters restart at 0.  Arrays the
 * reference also lists but this path does not form (dtalpha, KTH, lambda,
 * hdet, ...) are skipped.  Then call oa_initial_aux, as evolve.f90:128-195
 * runs symmetries, the auxiliary pass and the Gamma-driver initialisation
 * after initial.  The horizon finder starts afresh, as in a new run of the
 * reference.  oa_initial_data refuses idata = checkpoint (OA_ERR_STATE) and
 * points here. */
int oa_checkpoint_restart(oa_handle* h, const char* dir);
/* checkpointsave (checkpoint.f90:2-328) for one process into dir (created
 * if missing; the reference names it <directory>/checkpoint_t=<t>):
 * checkparam, checklist (parfile's name, then r, z, rr and the arrays this
 * spacetime evolves), each listed array's .2D dump as save2Dvariable writes
 * it there (9 significant digits, + 1e-30) and, but for the coordinates,
 * its _bound_rL/rR/zL/zR files with the level clocks (16 digits; rL holds
 * zeros, box 0 never fills it).  oa_checkpoint_restart reads it back.  The
 * reference's restart reads the arrays its checklist names, so from this
 * directory it would start the arrays this path does not form from its own
 * initial values; without a reference build that is not checked here.
 * Between coarse steps only. */
int oa_checkpoint_write(oa_handle* h, const char* dir, const char* parfile);

/* One apparent-horizon search: what ahshoot (src/geometry/horizon_finder.f90:
 * 246-1103) finds for horizon k and writes to its ah_area / ah_mass files. */
typedef struct oa_horizon {
  int    found;      /* ahfound(k) after the search.  As in the reference it keeps
                        its previous value when the bisection converges on a
                        non-horizon within dzlocal of the centre (:797-845).     */
  int    ntheta;     /* Ntheta: accepted steps of the last integration          */
  int    shots;      /* integrations of the horizon ODE this search ran         */
  int    errorflag;  /* errorflag of the last integration                       */
  double zcenter;    /* centre on the axis, zah(k) (:100-110)                   */
  double zinit;      /* axis position where the shooting stopped (:797-815)     */
  double dradlast;   /* dh/dtheta at the end of the last integration            */
  double area;       /* ah_area: 0 unless found; doubled with eqsym (:989-995)  */
  double radius;     /* RAH = sqrt(area/pi)/2 (:997)                            */
  double mass;       /* MAH = RAH/2: no angular momentum in scope (:1065-1076)  */
  int    method;     /* OA_AH_* of the search.  With spectral and both the
                        fields above describe the fast flow (ahspectral,
                        :1113-1972): ntheta = Ntheta = 10 ahNmodes, errorflag
                        its error flag (also set when ac(0) < dr0 + dz0), and
                        shots, zinit, dradlast those of the shooting search
                        that ran before it in this call (both; ahshootfirst
                        on the first call), else 0.                          */
  int    iterations; /* fast-flow iterations (0 for shoot)                      */
  double residual;   /* max_j |c_j| of the last iteration (0 for shoot)         */
} oa_horizon;

/* horizon_finder (horizon_finder.f90:2-237) for horizon k (0 <= k < N_ah,
 * centre 0, z_ah1, z_ah2) on the current state of box 0 (every level
 * projected): replaces the reference's call at evolve.f90:251 and
 * :429-432.  The metric quantities the horizon ODE interpolates are formed
 * on the device for every level.  ahmethod = shoot: the shots of the search
 * (the march along the axis and the bisection) are integrated on the
 * device, many per launch; the sequence of shots and every decision between
 * them are the reference's.  ahmethod = spectral: the fast flow of the
 * cosine coefficients (ahspectral, :1113-1972) runs whole on the device, one
 * workgroup per search, from the last horizon found (ack(k,:), kept between
 * searches) or a large circle (an ellipse with aheccentric), after a
 * shooting search on the first call with ahshootfirst (:1292-1303);
 * ahmethod = both: a shooting search, then the flow from its surface
 * (:223-226).  With ahmove the centres of horizons 1 and 2 first move to
 * the nearest maximum of phi on the axis (:117-180). */
int oa_ahfind(oa_handle* h, int k, oa_horizon* out);
/* (theta_i, h(theta_i)), i = 0..ntheta, of horizon k's last search, h = 0
 * when not found.  Shooting: the last integration (what ahshoot writes to
 * ah_radius_SH, :956-966), run again on the quantities the search
 * interpolated, so it needs no state; a later search of the same spacetime
 * replaces them (OA_ERR_STATE: search again).  Spectral and both: the
 * surface of the flow's last iteration on theta_i = i pi / (s Ntheta),
 * s = 2 with equatorial symmetry about a centre at 0, else 1 (what
 * ahspectral writes to ah_radius_SP, :1766-1790).  cap is the room in
 * theta and rad; *n receives ntheta + 1 (also when the arrays are null, to
 * size them). */
int oa_horizon_shape(oa_handle* h, int k, double* theta, double* rad, int cap, int* n);
/* oa_ahfind in two calls, so that a search can run while the spacetime
 * evolves on: oa_ahfind_begin forms the quantities the search interpolates
 * from the current state (in order with the spacetime's work, so a later
 * step cannot change what it reads) and returns a search;
 * oa_ahfind_end runs its shots on a stream of their own, concurrently with
 * whatever the caller enqueued in between (a coarse step, typically), and
 * returns the result oa_ahfind would have returned at the begin.  No
 * counterpart in the reference, which searches and then evolves; for the
 * serving loop, whose searches would otherwise leave the GPU idle between
 * steps.  A spacetime with a begun search may not begin another, be
 * destroyed or leave its batch before the search ends (OA_ERR_STATE). */
typedef struct oa_ahsearch oa_ahsearch;
int oa_ahfind_begin(oa_handle* h, int k, oa_ahsearch** search);
int oa_ahfind_end(oa_ahsearch* search, oa_horizon* out);

/* Device memory (bytes) oa_create allocates for these parameters: what a
 * scheduler admitting spacetimes by HBM footprint needs. */
int oa_params_footprint(const oa_params* p, double* bytes);

/* ---- batches of spacetimes
 * No counterpart in the reference, which evolves one spacetime per run.
 * For callers that evolve many independent spacetimes at once (parameter
 * scans, the serving loop): every member advances by one coarse step per
 * oa_batch_coarse_step, and each kernel of the step runs as ONE launch over
 * the work items (spacetime, level) of all members that take part in it.
 * Members may differ in grid size, level count, integrator and physics.
 * Each member's arithmetic is that of oa_coarse_step on the member alone:
 * results are bit-identical.  While a member, a handle's own calls
 * (download, upload, ...) run on the batch's stream, in order with the
 * batch's steps; oa_coarse_step / oa_evolve on a member are refused. */
typedef struct oa_batch oa_batch;
int oa_batch_create(oa_batch** out);
int oa_batch_destroy(oa_batch* b);          /* members leave first; none is destroyed */
/* h must have its state projected (oa_initial_aux done); it joins at its
 * current level clocks. */
int oa_batch_add(oa_batch* b, oa_handle* h);
int oa_batch_remove(oa_batch* b, oa_handle* h);
int oa_batch_size(const oa_batch* b, int* n);
/* One coarse step of every member (onestep(0) of each, level by level in
 * lockstep).  Returns after the work is enqueued; oa_batch_sync waits. */
int oa_batch_coarse_step(oa_batch* b);
int oa_batch_evolve(oa_batch* b, int nsteps);
int oa_batch_sync(oa_batch* b);
/* Device time (ms) of the last oa_batch_evolve, from HIP events. */
int oa_batch_last_elapsed_ms(const oa_batch* b, double* ms);
/* Grid-point updates by all members during this batch's steps. */
int oa_batch_point_updates(const oa_batch* b, double* updates);
/* Launches enqueued by the last oa_batch_coarse_step (one per kernel kind
 * and physics group) and the work items they covered. */
int oa_batch_launch_stats(const oa_batch* b, long* launches, long* items);
/* horizon_finder for horizon k[i] of members[i], i < n (every one a member
 * of b), as oa_ahfind does for each alone, with the same results: the
 * searches advance in lockstep and each round of shots of all of them is
 * one launch.  For the serving loop, which checks many spacetimes for
 * horizon formation at once. */
int oa_batch_ahfind(oa_batch* b, int n, oa_handle* const* members, const int* k, oa_horizon* out);
/* oa_batch_ahfind in two calls, as oa_ahfind_begin / oa_ahfind_end (n >= 1;
 * out of oa_ahfind_end then holds n results, in the order of members). */
int oa_batch_ahfind_begin(oa_batch* b, int n, oa_handle* const* members, const int* k, oa_ahsearch** search);
/* HIP-event timing of every RHS launch during oa_batch_evolve (off by
 * default): total kernel time (ms), launches, and grid points (Nr x Nz of
 * each item) those launches advanced by one substep. */
int oa_batch_set_kernel_timing(oa_batch* b, int on);
/* As oa_set_graphs / oa_schedule, for the batch's coarse steps. */
int oa_batch_set_graphs(oa_batch* b, int on);
int oa_batch_schedule(const oa_batch* b, int* graphs, int* xcd_remap, int* tuned, double* step_ms);
int oa_batch_rhs_kernel_stats(const oa_batch* b, double* total_ms, long* launches, double* points);

#ifdef __cplusplus
}
#endif
#endif /* OAXIS_H */
