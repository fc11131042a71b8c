// Internal declarations shared by the host runtime and the HIP kernels.
#pragma once

#include <string>

#include "oaxis.h"

// thread-local last-error text behind oa_last_error()
void oa_set_error(const std::string& msg);

namespace oax {

// Device state slots.  Every level stores its evolved fields as z-fastest
// planes, one plane per field, fields of one slab contiguous.  Vacuum uses
// the first NBASE slots; the Gamma-driver shift and the real scalar field
// append their slots in that order.
enum Slot {
  S_alpha = 0, S_chi, S_phi, S_A, S_B, S_H, S_C, S_trK, S_KTA, S_KTB, S_KTC,
  S_Delta_r, S_Delta_z,
  NBASE
};
enum ShiftSlot { SH_beta_r = 0, SH_beta_z, SH_dtbeta_r, SH_dtbeta_z, NSHIFT };
static_assert(SH_dtbeta_z == SH_dtbeta_r + 1, "dtbeta_r and dtbeta_z planes are adjacent");
enum ScalarSlot { SC_phi = 0, SC_xi_r, SC_xi_z, SC_pi, NSCALAR };

// Planes of the per-level buffer of geometric auxiliaries that the first
// phase of the RHS kernel writes and the second reads: the Ricci tensor (A,
// B, C components and the scalar) and the divergences DIV_KT, DIVA_KT.
enum AuxSlot {
  AUX_RA = 0, AUX_RB, AUX_RC, AUX_RSCAL, AUX_DIV_r, AUX_DIV_z, AUX_DIVA_r, AUX_DIVA_z, NAUX
};

constexpr int nslots(bool shift, bool scalar) {
  return NBASE + (shift ? NSHIFT : 0) + (scalar ? NSCALAR : 0);
}
constexpr int shift_slot(int s) { return NBASE + s; }
constexpr int scalar_slot(bool shift, int s) { return NBASE + (shift ? NSHIFT : 0) + s; }

// Geometry of one level as the kernels see it.  Point (i, j) of the
// reference's (1-g:Nr, 1-g:Nz) range lives at (i+g-1)*ld + (j+g-1).
struct KGrid {
  int Nr, Nz, g, ld;
  long plane;           // elements of one field plane (nr * ld)
  int eqsym;
  int jlo;              // first computed z index: 1 with equatorial symmetry, 1-g without
  double dr, dz, dt;
  double idr, idz;      // 1/dr, 1/dz
  double zoff;          // z_j = (j - zoff) * dz
};

// Physics switches and constants of one right-hand-side launch.
struct KPhys {
  int slicing;
  double gauge_f, gauge_eta;
  double geodiss, scalardiss, eta;
  double drivercsi, drivereta;
  int driveradv, shiftadvect;
  int scalarpotential, scalarmethod;
  double scalar_mass, scalar_lambda;
  int chimethod;
  int radiative;        // level 0 with boundtype = radiative
};

// RK bookkeeping of one internal iteration.
struct KStep {
  int iter, niter, rk4;
  double dtw, w;
  int proj_t0;          // the auxiliary call ending this iteration runs at clock 0
  // Non-null on a level's first substep after the initial auxiliary pass:
  // the dtbeta_r, dtbeta_z planes as they were before the Gamma-driver
  // initialisation dtbeta = drivercsi * Delta.  The reference differentiates
  // dtbeta in that auxiliary pass, before the initialisation, so the first
  // substep's dtbeta derivatives come from these planes.
  const double* dtb_pre;
};

// Time-interpolation coefficients of one finebound call (computed on the
// host exactly as the generated bound_interp.inc forms them).
struct KTime {
  int border;           // 1: linear with var_p, 2: quadratic with strips
  int depth;            // imax = jmax
  int same;             // tp == tl: strip terms carry zero weight
  double c_aux, c_p;    // border 1
  double c0, c1, c2;    // border 2: coarse value, strip t0, strip tm1
};

struct FbPoint;   // one fine-boundary point with its coarse cell (oaxis_device.h)
long fb_points(const KGrid& Gf, int depth);   // records of a zone of that depth
size_t fb_record_bytes();                     // sizeof(FbPoint)

// ------------------------------------------------------------ work items
// Each kernel of a substep runs over a table of work items in device memory,
// one item per (spacetime, level): everything one level's share of the
// launch needs.  Item k owns the blocks [block0, block0 + blocks) of the
// grid.  A spacetime alone launches tables of one item; a batch of
// spacetimes launches one table over all its members.
struct RhsItem {                  // one RK substep of a level
  KGrid G; KPhys P; KStep S;
  const double* yin; const double* yp; double* acc; double* yout; double* gaux;
  int nbi, nbj;                   // 16 x 16 tiles along r and z
  int block0;
};
struct ProjItem {                 // state projection (chi/phi, factordet) of a level
  KGrid G; KPhys P; int time0; double* y;
  int block0;
};
struct DelItem {                  // Delta = DelDef on a level at clock 0
  KGrid G; double* y;
  int block0;
};
struct SaveItem {                 // saveold of the fine-boundary strips
  KGrid G; int nf; const double* y; double* rR; double* zL; double* zR;
  int block0;
};
struct FbItem {                   // finebound of a fine level
  KGrid Gf, Gc; KPhys P; KTime T; int time0;
  const FbPoint* table; long n;   // the zone's point records
  const double* yc; double* yf;
  const double* rR0; const double* rR1; const double* zL0; const double* zL1; const double* zR0; const double* zR1;
  int block0;
};
struct RsItem {                   // restriction of a fine level onto its parent
  KGrid Gf, Gc; int nf, interporder; const double* yf; double* yc;
  // coarse points ic0 .. ic0 + nic - 1, jc0 .. jc0 + njc - 1 (inside the
  // coarse level): a box around the points the fine level restricts onto
  int ic0, nic, jc0, njc;
  int block0;
};

// blocks one item occupies in a launch (host side, for block0 prefix sums)
int rhs_blocks(const RhsItem& it);
int proj_blocks(const ProjItem& it);
int del_blocks(const DelItem& it);
int save_blocks(const SaveItem& it);
int fb_blocks(const FbItem& it, bool shift, bool scalar);
int rs_blocks(const RsItem& it);
// tiles of a level for the RHS kernel; OA_ERR_ARG if a direction has fewer
// than 16 points
int rhs_tiles(const KGrid& G, int* nbi, int* nbj);

// raise the LDS limits of the RHS kernels (once, before any launch)
int prepare_rhs_kernels();

// launchers (oaxis_kernels.hip): n items in device memory whose block0
// fields are the prefix sums of their blocks, and the total
// xcd_remap: renumber blocks so that each XCD takes one contiguous range of tiles
int launch_rhs_table(const RhsItem* items, int n, int blocks, bool shift, bool scalar, bool phimode,
                     bool any_dtb_pre, bool xcd_remap, void* stream);
int launch_project_table(const ProjItem* items, int n, int blocks, bool shift, bool scalar, void* stream);
int launch_deldef_table(const DelItem* items, int n, int blocks, void* stream);
int launch_saveold_table(const SaveItem* items, int n, int blocks, void* stream);
int launch_finebound_table(const FbItem* items, int n, int blocks, bool shift, bool scalar, void* stream);
int launch_restrict_table(const RsItem* items, int n, int blocks, void* stream);
// constraints of the state y (constraints.f90): the geometric auxiliaries
// into gaux, then ham, mom_r, mom_z, CDelta_r, CDelta_z planes into out;
// sc0: slot of scalar_phi when scalar
int launch_constraints(const KGrid& G, const KPhys& P, const KStep& S, bool phimode, bool scalar, int sc0,
                       const double* y, double* gaux, double* out, void* stream);

// 0D norms (save0Ddata): norms_blocks() partial records of 6 doubles
int norms_blocks();
int launch_norms_partial(const KGrid& G, const double* f, double* part, void* stream);
// the same norms for n fields of one level, with the partial records in the
// library's fixed device array on device dev (no allocation); out is device
// memory of 5 doubles per field (max, min, max|f|, rms, NaN flag)
int norms_max_fields();
int launch_norms_fields_noalloc(int dev, const KGrid& G, int n, const double* const* planes, double* out,
                                void* stream);
int launch_dtbeta_init(const KGrid& G, double csi, bool scalar, double* y, void* stream);
int launch_fb_table(const KGrid& Gf, const KGrid& Gc, int depth, int interporder, FbPoint* table, long n,
                    void* stream);

}  // namespace oax
