// Host-side state of a spacetime handle and the pieces of one Berger-Oliger
// step, shared by the single-spacetime runtime and the batch runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "oa_internal.h"
#include "oa_plan.h"

namespace oax {

struct Level {
  KGrid G{};
  double* slab[3] = {nullptr, nullptr, nullptr};   // state buffers
  double* acc = nullptr;                            // RK4 accumulator
  double* strip[3] = {nullptr, nullptr, nullptr};  // fine-boundary time levels (ring)
  int ring[3] = {0, 1, 2};                          // ring[k]: strip set of time level -k
  int cur = 0;                                      // slab holding the current state
  int step_a = 0, last_out = 0;                     // roles inside a step
  bool in_step = false;
  bool dirty = true;                                // state needs the full projection
  bool phi_valid = true;                            // phi plane holds the reference's phi
  bool aux_phi = true;                              // last auxiliary ran at clock 0 (phi mode)
  double* gaux = nullptr;                           // geometric auxiliaries between the two RHS phases
  double* dtb_pre = nullptr;                        // dtbeta_r, dtbeta_z before the driver initialisation
  bool dtb_pre_valid = false;                       // the next substep differentiates dtb_pre
  double t = 0.0, t1 = 0.0, t2 = 0.0;
  int s = 0;
  long strip_r = 0, strip_z = 0;                    // doubles per strip set piece
  double* cons = nullptr;                           // ham, mom_r, mom_z, CDelta_r, CDelta_z (on demand)
  bool cons_valid = false;                          // cons holds the constraints of the current state
  FbPoint* fbtab[2] = {nullptr, nullptr};           // fine-boundary geometry: depth g+2, depth g-1
  long fbn[2] = {0, 0};
};

struct AhState;                                     // apparent-horizon finder (oaxis_ah.hip)

}  // namespace oax

struct oa_batch;

struct oa_handle {
  oa_params p{};
  int nlev = 0;
  int nf = 0;
  bool shift = false, scalar = false;
  double dt0 = 0.0;
  std::vector<oax::Level> L;
  int device = 0;                    // the HIP device holding the state
  hipStream_t stream = nullptr;      // the stream every call of this handle uses
  hipStream_t own_stream = nullptr;  // its own; a batch lends its stream while the handle is a member
  oa_batch* batch = nullptr;
  oax::KPhys phys{};
  oax::Plan plan;                    // launches of this handle's own calls
  double* norms_part = nullptr;      // partial records of the 0D norms
  oax::AhState* ah = nullptr;        // horizon finder state (on demand)
  // timing
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  double last_ms = 0.0;
  double rhs_avg_ms = 0.0;
  long rhs_launches = 0;
  double point_updates = 0.0;
  // initial data's elliptic solve: stopping rule and outcome
  int ell_stationary = 1;            // 1: the all-levels phase may end on the stationary rule
  int ell_steps[2] = {0, 0};         // iterations: base level alone, all levels
  double ell_res = 0.0;              // residual of the last iteration
};

namespace oax {

// The pieces of onestep.f90 for one level, split into the bookkeeping that
// fills a work item (prepare_*) and its launch, so that a batch can gather
// the items of many spacetimes into one launch.  Each prepare_* advances the
// handle's bookkeeping exactly as the launch it describes will advance the
// device state.
int niter_of(const oa_handle* h);
void substep_weights(const oa_handle* h, int level, int iter, double* dtw, double* w);
int prepare_saveold(oa_handle* h, int level, SaveItem* it);
int prepare_substep(oa_handle* h, int level, int iter, int niter, double dtw, double w, RhsItem* it,
                    bool* phimode);
bool prepare_finebound(oa_handle* h, int level, double dtw, FbItem* it);   // false on level 0
// auxiliary(level): *proj / *del say whether the projection / DelDef run
void prepare_auxiliary(oa_handle* h, int level, ProjItem* pj, bool* proj, DelItem* dd, bool* del);
void prepare_restrict(oa_handle* h, int level, RsItem* it);
// box of coarse points a restriction of the fine level Gf onto Gc may write
void restrict_box(const KGrid& Gf, const KGrid& Gc, int* ic0, int* nic, int* jc0, int* njc);
void finish_level_step(oa_handle* h, int level);   // advance the level clock, count point updates
// onestep(level) of every spacetime of `all` that has the level, in lockstep,
// appended to `plan` (oa_plan.cpp)
int plan_level(Plan& plan, const std::vector<oa_handle*>& all, int level);
void ah_release(oa_handle* h);                       // free the horizon finder's device memory
int ah_pending(const oa_handle* h);                  // begun horizon searches not yet ended
int ah_record_doubles();                             // horizon finder: doubles per grid point
double ah_fixed_bytes(int n_ah, int method);                     // horizon finder: per-handle device bytes besides those

}  // namespace oax
