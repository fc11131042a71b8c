// Batched evolution of many independent spacetimes on one device.
//
// A coarse step of the batch is onestep(0) of every member, run level by
// level in lockstep: at each point of the Berger-Oliger recursion the members
// that have that level take part, and every kernel of that point (saveold,
// each RK substep, finebound, DelDef, restriction, projection) is ONE launch
// over the work items of all of them, grouped only by the template variant
// the item needs (shift, scalar field, phi mode).  Each member's own sequence
// of operations, and so its arithmetic, is that of onestep() on the member
// alone; members share no memory, so the interleaving between members does
// not matter.
//
// The whole coarse step is planned on the host first (plan_level over all
// members): the prepare_* pieces advance each member's bookkeeping and fill
// work items, which go into one arena, copied to the device in one transfer;
// the launches follow on the batch's stream.  A spacetime alone runs the
// same planner with itself as the only member, so it executes the same
// kernels on the same items.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "oa_state.h"
#include "oaxis.h"

using namespace oax;

namespace {

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return OA_OK;
  oa_set_error(std::string(what) + ": " + hipGetErrorString(e));
  return OA_ERR_HIP;
}

#define HIPCHK(call)                                  \
  do {                                                \
    int rc_ = hip_check((call), #call);               \
    if (rc_ != OA_OK) return rc_;                     \
  } while (0)

#define RC(call)                                      \
  do {                                                \
    int rc_ = (call);                                 \
    if (rc_ != OA_OK) return rc_;                     \
  } while (0)

}  // namespace

struct oa_batch {
  std::vector<oa_handle*> members;
  hipStream_t stream = nullptr;
  Plan plan;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  double last_ms = 0.0;
  double point_updates = 0.0;
  long last_launches = 0, last_items = 0;
  double rhs_ms = 0.0, rhs_points = 0.0;
  long rhs_launches = 0;
};

namespace {

int batch_step(oa_batch* b) {
  double before = 0.0;
  for (oa_handle* h : b->members) before += h->point_updates;
  b->plan.begin();
  RC(plan_level(b->plan, b->members, 0));
  double after = 0.0;
  for (oa_handle* h : b->members) after += h->point_updates;
  b->point_updates += after - before;
  b->last_launches = static_cast<long>(b->plan.launches.size());
  b->last_items = b->plan.items();
  return b->plan.submit(b->stream, true);
}

bool member_of(const oa_batch* b, const oa_handle* h) {
  return std::find(b->members.begin(), b->members.end(), h) != b->members.end();
}

}  // namespace

// ================================================================== C ABI
extern "C" int oa_batch_create(oa_batch** out) {
  if (!out) {
    oa_set_error("oa_batch_create: null argument");
    return OA_ERR_ARG;
  }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    oa_set_error("no HIP device available: this library runs on the GPU only");
    return OA_ERR_HIP;
  }
  oa_batch* b = new oa_batch();
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev_begin) != hipSuccess || hipEventCreate(&b->ev_end) != hipSuccess ||
      b->plan.init() != OA_OK) {
    oa_batch_destroy(b);
    oa_set_error("oa_batch_create: HIP stream/event creation failed");
    return OA_ERR_HIP;
  }
  *out = b;
  return OA_OK;
}

extern "C" int oa_batch_add(oa_batch* b, oa_handle* h) {
  if (!b || !h) {
    oa_set_error("oa_batch_add: null argument");
    return OA_ERR_ARG;
  }
  if (h->batch) {
    oa_set_error("oa_batch_add: the spacetime is already a member of a batch");
    return OA_ERR_STATE;
  }
  for (const Level& L : h->L)
    if (L.dirty || L.in_step) {
      oa_set_error("oa_batch_add: state not projected or inside a step; call oa_initial_aux first");
      return OA_ERR_STATE;
    }
  // the handle's queued work completes before the batch's stream takes over
  HIPCHK(hipStreamSynchronize(h->own_stream));
  h->stream = b->stream;
  h->batch = b;
  b->members.push_back(h);
  return OA_OK;
}

extern "C" int oa_batch_remove(oa_batch* b, oa_handle* h) {
  if (!b || !h || !member_of(b, h)) {
    oa_set_error("oa_batch_remove: not a member of this batch");
    return OA_ERR_ARG;
  }
  if (ah_pending(h)) {
    oa_set_error("oa_batch_remove: a begun horizon search of this spacetime has not ended (oa_ahfind_end)");
    return OA_ERR_STATE;
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  h->stream = h->own_stream;
  h->batch = nullptr;
  b->members.erase(std::find(b->members.begin(), b->members.end(), h));
  return OA_OK;
}

extern "C" int oa_batch_size(const oa_batch* b, int* n) {
  if (!b || !n) return OA_ERR_ARG;
  *n = static_cast<int>(b->members.size());
  return OA_OK;
}

extern "C" int oa_batch_coarse_step(oa_batch* b) {
  if (!b) return OA_ERR_ARG;
  return batch_step(b);
}

extern "C" int oa_batch_evolve(oa_batch* b, int nsteps) {
  if (!b || nsteps < 0) return OA_ERR_ARG;
  b->plan.reset_timing();
  HIPCHK(hipEventRecord(b->ev_begin, b->stream));
  for (int s = 0; s < nsteps; ++s) RC(batch_step(b));
  HIPCHK(hipEventRecord(b->ev_end, b->stream));
  HIPCHK(hipEventSynchronize(b->ev_end));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev_begin, b->ev_end));
  b->last_ms = ms;
  if (b->plan.kernel_timing) RC(b->plan.rhs_stats(&b->rhs_ms, &b->rhs_launches, &b->rhs_points));
  return OA_OK;
}

extern "C" int oa_batch_set_graphs(oa_batch* b, int on) {
  if (!b) return OA_ERR_ARG;
  b->plan.graphs = (on != 0);
  b->plan.tuner.enabled = false;     // the caller's choice stands
  return OA_OK;
}

extern "C" int oa_batch_schedule(const oa_batch* b, int* graphs, int* xcd_remap, int* tuned, double* step_ms) {
  if (!b) return OA_ERR_ARG;
  const Plan& pl = b->plan;
  if (graphs) *graphs = pl.graphs ? 1 : 0;
  if (xcd_remap) *xcd_remap = pl.xcd_remap ? 1 : 0;
  if (tuned) *tuned = pl.tuner.done ? 1 : 0;
  if (step_ms)
    for (int c = 0; c < Plan::kTuneCands; ++c) step_ms[c] = pl.tuner.ms[c];
  return OA_OK;
}

extern "C" int oa_batch_set_kernel_timing(oa_batch* b, int on) {
  if (!b) return OA_ERR_ARG;
  b->plan.kernel_timing = (on != 0);
  return OA_OK;
}

extern "C" int oa_batch_rhs_kernel_stats(const oa_batch* b, double* total_ms, long* launches, double* points) {
  if (!b) return OA_ERR_ARG;
  if (total_ms) *total_ms = b->rhs_ms;
  if (launches) *launches = b->rhs_launches;
  if (points) *points = b->rhs_points;
  return OA_OK;
}

extern "C" int oa_batch_sync(oa_batch* b) {
  if (!b) return OA_ERR_ARG;
  HIPCHK(hipStreamSynchronize(b->stream));
  return OA_OK;
}

extern "C" int oa_batch_last_elapsed_ms(const oa_batch* b, double* ms) {
  if (!b || !ms) return OA_ERR_ARG;
  *ms = b->last_ms;
  return OA_OK;
}

extern "C" int oa_batch_point_updates(const oa_batch* b, double* updates) {
  if (!b || !updates) return OA_ERR_ARG;
  *updates = b->point_updates;
  return OA_OK;
}

extern "C" int oa_batch_launch_stats(const oa_batch* b, long* launches, long* items) {
  if (!b) return OA_ERR_ARG;
  if (launches) *launches = b->last_launches;
  if (items) *items = b->last_items;
  return OA_OK;
}

extern "C" int oa_batch_destroy(oa_batch* b) {
  if (!b) return OA_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (oa_handle* h : b->members) {
    h->stream = h->own_stream;
    h->batch = nullptr;
  }
  b->members.clear();
  b->plan.release();
  if (b->ev_begin) (void)hipEventDestroy(b->ev_begin);
  if (b->ev_end) (void)hipEventDestroy(b->ev_end);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return OA_OK;
}
