// Launch plans: transfer of the work-item arena and the launches, and the
// lockstep Berger-Oliger schedule over a set of spacetimes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "oa_plan.h"
#include "oa_state.h"

namespace oax {

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

int issue(const LaunchRec& r, const char* dev, bool xcd, hipStream_t st) {
  const char* base = dev + r.offset;
  switch (r.kind) {
    case K_SAVE: return launch_saveold_table(reinterpret_cast<const SaveItem*>(base), r.n, r.blocks, st);
    case K_RHS:
      return launch_rhs_table(reinterpret_cast<const RhsItem*>(base), r.n, r.blocks, r.shift, r.scalar, r.phimode,
                              r.any_pre, xcd, st);
    case K_FB:
      return launch_finebound_table(reinterpret_cast<const FbItem*>(base), r.n, r.blocks, r.shift, r.scalar, st);
    case K_DEL: return launch_deldef_table(reinterpret_cast<const DelItem*>(base), r.n, r.blocks, st);
    case K_PROJ:
      return launch_project_table(reinterpret_cast<const ProjItem*>(base), r.n, r.blocks, r.shift, r.scalar, st);
    default: return launch_restrict_table(reinterpret_cast<const RsItem*>(base), r.n, r.blocks, st);
  }
}

int group_code(bool shift, bool scalar) { return (shift ? 2 : 0) + (scalar ? 1 : 0); }

}  // namespace

int Plan::init() {
  for (int k = 0; k < kSlots; ++k)
    HIPCHK(hipEventCreateWithFlags(&copied[k], hipEventDisableTiming));
  for (GraphSlot& g : gslot) HIPCHK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&tuner.t0));
  HIPCHK(hipEventCreate(&tuner.t1));
  return prepare_rhs_kernels();
}

namespace {

bool same_launches(const std::vector<LaunchRec>& a, const std::vector<LaunchRec>& b) {
  if (a.size() != b.size()) return false;
  for (size_t q = 0; q < a.size(); ++q) {
    const LaunchRec &x = a[q], &y = b[q];
    if (x.kind != y.kind || x.shift != y.shift || x.scalar != y.scalar || x.phimode != y.phimode ||
        x.offset != y.offset || x.n != y.n || x.blocks != y.blocks || x.any_pre != y.any_pre)
      return false;
  }
  return true;
}

void drop_graph(Plan::GraphSlot& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
  g.sig.clear();
  g.bytes = 0;
  g.dev = nullptr;
}

}  // namespace

// A larger device arena for the work items.  Every earlier launch on the
// stream has finished with the old arena once the stream is idle; the
// captured graphs of both slots copy into the old arena and launch kernels
// that read it, so they are destroyed with it (no exec outlives the memory
// it names) and are captured again on their next use.
int Plan::grow_arena(hipStream_t stream, size_t bytes) {
  if (dev) {
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipFree(dev));
  }
  for (GraphSlot& g : gslot) drop_graph(g);
  dev = nullptr;
  dev_cap = 0;
  const size_t cap = std::max(bytes * 2, static_cast<size_t>(64) << 10);
  HIPCHK(hipMalloc(reinterpret_cast<void**>(&dev), cap));
  dev_cap = cap;
  return OA_OK;
}

void Plan::release() {
  for (int k = 0; k < kSlots; ++k) {
    if (pinned[k]) (void)hipHostFree(pinned[k]);
    if (copied[k]) (void)hipEventDestroy(copied[k]);
    pinned[k] = nullptr;
    copied[k] = nullptr;
    pinned_cap[k] = 0;
  }
  for (GraphSlot& g : gslot) {
    drop_graph(g);
    if (g.pinned) (void)hipHostFree(g.pinned);
    if (g.done) (void)hipEventDestroy(g.done);
    g = GraphSlot{};
  }
  if (tuner.t0) (void)hipEventDestroy(tuner.t0);
  if (tuner.t1) (void)hipEventDestroy(tuner.t1);
  tuner.t0 = tuner.t1 = nullptr;
  if (dev) (void)hipFree(dev);
  dev = nullptr;
  dev_cap = 0;
  for (hipEvent_t e : kev) (void)hipEventDestroy(e);
  kev.clear();
}

int Plan::submit(hipStream_t stream, bool as_graph) {
  if (launches.empty()) return OA_OK;
  Tuner& t = tuner;
  if (!as_graph || !t.enabled || t.done || kernel_timing) return submit_once(stream, as_graph);
  if (t.rep == -2 || !same_launches(t.sig, launches)) {
    // wait for a step that repeats the previous one
    const bool repeat = (t.rep == -2) && same_launches(t.sig, launches);
    t.sig = launches;
    // a restart discards what was timed of the abandoned sequence
    t.acc = 0.0;
    for (double& m : t.ms) m = 0.0;
    if (!repeat) {
      t.rep = -2;
      return submit_once(stream, as_graph);
    }
    t.cand = 0;
    t.rep = -1;
    t.best_cand = -1;
  }
  graphs = (t.cand & 1) != 0;
  xcd_remap = (t.cand & 2) != 0;
  HIPCHK(hipEventRecord(t.t0, stream));
  RC(submit_once(stream, as_graph));
  HIPCHK(hipEventRecord(t.t1, stream));
  HIPCHK(hipEventSynchronize(t.t1));
  float m = 0.0f;
  HIPCHK(hipEventElapsedTime(&m, t.t0, t.t1));
  if (t.rep >= 0) t.acc += m;
  if (++t.rep == kTuneReps) {
    t.ms[t.cand] = t.acc / kTuneReps;
    if (t.best_cand < 0 || t.ms[t.cand] < t.best) {
      t.best = t.ms[t.cand];
      t.best_cand = t.cand;
    }
    t.acc = 0.0;
    t.rep = -1;
    if (++t.cand == kTuneCands) {
      t.done = true;
      graphs = (t.best_cand & 1) != 0;
      xcd_remap = (t.best_cand & 2) != 0;
    }
  }
  return OA_OK;
}

int Plan::submit_once(hipStream_t stream, bool as_graph) {
  const size_t bytes = arena.size();
  if (as_graph && graphs && !kernel_timing) {
    GraphSlot& g = gslot[gnext];
    // this slot's pinned buffer is free once its previous launch has finished
    if (g.pending) {
      HIPCHK(hipEventSynchronize(g.done));
      g.pending = false;
    }
    if (bytes > dev_cap) RC(grow_arena(stream, bytes));
    if (bytes > g.pinned_cap) {
      drop_graph(g);
      if (g.pinned) HIPCHK(hipHostFree(g.pinned));
      g.pinned = nullptr;
      g.pinned_cap = 0;
      const size_t cap = std::max(bytes * 2, static_cast<size_t>(64) << 10);
      HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&g.pinned), cap, hipHostMallocDefault));
      g.pinned_cap = cap;
    }
    if (!g.exec || g.dev != dev || g.bytes != bytes || g.xcd != xcd_remap || !same_launches(g.sig, launches)) {
      drop_graph(g);
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      int rc = hip_check(hipMemcpyAsync(dev, g.pinned, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
      for (size_t q = 0; rc == OA_OK && q < launches.size(); ++q) rc = issue(launches[q], dev, xcd_remap, stream);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(stream, &graph);
      if (rc != OA_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIPCHK(e);
      g.graph = graph;
      HIPCHK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
      g.sig = launches;
      g.bytes = bytes;
      g.dev = dev;
      g.xcd = xcd_remap;
      ++graph_captures;
    }
    std::memcpy(g.pinned, arena.data(), bytes);
    HIPCHK(hipGraphLaunch(g.exec, stream));
    HIPCHK(hipEventRecord(g.done, stream));
    g.pending = true;
    gnext ^= 1;
    ++graph_launches;
    return OA_OK;
  }
  const int k = slot;
  // the pinned buffer of this slot is free once its last transfer is done
  if (copied_pending[k]) {
    HIPCHK(hipEventSynchronize(copied[k]));
    copied_pending[k] = false;
  }
  if (bytes > pinned_cap[k]) {
    if (pinned[k]) HIPCHK(hipHostFree(pinned[k]));
    pinned[k] = nullptr;
    pinned_cap[k] = 0;
    const size_t cap = std::max(bytes, static_cast<size_t>(64) << 10);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&pinned[k]), cap, hipHostMallocDefault));
    pinned_cap[k] = cap;
  }
  if (bytes > dev_cap) RC(grow_arena(stream, bytes));
  std::memcpy(pinned[k], arena.data(), bytes);
  // stream order: this transfer starts after every earlier launch, which has
  // then finished reading the previous contents of the device arena
  HIPCHK(hipMemcpyAsync(dev, pinned[k], bytes, hipMemcpyHostToDevice, stream));
  HIPCHK(hipEventRecord(copied[k], stream));
  copied_pending[k] = true;
  slot = (slot + 1) % kSlots;
  for (const LaunchRec& r : launches) {
    const bool timed = kernel_timing && r.kind == K_RHS;
    if (timed) {
      if (kev_used + 2 > kev.size()) {
        for (int q = 0; q < 256; ++q) {
          hipEvent_t e;
          HIPCHK(hipEventCreate(&e));
          kev.push_back(e);
        }
        kev_points.resize(kev.size() / 2);
      }
      HIPCHK(hipEventRecord(kev[kev_used], stream));
    }
    RC(issue(r, dev, xcd_remap, stream));
    if (timed) {
      HIPCHK(hipEventRecord(kev[kev_used + 1], stream));
      kev_points[kev_used / 2] = r.points;
      kev_used += 2;
    }
  }
  return OA_OK;
}

int Plan::rhs_stats(double* total_ms, long* nlaunch, double* points) const {
  double sum = 0.0, pts = 0.0;
  for (size_t q = 0; q + 1 < kev_used; q += 2) {
    float m = 0.0f;
    HIPCHK(hipEventElapsedTime(&m, kev[q], kev[q + 1]));
    sum += m;
    pts += kev_points[q / 2];
  }
  if (total_ms) *total_ms = sum;
  if (nlaunch) *nlaunch = static_cast<long>(kev_used / 2);
  if (points) *points = pts;
  return OA_OK;
}

// onestep(level) (onestep.f90:78-469) of every spacetime in M that has the
// level, in lockstep: each kernel is one launch over all of their items.
int plan_level(Plan& plan, const std::vector<oa_handle*>& all, int level) {
  std::vector<oa_handle*> M;
  for (oa_handle* h : all)
    if (h->nlev > level) M.push_back(h);
  if (M.empty()) return OA_OK;
  const size_t nm = M.size();

  std::vector<SaveItem> sv(nm);
  for (size_t k = 0; k < nm; ++k) RC(prepare_saveold(M[k], level, &sv[k]));
  plan.add(K_SAVE, false, false, false, sv);

  std::vector<char> time0(nm);
  int nmax = 0;
  for (size_t k = 0; k < nm; ++k) {
    time0[k] = (M[k]->L[level].t == 0.0);
    nmax = std::max(nmax, niter_of(M[k]));
  }
  std::vector<double> dtws(nm);
  for (int iter = 1; iter <= nmax; ++iter) {
    std::vector<RhsItem> rhs[8];
    for (size_t k = 0; k < nm; ++k) {
      oa_handle* h = M[k];
      const int ni = niter_of(h);
      if (iter > ni) continue;
      double dtw, w;
      substep_weights(h, level, iter, &dtw, &w);
      dtws[k] = dtw;
      RhsItem it;
      bool pm = false;
      RC(prepare_substep(h, level, iter, ni, dtw, w, &it, &pm));
      rhs[group_code(h->shift, h->scalar) * 2 + (pm ? 1 : 0)].push_back(it);
    }
    for (int c = 0; c < 8; ++c) plan.add(K_RHS, (c & 4) != 0, (c & 2) != 0, (c & 1) != 0, rhs[c]);
    if (level > 0) {
      std::vector<FbItem> fb[4];
      for (size_t k = 0; k < nm; ++k) {
        oa_handle* h = M[k];
        if (iter > niter_of(h)) continue;
        FbItem it;
        if (prepare_finebound(h, level, dtws[k], &it)) fb[group_code(h->shift, h->scalar)].push_back(it);
      }
      for (int c = 0; c < 4; ++c) plan.add(K_FB, (c & 2) != 0, (c & 1) != 0, false, fb[c]);
    }
    std::vector<DelItem> del;
    for (size_t k = 0; k < nm; ++k) {
      oa_handle* h = M[k];
      if (iter > niter_of(h) || !time0[k]) continue;
      const Level& L = h->L[level];
      del.push_back(DelItem{L.G, L.slab[L.last_out], 0});
    }
    plan.add(K_DEL, false, false, false, del);
  }
  bool deeper = false;
  for (oa_handle* h : M) {
    finish_level_step(h, level);
    deeper = deeper || (h->nlev > level + 1);
  }
  if (deeper) {
    RC(plan_level(plan, M, level + 1));
    RC(plan_level(plan, M, level + 1));
  }
  if (level > 0) {
    std::vector<oa_handle*> R;
    for (oa_handle* h : M)
      if ((h->L[level].s % 2) == 0) R.push_back(h);
    std::vector<RsItem> rs(R.size());
    for (size_t k = 0; k < R.size(); ++k) prepare_restrict(R[k], level, &rs[k]);
    plan.add(K_RS, false, false, false, rs);
    std::vector<ProjItem> pj[4];
    std::vector<DelItem> del;
    for (oa_handle* h : R) {
      ProjItem p;
      DelItem d;
      bool proj = false, dd = false;
      prepare_auxiliary(h, level - 1, &p, &proj, &d, &dd);
      if (proj) pj[group_code(h->shift, h->scalar)].push_back(p);
      if (dd) del.push_back(d);
    }
    for (int c = 0; c < 4; ++c) plan.add(K_PROJ, (c & 2) != 0, (c & 1) != 0, false, pj[c]);
    plan.add(K_DEL, false, false, false, del);
  }
  return OA_OK;
}

}  // namespace oax
