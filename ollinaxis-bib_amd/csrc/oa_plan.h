// Launch planning shared by a single spacetime and a batch of them.
//
// Every kernel of a substep runs from a table of work items in device memory
// (one item per (spacetime, level)), so a spacetime alone and a spacetime in
// a batch execute the same kernels on the same items: their arithmetic, and
// so their results, are identical.  A plan collects the items of a sequence
// of launches in one host arena; submit() copies the arena to the device in
// one transfer (from pinned memory) and enqueues the launches in order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <vector>

#include "oa_internal.h"

namespace oax {

enum LaunchKind { K_SAVE, K_RHS, K_FB, K_DEL, K_PROJ, K_RS };

struct LaunchRec {
  int kind;
  bool shift, scalar, phimode;
  size_t offset;    // bytes into the arena
  int n;            // items
  int blocks;       // total blocks
  bool any_pre;     // an RHS item stages the pre-initialisation dtbeta
  double points;    // RHS: grid points (Nr x Nz) the launch advances one substep
};

inline int blocks_of(const RhsItem& it, bool, bool) { return rhs_blocks(it); }
inline int blocks_of(const SaveItem& it, bool, bool) { return save_blocks(it); }
inline int blocks_of(const FbItem& it, bool sh, bool sc) { return fb_blocks(it, sh, sc); }
inline int blocks_of(const DelItem& it, bool, bool) { return del_blocks(it); }
inline int blocks_of(const ProjItem& it, bool, bool) { return proj_blocks(it); }
inline int blocks_of(const RsItem& it, bool, bool) { return rs_blocks(it); }

struct Plan {
  static constexpr int kSlots = 8;        // pinned transfer buffers in flight
  static constexpr size_t kAlign = 256;

  std::vector<char> arena;
  std::vector<LaunchRec> launches;
  char* pinned[kSlots] = {};
  size_t pinned_cap[kSlots] = {};
  hipEvent_t copied[kSlots] = {};
  bool copied_pending[kSlots] = {};
  int slot = 0;
  char* dev = nullptr;
  size_t dev_cap = 0;
  // Whole coarse steps can run as captured HIP graphs: the transfer of the
  // arena plus every launch.  A graph is reused while the sequence of
  // launches (kinds, variants, item counts, grids, arena layout) repeats,
  // which it does from step to step once every level's clock is past 0; only
  // the work items change, and they reach the device through the graph's own
  // transfer.  Two graphs alternate, each with its own pinned source buffer,
  // so the host can prepare one step while the previous one runs.
  struct GraphSlot {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<LaunchRec> sig;
    size_t bytes = 0;
    char* pinned = nullptr;
    size_t pinned_cap = 0;
    char* dev = nullptr;          // device arena the graph copies to
    bool xcd = false;             // the RHS block numbering it was captured with
    hipEvent_t done = nullptr;    // recorded after the graph's last launch
    bool pending = false;
  };
  bool graphs = true;
  bool xcd_remap = true;          // RHS blocks renumbered per XCD
  GraphSlot gslot[2];
  // Schedule autotuner.  The schedule of a coarse step — launched as a graph
  // or kernel by kernel, RHS blocks renumbered per XCD or not — does not
  // change any arithmetic (the same kernels run on the same tiles and items),
  // so it is chosen by measurement: once the sequence of launches repeats,
  // each of the four schedules runs kTuneReps timed steps (after one untimed
  // step, which pays for a graph capture), and the fastest is kept for the
  // rest of the plan's life.
  static constexpr int kTuneCands = 4, kTuneReps = 2;
  struct Tuner {
    bool enabled = true, done = false;
    std::vector<LaunchRec> sig;   // the repeating step being timed
    int cand = 0, rep = -2;       // rep -2: waiting for the step to repeat; -1: untimed step
    double acc = 0.0, best = 0.0;
    int best_cand = 0;
    double ms[kTuneCands] = {};
    hipEvent_t t0 = nullptr, t1 = nullptr;
  } tuner;
  int gnext = 0;
  long graph_launches = 0, graph_captures = 0;
  // optional HIP-event timing of the RHS launches (direct launches only)
  bool kernel_timing = false;
  std::vector<hipEvent_t> kev;
  std::vector<double> kev_points;
  size_t kev_used = 0;

  int init();                  // events; OA_ERR_HIP on failure
  void release();              // frees everything (the stream must be idle)
  void begin() {
    arena.clear();
    launches.clear();
  }
  template <class Item>
  void add(int kind, bool shift, bool scalar, bool phimode, std::vector<Item>& items) {
    if (items.empty()) return;
    int total = 0;
    bool any_pre = false;
    double points = 0.0;
    for (Item& it : items) {
      it.block0 = total;
      total += blocks_of(it, shift, scalar);
    }
    if constexpr (std::is_same<Item, RhsItem>::value)
      for (const Item& it : items) {
        any_pre = any_pre || (it.S.dtb_pre != nullptr);
        points += static_cast<double>(it.G.Nr) * static_cast<double>(it.G.Nz);
      }
    const size_t off = (arena.size() + kAlign - 1) / kAlign * kAlign;
    arena.resize(off + sizeof(Item) * items.size());
    std::memcpy(arena.data() + off, items.data(), sizeof(Item) * items.size());
    launches.push_back(LaunchRec{kind, shift, scalar, phimode, off, static_cast<int>(items.size()), total,
                                 any_pre, points});
  }
  template <class Item>
  void add_one(int kind, bool shift, bool scalar, bool phimode, const Item& it) {
    std::vector<Item> v(1, it);
    add(kind, shift, scalar, phimode, v);
  }
  // transfer the arena and enqueue the launches; as_graph: a whole coarse
  // step that may run as a captured graph
  int submit(hipStream_t stream, bool as_graph = false);
  int submit_once(hipStream_t stream, bool as_graph);
  int grow_arena(hipStream_t stream, size_t bytes);
  // RHS timing over the launches since the last reset
  void reset_timing() { kev_used = 0; }
  int rhs_stats(double* total_ms, long* launches, double* points) const;
  long items() const {
    long n = 0;
    for (const LaunchRec& r : launches) n += r.n;
    return n;
  }
};

}  // namespace oax
