// HIP kernels of the evolution hot path (gfx950, FP64 throughout).
//
// One launch of k_rhs performs, for every point of a level, what the
// reference does in one internal iteration of onestep.f90:129-342:
//   the derivative stencils and algebra of auxiliary_geometry.f90,
//   sources_geometry/lapse/shift/scalar, Kreiss-Oliger dissipation,
//   radiative boundaries on level 0, RK4 accumulate / update,
//   the axis (and equator) parity images, and the pointwise part of the
//   next auxiliary call (factordet, phi/chi, KT_H / lambda are implicit).
// The state stays projected in memory, as in the reference.
//
// Layout: a level holds each evolved field as one z-fastest plane; point
// (i, j) of the reference's (1-g:Nr, 1-g:Nz) range sits at
// (i+g-1)*ld + (j+g-1).  Ghost columns i <= 0 (axis) and, with equatorial
// symmetry, ghost rows j <= 0 are stored and kept equal to the parity images
// of the interior, so every centred stencil reads them directly.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "oa_internal.h"

#define OA_HD __device__ __forceinline__
#define OA_HHD __host__ __device__ inline
#include "oaxis_device.h"

namespace oax {

#define RC_K(call)                \
  do {                            \
    const int rc_ = (call);       \
    if (rc_ != OA_OK) return rc_; \
  } while (0)

// ------------------------------------------------------------ RHS kernel
// Each 16 x 16 block stages its tile of every field, with a 3-point halo, in
// LDS and reads all stencils from there.  Centred stencils reach 3 points.
// The one-sided closures at i = Nr-1..Nr, j = Nz-1..Nz and (without
// equatorial symmetry) j = 1-ghost..2-ghost reach up to 5 points, always
// towards the interior; the first and the last tile of each direction are
// anchored to the ends of the level, so those reads stay inside the tile.
// The last tile of a direction may overlap the one before it; its points
// already covered there are skipped (the RK accumulator is read-modify-write).
constexpr int kTile = 16, kHalo = 3;
constexpr int kTileW = kTile + 2 * kHalo;                 // staged extent per direction
constexpr int kTileSize = kTileW * kTileW;                // values per field

// first point of tile b of nb along [lo, hi] (extent >= kTile)
__host__ __device__ inline int tile_start(int b, int nb, int lo, int hi) {
  return (b < nb - 1) ? lo + b * kTile : hi - (kTile - 1);
}

// A tile wholly inside i in [1, Nr-3], j in [1, Nz-3]: every stencil there
// takes its centred form, so the closure branches compile out (IN).
__device__ __forceinline__ bool interior_tile(const KGrid& G, int i0, int j0) {
  return i0 >= 1 && i0 + kTile - 1 <= G.Nr - 3 && j0 >= 1 && j0 + kTile - 1 <= G.Nz - 3;
}

// Staging an RHS input tile into LDS.  Each thread owns the same two points
// of every staged plane, q0 = tid and q1 = tid + 256 (the tile has 484), so
// their offsets in a plane are worked out once.  Points of the tile outside
// the level, which no stencil reads, take the value of the nearest point
// inside it, so that every load is in bounds and none needs a branch.
struct StageSlots {
  long o0, o1;          // plane offsets of the two points
  int q0, q1;           // their tile positions
  bool two;             // q1 lies inside the tile
};

__device__ __forceinline__ StageSlots stage_slots(const KGrid& G, int ib, int jb) {
  StageSlots s;
  const int tid = static_cast<int>(threadIdx.y) * kTile + static_cast<int>(threadIdx.x);
  const int lo = 1 - G.g;
  s.q0 = tid;
  s.q1 = tid + kTile * kTile;
  s.two = s.q1 < kTileSize;
  auto offset = [&](int q) {
    const int ii = min(max(ib + q / kTileW, lo), G.Nr), jj = min(max(jb + q % kTileW, lo), G.Nz);
    return IDX(G, ii, jj);
  };
  s.o0 = offset(s.q0);
  s.o1 = s.two ? offset(s.q1) : s.o0;
  return s;
}

// Stage NST planes, plane f from src(f), into tile slots first .. first +
// NST - 1.  Every load of the batch is issued before the first LDS store, so
// their latencies overlap instead of adding up (a load-wait-store loop per
// plane waits out one memory round trip per plane and point).
template <int NST, class Src>
__device__ __forceinline__ void stage_planes(double* tile, int first, const StageSlots& s, Src src) {
  double v0[NST], v1[NST];
#pragma unroll
  for (int f = 0; f < NST; ++f) {
    const double* __restrict__ p = src(f);
    v0[f] = p[s.o0];
    v1[f] = p[s.o1];
  }
#pragma unroll
  for (int f = 0; f < NST; ++f) {
    double* dst = tile + (first + f) * kTileSize;
    dst[s.q0] = v0[f];
    if (s.two) dst[s.q1] = v1[f];
  }
}

// Pointers read from a work item are generic; every one of them addresses
// device (global) memory and is uniform across the block.  Passing each
// through the global address space, behind an empty asm so the two casts do
// not fold away, lets the compiler emit global loads and stores instead of
// flat ones (flat accesses also count against the LDS wait counter, so every
// LDS wait would also wait for outstanding global traffic).
template <class T>
__device__ __forceinline__ T* gmem(T* p) {
  __attribute__((address_space(1))) T* g = (__attribute__((address_space(1))) T*)p;
  asm volatile("" : "+s"(g));
  return (T*)g;
}

// The tile address a thread's stencils are read relative to: its point's
// value, kHalo rows and columns back (di, dj: the point's position in the
// staged tile).  The centred stencils then reach offsets 0 .. 6 rows and
// columns from it, all non-negative, so every read is one LDS instruction
// with an immediate offset; the empty asm keeps the compiler from splitting
// the address back into a tile base and a signed per-read offset.
__device__ __forceinline__ const double* corner_of(const double* tile, int di, int dj) {
  auto l = (__attribute__((address_space(3))) const double*)(tile + (di - kHalo) * kTileW + (dj - kHalo));
  asm volatile("" : "+v"(l));
  return (const double*)l;
}

// the item of block b in a table launch: the last one whose block0 <= b
// (block0 ascends; the loads are uniform across the block)
template <class Item>
__device__ __forceinline__ const Item& find_item(const Item* __restrict__ items, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return items[lo];
}

// One RK substep of a level: each 16 x 16 block stages its tile and runs two
// phases per point.  Phase 1 forms the geometric auxiliaries (the Ricci
// tensor and div KT: the longest expressions, which need only the vacuum
// fields) and passes them through the per-level buffer gaux; phase 2 forms
// D_i D_j alpha, every source, the dissipation, the radiative boundaries and
// the RK update.  Kept apart, the phases need at most the registers of the
// larger one.
// Occupancy of the RHS kernel.  With the Gamma-driver or the scalar field the
// staged tile (17 or 21 fields, 66-81 KB) allows two blocks per CU, and the
// two phases fit 256 VGPRs: two waves per SIMD.  In vacuum the tile (13
// fields, 50 KB) allows three blocks, and phase 1 split in two halves fits
// 168 VGPRs: three waves per SIMD.
template <bool SHIFT, bool SCALAR>
constexpr bool kSplitPhase1 = !SHIFT && !SCALAR;
template <bool SHIFT, bool SCALAR>
constexpr int kRhsBlocksPerCu = kSplitPhase1<SHIFT, SCALAR> ? 3 : 2;

template <bool SHIFT, bool SCALAR, bool PHIMODE>
__device__ __forceinline__ void rhs_tile(const RhsItem& it, int bi, int bj, double* tile) {
  constexpr int nf = nslots(SHIFT, SCALAR);
  const KGrid& G = it.G;
  KStep S = it.S;
  S.dtb_pre = gmem(S.dtb_pre);            // null stays null
  const int nbj = it.nbj, nbi = it.nbi;
  const int j0 = tile_start(bj, nbj, G.jlo, G.Nz);
  const int i0 = tile_start(bi, nbi, 1, G.Nr);
  const int ib = i0 - kHalo, jb = j0 - kHalo;
  // the evolved fields, then (first substep only) the dtbeta planes it differentiates
  const double* __restrict__ yin = gmem(it.yin);
  double* const gaux = gmem(it.gaux);
  {
    const StageSlots ss = stage_slots(G, ib, jb);
    stage_planes<nf>(tile, 0, ss, [&](int f) { return yin + f * G.plane; });
    if (S.dtb_pre) stage_planes<2>(tile, nf, ss, [&](int f) { return S.dtb_pre + f * G.plane; });
    __syncthreads();
  }
  const int i = i0 + static_cast<int>(threadIdx.y), j = j0 + static_cast<int>(threadIdx.x);
  if (bi == nbi - 1 && i < 1 + (nbi - 1) * kTile) return;        // done by the tile before
  if (bj == nbj - 1 && j < G.jlo + (nbj - 1) * kTile) return;
  const TileView V{corner_of(tile, i - ib, j - jb), i - kHalo, j - kHalo, kTileW, kTileSize, nf};
  // phase 1: the geometric auxiliaries of this point, through gaux (own point
  // only, so no synchronisation); the memory barrier keeps the compiler from
  // carrying phase-1 values or loads into phase 2, so the two phases do not
  // add up in registers
  if constexpr (kSplitPhase1<SHIFT, SCALAR>) {
    // vacuum: phase 1 in its two halves (the Ricci tensor, then div KT),
    // so that no part needs more than 168 VGPRs: three waves per SIMD
    if (interior_tile(G, i0, j0)) {
      rhs_point<false, false, PHIMODE, true, 11>(G, it.P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
      asm volatile("" ::: "memory");
      rhs_point<false, false, PHIMODE, true, 12>(G, it.P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
    } else {
      rhs_point<false, false, PHIMODE, false, 11>(G, it.P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
      asm volatile("" ::: "memory");
      rhs_point<false, false, PHIMODE, false, 12>(G, it.P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
    }
  } else {
    if (interior_tile(G, i0, j0))
      rhs_point<false, false, PHIMODE, true, 1>(G, it.P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
    else
      rhs_point<false, false, PHIMODE, false, 1>(G, it.P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
  }
  asm volatile("" ::: "memory");
  rhs_point<SHIFT, SCALAR, PHIMODE, false, 2>(G, it.P, S, V, gmem(it.yp), gmem(it.acc), gmem(it.yout), gaux, i, j);
}

// Workgroups are dispatched to the 8 XCDs round robin (block b to XCD b % 8),
// and each XCD has its own L2.  Renumbering the blocks so that every XCD
// receives one contiguous range of tiles keeps neighbouring tiles, whose
// halos overlap, in the same L2.  A bijection of [0, total).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int total) {
  const int q = total / kXcds, r = total % kXcds;
  const int x = b % kXcds, l = b / kXcds;
  return x * q + (x < r ? x : r) + l;
}

template <bool SHIFT, bool SCALAR, bool PHIMODE>
__global__ void __launch_bounds__(kTile * kTile, (kRhsBlocksPerCu<SHIFT, SCALAR>)) k_rhs_table(const RhsItem* __restrict__ items, int n,
                                                                int xcd_remap) {
  extern __shared__ double tile[];
  const int b = xcd_remap ? xcd_block(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x))
                          : static_cast<int>(blockIdx.x);
  const RhsItem& it = find_item(items, n, b);
  const int t = b - it.block0;
  rhs_tile<SHIFT, SCALAR, PHIMODE>(it, t / it.nbj, t % it.nbj, tile);
}

// ------------------------------------------------------ constraints
// The geometric auxiliaries of the current state (phase 1 of the RHS kernel
// alone, on the vacuum fields), for the constraints.
template <bool PHIMODE>
__global__ void __launch_bounds__(kTile * kTile, 2) k_aux(RhsItem it) {
  extern __shared__ double tile[];
  const KGrid& G = it.G;
  const int bi = static_cast<int>(blockIdx.y), bj = static_cast<int>(blockIdx.x);
  const int j0 = tile_start(bj, it.nbj, G.jlo, G.Nz);
  const int i0 = tile_start(bi, it.nbi, 1, G.Nr);
  const int ib = i0 - kHalo, jb = j0 - kHalo;
  const double* __restrict__ yin = it.yin;
  stage_planes<NBASE>(tile, 0, stage_slots(G, ib, jb), [&](int f) { return yin + f * G.plane; });
  __syncthreads();
  const int i = i0 + static_cast<int>(threadIdx.y), j = j0 + static_cast<int>(threadIdx.x);
  if (bi == it.nbi - 1 && i < 1 + (it.nbi - 1) * kTile) return;     // done by the tile before
  if (bj == it.nbj - 1 && j < G.jlo + (it.nbj - 1) * kTile) return;
  const TileView V{corner_of(tile, i - ib, j - jb), i - kHalo, j - kHalo, kTileW, kTileSize, NBASE};
  if (interior_tile(G, i0, j0))
    rhs_point<false, false, PHIMODE, true, 1>(G, it.P, it.S, V, nullptr, nullptr, nullptr, it.gaux, i, j);
  else
    rhs_point<false, false, PHIMODE, false, 1>(G, it.P, it.S, V, nullptr, nullptr, nullptr, it.gaux, i, j);
}

struct ConsArgs {
  KGrid G; KPhys P; int sc0;
  const double* y; const double* gaux; double* out;
};

template <bool PHIMODE, bool SCALAR>
__global__ void __launch_bounds__(256) k_constraints(ConsArgs c) {
  const int j = c.G.jlo + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = 1 + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
  if (j > c.G.Nz || i > c.G.Nr) return;
  constraints_point<PHIMODE, SCALAR>(c.G, c.P, c.sc0, c.y, c.gaux, c.out, i, j);
}

// constraints.f90:68-93, in the reference's order: ham copied from the
// neighbouring row at i = 1-ghost, i = Nr, j = 1-ghost, j = Nz, then the
// axis images and (equatorial symmetry) the equator images.  One block;
// each step sees the previous one's writes.
__global__ void __launch_bounds__(256) k_ham_fix(KGrid G, double* __restrict__ ham) {
  const int t = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
  const int g = G.g;
  for (int j = 1 - g + t; j <= G.Nz; j += nt) ham[IDX(G, 1 - g, j)] = ham[IDX(G, 2 - g, j)];
  __syncthreads();
  for (int j = 1 - g + t; j <= G.Nz; j += nt) ham[IDX(G, G.Nr, j)] = ham[IDX(G, G.Nr - 1, j)];
  __syncthreads();
  for (int i = 1 - g + t; i <= G.Nr; i += nt) ham[IDX(G, i, 1 - g)] = ham[IDX(G, i, 2 - g)];
  __syncthreads();
  for (int i = 1 - g + t; i <= G.Nr; i += nt) ham[IDX(G, i, G.Nz)] = ham[IDX(G, i, G.Nz - 1)];
  __syncthreads();
  for (int q = t; q < g * (G.Nz + g); q += nt) {
    const int i = 1 + q / (G.Nz + g), j = 1 - g + q % (G.Nz + g);
    ham[IDX(G, 1 - i, j)] = ham[IDX(G, i, j)];
  }
  __syncthreads();
  if (G.eqsym)
    for (int q = t; q < g * (G.Nr + g); q += nt) {
      const int j = 1 + q / (G.Nr + g), i = 1 - g + q % (G.Nr + g);
      ham[IDX(G, i, 1 - j)] = ham[IDX(G, i, j)];
    }
}

// ----------------------------------------------- pointwise kernels
// One thread per point (or record) of an item, 256 threads per block; q is
// the thread's index within its item.
constexpr int kPtBlock = 256;

__device__ __forceinline__ bool level_point(const KGrid& G, long q, int& i, int& j) {
  const int nj = G.Nz - G.jlo + 1;
  i = 1 + static_cast<int>(q / nj);
  j = G.jlo + static_cast<int>(q % nj);
  return i <= G.Nr;
}

template <bool SHIFT, bool SCALAR>
__device__ __forceinline__ void project_body(const ProjItem& it, long q) {
  int i, j;
  if (level_point(it.G, q, i, j)) project_point<SHIFT, SCALAR>(it.G, it.P, it.time0, gmem(it.y), i, j);
}

__device__ __forceinline__ void deldef_body(const DelItem& it, long q) {
  int i, j;
  if (level_point(it.G, q, i, j)) deldef_point(it.G, gmem(it.y), i, j);
}

// saveold for the fine-boundary strips: blocks of one field, then the next
__device__ __forceinline__ int save_per_field(const KGrid& G) {
  const int nrt = G.Nr + G.g, nzt = G.Nz + G.g;
  return G.g * (nrt > nzt ? nrt : nzt);
}
__device__ __forceinline__ void saveold_body(const SaveItem& it, int blk) {
  const int per = save_per_field(it.G);
  const int bpf = (per + kPtBlock - 1) / kPtBlock;
  const int f = blk / bpf;
  const int q = (blk % bpf) * kPtBlock + static_cast<int>(threadIdx.x);
  if (f < it.nf && q < per) saveold_point(it.G, gmem(it.y), gmem(it.rR), gmem(it.zL), gmem(it.zR), f, q);
}

// finebound: blocks of one field group, then the next (the group stays
// uniform within a block)
template <bool SHIFT, bool SCALAR>
__device__ __forceinline__ void finebound_body(const FbItem& it, int blk) {
  const int bpg = static_cast<int>((it.n + kPtBlock - 1) / kPtBlock);
  const int grp = blk / bpg;
  const long q = static_cast<long>(blk % bpg) * kPtBlock + threadIdx.x;
  if (grp >= finebound_groups(SHIFT, SCALAR) || q >= it.n) return;
  finebound_point<SHIFT, SCALAR>(it.Gf, it.Gc, it.P, it.T, it.time0, gmem(it.yc), gmem(it.yf), gmem(it.rR0),
                                 gmem(it.rR1), gmem(it.zL0), gmem(it.zL1), gmem(it.zR0), gmem(it.zR1),
                                 gmem(it.table)[q], grp);
}

// restrict: blocks of one field, then the next; one thread per coarse point
// of the item's box (the points of the coarse level the fine level may
// restrict onto; restrict_point tests each against the exact bounds)
__device__ __forceinline__ void restrict_body(const RsItem& it, int blk) {
  const long per = static_cast<long>(it.nic) * it.njc;
  const int bpf = static_cast<int>((per + kPtBlock - 1) / kPtBlock);
  const int s = blk / bpf;
  const long q = static_cast<long>(blk % bpf) * kPtBlock + threadIdx.x;
  if (s >= it.nf || q >= per) return;
  const int ic = it.ic0 + static_cast<int>(q / it.njc), jc = it.jc0 + static_cast<int>(q % it.njc);
  restrict_point(it.Gf, it.Gc, s, it.interporder, gmem(it.yf), gmem(it.yc), ic, jc);
}

template <bool SHIFT, bool SCALAR>
__global__ void __launch_bounds__(kPtBlock) k_project_table(const ProjItem* __restrict__ items, int n) {
  const ProjItem& it = find_item(items, n, static_cast<int>(blockIdx.x));
  project_body<SHIFT, SCALAR>(it, static_cast<long>(blockIdx.x - it.block0) * kPtBlock + threadIdx.x);
}

__global__ void __launch_bounds__(kPtBlock) k_deldef_table(const DelItem* __restrict__ items, int n) {
  const DelItem& it = find_item(items, n, static_cast<int>(blockIdx.x));
  deldef_body(it, static_cast<long>(blockIdx.x - it.block0) * kPtBlock + threadIdx.x);
}

__global__ void __launch_bounds__(kPtBlock) k_saveold_table(const SaveItem* __restrict__ items, int n) {
  const SaveItem& it = find_item(items, n, static_cast<int>(blockIdx.x));
  saveold_body(it, static_cast<int>(blockIdx.x) - it.block0);
}

template <bool SHIFT, bool SCALAR>
__global__ void __launch_bounds__(kPtBlock) k_finebound_table(const FbItem* __restrict__ items, int n) {
  const FbItem& it = find_item(items, n, static_cast<int>(blockIdx.x));
  finebound_body<SHIFT, SCALAR>(it, static_cast<int>(blockIdx.x) - it.block0);
}

__global__ void __launch_bounds__(kPtBlock) k_restrict_table(const RsItem* __restrict__ items, int n) {
  const RsItem& it = find_item(items, n, static_cast<int>(blockIdx.x));
  restrict_body(it, static_cast<int>(blockIdx.x) - it.block0);
}

// 0D norms of one field over i in [1, Nr-g], j in [1, Nz-g] (the range
// save0Ddata.f90:221-246 reduces): per block max, min, max|f|, sum f^2,
// point count and a NaN flag, in a fixed tree order; the host combines the
// blocks in block order.
constexpr int kNormBlocks = 256;
// partial record b of one field: a grid-stride pass of block b of nb
__device__ __forceinline__ void norms_block(const KGrid& G, const double* __restrict__ f, double* __restrict__ part,
                                            int b, int nb) {
  __shared__ double smax[kPtBlock], smin[kPtBlock], sabs[kPtBlock], ssq[kPtBlock], scnt[kPtBlock], snan[kPtBlock];
  const int ni = G.Nr - G.g, nj = G.Nz - G.g;
  const long n = static_cast<long>(ni) * nj;
  double vmax = -1.0e300, vmin = 1.0e300, vabs = 0.0, sq = 0.0, cnt = 0.0, nan = 0.0;
  for (long q = static_cast<long>(b) * kPtBlock + threadIdx.x; q < n; q += static_cast<long>(nb) * kPtBlock) {
    const int i = 1 + static_cast<int>(q / nj), j = 1 + static_cast<int>(q % nj);
    const double v = f[IDX(G, i, j)];
    if (v != v) nan = 1.0;
    vmax = fmax(vmax, v);
    vmin = fmin(vmin, v);
    vabs = fmax(vabs, fabs(v));
    sq += v * v;
    cnt += 1.0;
  }
  const int t = static_cast<int>(threadIdx.x);
  smax[t] = vmax; smin[t] = vmin; sabs[t] = vabs; ssq[t] = sq; scnt[t] = cnt; snan[t] = nan;
  __syncthreads();
  for (int h = kPtBlock / 2; h > 0; h >>= 1) {
    if (t < h) {
      smax[t] = fmax(smax[t], smax[t + h]);
      smin[t] = fmin(smin[t], smin[t + h]);
      sabs[t] = fmax(sabs[t], sabs[t + h]);
      ssq[t] += ssq[t + h];
      scnt[t] += scnt[t + h];
      snan[t] = fmax(snan[t], snan[t + h]);
    }
    __syncthreads();
  }
  if (t == 0) {
    double* o = part + 6 * b;
    o[0] = smax[0]; o[1] = smin[0]; o[2] = sabs[0]; o[3] = ssq[0]; o[4] = scnt[0]; o[5] = snan[0];
  }
}

__global__ void __launch_bounds__(kPtBlock) k_norms_partial(KGrid G, const double* __restrict__ f,
                                                            double* __restrict__ part) {
  norms_block(G, f, part, static_cast<int>(blockIdx.x), kNormBlocks);
}

// The norms of several fields of one level: blockIdx.y picks the field,
// whose kNormBlocks partial records start at part + y * 6 * kNormBlocks.
constexpr int kNormsMaxFields = 32;
struct NormsFields {
  KGrid G;
  int n;
  const double* f[kNormsMaxFields];
};

__global__ void __launch_bounds__(kPtBlock) k_norms_fields(NormsFields a, double* __restrict__ part) {
  const int y = static_cast<int>(blockIdx.y);
  norms_block(a.G, a.f[y], part + static_cast<long>(6) * kNormBlocks * y, static_cast<int>(blockIdx.x),
              kNormBlocks);
}

// Thread y combines the partial records of field y in block order, as the
// host does for oa_norms0d: out[5y..5y+4] = max, min, max|f|, rms, NaN flag.
__global__ void __launch_bounds__(64) k_norms_combine(int n, const double* __restrict__ part,
                                                      double* __restrict__ out) {
  const int y = static_cast<int>(threadIdx.x);
  if (y >= n) return;
  const double* p = part + static_cast<long>(6) * kNormBlocks * y;
  double vmax = -1.0e300, vmin = 1.0e300, vabs = 0.0, sq = 0.0, cnt = 0.0, nan = 0.0;
  for (int b = 0; b < kNormBlocks; ++b) {
    const double* o = p + 6 * b;
    vmax = fmax(vmax, o[0]);
    vmin = fmin(vmin, o[1]);
    vabs = fmax(vabs, o[2]);
    sq += o[3];
    cnt += o[4];
    nan = fmax(nan, o[5]);
  }
  double* r = out + 5 * y;
  r[0] = vmax;
  r[1] = vmin;
  r[2] = vabs;
  r[3] = (cnt > 0.0) ? sqrt(sq / cnt) : 0.0;
  r[4] = nan;
}

// dtbeta = drivercsi * Delta (evolve.f90:174-195), whole arrays incl. images
__global__ void k_dtbeta_init(long n, double csi, const double* __restrict__ dr, const double* __restrict__ dz,
                              double* __restrict__ tr, double* __restrict__ tz) {
  const long q = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  tr[q] = csi * dr[q];
  tz[q] = csi * dz[q];
}

// the fixed geometry of the fine boundary zone of depth D (once per level)
__global__ void __launch_bounds__(256) k_fb_table(KGrid Gf, KGrid Gc, int D, int interporder, long n,
                                                  FbPoint* __restrict__ table) {
  const long q = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  FbPoint fp;
  fb_geometry(Gf, Gc, D, interporder, q, fp);
  table[q] = fp;
}

// ------------------------------------------------------------- launchers
static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    oa_set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return OA_ERR_HIP;
  }
  return OA_OK;
}

static int blocks_for(long threads) { return static_cast<int>((threads + kPtBlock - 1) / kPtBlock); }

int rhs_tiles(const KGrid& G, int* nbi, int* nbj) {
  const int nj = G.Nz - G.jlo + 1;
  if (G.Nr < kTile || nj < kTile) {
    oa_set_error("RHS kernel: a level needs at least 16 points in each direction");
    return OA_ERR_ARG;
  }
  *nbi = (G.Nr + kTile - 1) / kTile;
  *nbj = (nj + kTile - 1) / kTile;
  return OA_OK;
}

int rhs_blocks(const RhsItem& it) { return it.nbi * it.nbj; }
int proj_blocks(const ProjItem& it) { return blocks_for(static_cast<long>(it.G.Nr) * (it.G.Nz - it.G.jlo + 1)); }
int del_blocks(const DelItem& it) { return blocks_for(static_cast<long>(it.G.Nr) * (it.G.Nz - it.G.jlo + 1)); }
int save_blocks(const SaveItem& it) {
  const int nrt = it.G.Nr + it.G.g, nzt = it.G.Nz + it.G.g;
  return it.nf * blocks_for(static_cast<long>(it.G.g) * (nrt > nzt ? nrt : nzt));
}
// Sized for the deeper of the two zones whichever this call fills, so that
// the grid of a finebound launch does not depend on whether the clocks meet
// (a floating comparison); the blocks past the item's own records exit.
int fb_blocks(const FbItem& it, bool shift, bool scalar) {
  return finebound_groups(shift, scalar) * blocks_for(fb_count(it.Gf, it.Gf.g + 2));
}
int rs_blocks(const RsItem& it) { return it.nf * blocks_for(static_cast<long>(it.nic) * it.njc); }

template <class K>
static int raise_lds_limit(K* kernel, size_t bytes) {
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          static_cast<int>(bytes)) != hipSuccess) {
    oa_set_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed for an RHS kernel");
    return OA_ERR_HIP;
  }
  return OA_OK;
}

// the dynamic LDS of one RHS block exceeds the 64 KiB default; two more
// planes on a level's first substep (KStep::dtb_pre)
template <bool SH, bool SC, bool PM>
static int rhs_lds(bool dtb_pre, size_t* lds) {
  const size_t per = sizeof(double) * kTileSize;
  static bool attr_set = false;
  if (!attr_set) {
    RC_K(raise_lds_limit(&k_rhs_table<SH, SC, PM>, per * static_cast<size_t>(nslots(SH, SC) + 2)));
    attr_set = true;
  }
  *lds = per * static_cast<size_t>(nslots(SH, SC) + (dtb_pre ? 2 : 0));
  return OA_OK;
}

template <bool SH, bool SC, bool PM>
static int launch_rhs_table_t(const RhsItem* items, int n, int blocks, bool any_pre, bool xcd, hipStream_t st) {
  size_t lds = 0;
  RC_K((rhs_lds<SH, SC, PM>(any_pre, &lds)));
  hipLaunchKernelGGL((k_rhs_table<SH, SC, PM>), dim3(blocks), dim3(kTile, kTile, 1), lds, st, items, n, xcd ? 1 : 0);
  return check_launch();
}

// The LDS limits of every RHS variant, raised once before any launch (and so
// outside any graph capture).
int prepare_rhs_kernels() {
  static int rc = -1;
  if (rc >= 0) return rc;
  size_t lds = 0;
  rc = OA_OK;
  if (rc == OA_OK) rc = rhs_lds<false, false, false>(false, &lds);
  if (rc == OA_OK) rc = rhs_lds<false, false, true>(false, &lds);
  if (rc == OA_OK) rc = rhs_lds<false, true, false>(false, &lds);
  if (rc == OA_OK) rc = rhs_lds<false, true, true>(false, &lds);
  if (rc == OA_OK) rc = rhs_lds<true, false, false>(false, &lds);
  if (rc == OA_OK) rc = rhs_lds<true, false, true>(false, &lds);
  if (rc == OA_OK) rc = rhs_lds<true, true, false>(false, &lds);
  if (rc == OA_OK) rc = rhs_lds<true, true, true>(false, &lds);
  return rc;
}

#define RHS_DISPATCH(FN, ...)                                   \
  switch ((shift ? 4 : 0) + (scalar ? 2 : 0) + (phimode ? 1 : 0)) { \
    case 0: return FN<false, false, false>(__VA_ARGS__);        \
    case 1: return FN<false, false, true>(__VA_ARGS__);         \
    case 2: return FN<false, true, false>(__VA_ARGS__);         \
    case 3: return FN<false, true, true>(__VA_ARGS__);          \
    case 4: return FN<true, false, false>(__VA_ARGS__);         \
    case 5: return FN<true, false, true>(__VA_ARGS__);          \
    case 6: return FN<true, true, false>(__VA_ARGS__);          \
    default: return FN<true, true, true>(__VA_ARGS__);          \
  }

int launch_rhs_table(const RhsItem* items, int n, int blocks, bool shift, bool scalar, bool phimode,
                     bool any_dtb_pre, bool xcd_remap, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  RHS_DISPATCH(launch_rhs_table_t, items, n, blocks, any_dtb_pre, xcd_remap, st)
}

#define SC_DISPATCH(K, GRID, ...)                                                                  \
  do {                                                                                             \
    if (shift && scalar) hipLaunchKernelGGL((K<true, true>), GRID, dim3(kPtBlock), 0, st, __VA_ARGS__);   \
    else if (shift) hipLaunchKernelGGL((K<true, false>), GRID, dim3(kPtBlock), 0, st, __VA_ARGS__);       \
    else if (scalar) hipLaunchKernelGGL((K<false, true>), GRID, dim3(kPtBlock), 0, st, __VA_ARGS__);      \
    else hipLaunchKernelGGL((K<false, false>), GRID, dim3(kPtBlock), 0, st, __VA_ARGS__);                 \
  } while (0)

int launch_project_table(const ProjItem* items, int n, int blocks, bool shift, bool scalar, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  SC_DISPATCH(k_project_table, dim3(blocks), items, n);
  return check_launch();
}

int launch_finebound_table(const FbItem* items, int n, int blocks, bool shift, bool scalar, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  SC_DISPATCH(k_finebound_table, dim3(blocks), items, n);
  return check_launch();
}

int launch_deldef_table(const DelItem* items, int n, int blocks, void* stream) {
  hipLaunchKernelGGL(k_deldef_table, dim3(blocks), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), items, n);
  return check_launch();
}

int launch_saveold_table(const SaveItem* items, int n, int blocks, void* stream) {
  hipLaunchKernelGGL(k_saveold_table, dim3(blocks), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), items, n);
  return check_launch();
}

int launch_restrict_table(const RsItem* items, int n, int blocks, void* stream) {
  hipLaunchKernelGGL(k_restrict_table, dim3(blocks), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), items,
                     n);
  return check_launch();
}

int launch_dtbeta_init(const KGrid& G, double csi, bool scalar, double* y, void* stream) {
  (void)scalar;
  const long n = G.plane;
  const long pl = G.plane;
  hipLaunchKernelGGL(k_dtbeta_init, dim3(blocks_for(n)), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), n,
                     csi, y + S_Delta_r * pl, y + S_Delta_z * pl, y + shift_slot(SH_dtbeta_r) * pl,
                     y + shift_slot(SH_dtbeta_z) * pl);
  return check_launch();
}

int norms_blocks() { return kNormBlocks; }

int launch_norms_partial(const KGrid& G, const double* f, double* part, void* stream) {
  hipLaunchKernelGGL(k_norms_partial, dim3(kNormBlocks), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), G, f,
                     part);
  return check_launch();
}

// Partial records of oa_norms0d_noalloc: kNormsMaxFields fields of
// kNormBlocks records of 6 doubles (384 KiB).  Each GPU holds its own copy for
// the life of the process.  Host code never uses this name as an address: it
// uses the device address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[kNormsMaxFields * kNormBlocks * 6];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the address of `scratch` on device dev, looked up once per device with that
// device current
int scratch_on(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    oa_set_error("oa_norms0d_noalloc: device index out of range");
    return OA_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) {
      oa_set_error("oa_norms0d_noalloc: hipGetDevice failed");
      return OA_ERR_HIP;
    }
    if (prev != dev && hipSetDevice(dev) != hipSuccess) {
      oa_set_error("oa_norms0d_noalloc: hipSetDevice failed");
      return OA_ERR_HIP;
    }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != dev) (void)hipSetDevice(prev);
    if (e != hipSuccess || !ptr) {
      oa_set_error(std::string("oa_norms0d_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return OA_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return OA_OK;
}
}  // namespace

int norms_max_fields() { return kNormsMaxFields; }

int launch_norms_fields_noalloc(int dev, const KGrid& G, int n, const double* const* planes, double* out,
                                void* stream) {
  if (n < 1 || n > kNormsMaxFields) {
    oa_set_error("oa_norms0d_noalloc: " + std::to_string(n) + " arrays requested; one call takes 1 to " +
                 std::to_string(kNormsMaxFields));
    return OA_ERR_ARG;
  }
  double* part = nullptr;
  RC_K(scratch_on(dev, &part));
  NormsFields a{};
  a.G = G;
  a.n = n;
  for (int k = 0; k < n; ++k) a.f[k] = planes[k];
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_norms_fields, dim3(kNormBlocks, static_cast<unsigned>(n)), dim3(kPtBlock), 0, st, a, part);
  RC_K(check_launch());
  hipLaunchKernelGGL(k_norms_combine, dim3(1), dim3(64), 0, st, n, part, out);
  return check_launch();
}

int launch_constraints(const KGrid& G, const KPhys& P, const KStep& S, bool phimode, bool scalar, int sc0,
                       const double* y, double* gaux, double* out, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  RhsItem it{};
  RC_K(rhs_tiles(G, &it.nbi, &it.nbj));
  it.G = G;
  it.P = P;
  it.S = S;
  it.S.dtb_pre = nullptr;
  it.yin = y;
  it.gaux = gaux;
  const size_t lds = sizeof(double) * kTileSize * NBASE;
  static bool attr = false;
  if (!attr) {
    RC_K(raise_lds_limit(&k_aux<false>, lds));
    RC_K(raise_lds_limit(&k_aux<true>, lds));
    attr = true;
  }
  const dim3 tg(static_cast<unsigned>(it.nbj), static_cast<unsigned>(it.nbi), 1);
  if (phimode) hipLaunchKernelGGL(k_aux<true>, tg, dim3(kTile, kTile, 1), lds, st, it);
  else hipLaunchKernelGGL(k_aux<false>, tg, dim3(kTile, kTile, 1), lds, st, it);
  RC_K(check_launch());
  const ConsArgs c{G, P, sc0, y, gaux, out};
  const dim3 blk(64, 4, 1);
  const dim3 grd((G.Nz - G.jlo + 1 + 63) / 64, (G.Nr + 3) / 4, 1);
  const int code = (phimode ? 2 : 0) + (scalar ? 1 : 0);
  if (code == 0) hipLaunchKernelGGL((k_constraints<false, false>), grd, blk, 0, st, c);
  else if (code == 1) hipLaunchKernelGGL((k_constraints<false, true>), grd, blk, 0, st, c);
  else if (code == 2) hipLaunchKernelGGL((k_constraints<true, false>), grd, blk, 0, st, c);
  else hipLaunchKernelGGL((k_constraints<true, true>), grd, blk, 0, st, c);
  RC_K(check_launch());
  hipLaunchKernelGGL(k_ham_fix, dim3(1), dim3(256), 0, st, G, out + CN_ham * G.plane);
  return check_launch();
}

long fb_points(const KGrid& Gf, int depth) { return fb_count(Gf, depth); }
size_t fb_record_bytes() { return sizeof(FbPoint); }

int launch_fb_table(const KGrid& Gf, const KGrid& Gc, int depth, int interporder, FbPoint* table, long n,
                    void* stream) {
  hipLaunchKernelGGL(k_fb_table, dim3(blocks_for(n)), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), Gf, Gc,
                     depth, interporder, n, table);
  return check_launch();
}

}  // namespace oax
