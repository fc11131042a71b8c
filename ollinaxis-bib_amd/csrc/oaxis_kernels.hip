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

#include "oa_internal.h"

#define OA_HD __device__ __forceinline__
#define OA_HHD __host__ __device__ inline
#include "oaxis_device.h"

namespace oax {

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

// Stage one RHS input tile: nst planes, plane f from src(f), into LDS.
template <class Src>
__device__ __forceinline__ void stage_tile(const KGrid& G, double* tile, int nst, Src src, int ib, int jb) {
  const int tid = static_cast<int>(threadIdx.y) * kTile + static_cast<int>(threadIdx.x);
  const int lo = 1 - G.g;
  for (int f = 0; f < nst; ++f) {
    const double* __restrict__ p = src(f);
    double* dst = tile + f * kTileSize;
    for (int q = tid; q < kTileSize; q += kTile * kTile) {
      const int ii = ib + q / kTileW, jj = jb + q % kTileW;
      if (ii >= lo && ii <= G.Nr && jj >= lo && jj <= G.Nz) dst[q] = p[IDX(G, ii, jj)];
    }
  }
  __syncthreads();
}

// One RK substep of a level: each 16 x 16 block stages its tile and runs two
// phases per point.  Phase 1 forms the geometric auxiliaries (D_i D_j alpha,
// the Ricci tensor, div KT: the longest expressions, which need only the
// vacuum fields) and passes them through the per-level buffer gaux; phase 2
// forms every source, the dissipation, the radiative boundaries and the RK
// update.  Kept apart, the phases need at most the registers of the larger
// one, so the kernel runs at two waves per SIMD without spilling.
template <bool SHIFT, bool SCALAR, bool PHIMODE>
__global__ void __launch_bounds__(kTile * kTile, 2) k_rhs(KGrid G, KPhys P, KStep S, const double* __restrict__ yin,
                                                          const double* __restrict__ yp, double* __restrict__ acc,
                                                          double* __restrict__ yout, double* __restrict__ gaux) {
  extern __shared__ double tile[];
  constexpr int nf = nslots(SHIFT, SCALAR);
  const int nbj = static_cast<int>(gridDim.x), nbi = static_cast<int>(gridDim.y);
  const int bj = static_cast<int>(blockIdx.x), bi = static_cast<int>(blockIdx.y);
  const int j0 = tile_start(bj, nbj, G.jlo, G.Nz);
  const int i0 = tile_start(bi, nbi, 1, G.Nr);
  const int ib = i0 - kHalo, jb = j0 - kHalo;
  // the evolved fields, then (first substep only) the dtbeta planes it differentiates
  const int nst = nf + (S.dtb_pre ? 2 : 0);
  stage_tile(G, tile, nst, [&](int f) { return (f < nf) ? yin + f * G.plane : S.dtb_pre + (f - nf) * G.plane; },
             ib, jb);
  const int i = i0 + static_cast<int>(threadIdx.y), j = j0 + static_cast<int>(threadIdx.x);
  if (bi == nbi - 1 && i < 1 + (nbi - 1) * kTile) return;        // done by the tile before
  if (bj == nbj - 1 && j < G.jlo + (nbj - 1) * kTile) return;
  const TileView V{tile, ib, jb, kTileW, kTileSize, nf};
  // phase 1: the geometric auxiliaries of this point, through gaux (own point
  // only, so no synchronisation); the memory barrier keeps the compiler from
  // carrying phase-1 values or loads into phase 2, so the two phases do not
  // add up in registers
  if (interior_tile(G, i0, j0))
    rhs_point<false, false, PHIMODE, true, 1>(G, P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
  else
    rhs_point<false, false, PHIMODE, false, 1>(G, P, S, V, nullptr, nullptr, nullptr, gaux, i, j);
  asm volatile("" ::: "memory");
  rhs_point<SHIFT, SCALAR, PHIMODE, false, 2>(G, P, S, V, yp, acc, yout, gaux, i, j);
}

// ------------------------------------------------------- projection pass
template <bool SHIFT, bool SCALAR>
__global__ void __launch_bounds__(256) k_project(KGrid G, KPhys P, int time0, double* __restrict__ y) {
  const int j = G.jlo + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = 1 + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
  if (j > G.Nz || i > G.Nr) return;
  project_point<SHIFT, SCALAR>(G, P, time0, y, i, j);
}

__global__ void __launch_bounds__(256) k_deldef(KGrid G, double* __restrict__ y) {
  const int j = G.jlo + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = 1 + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
  if (j > G.Nz || i > G.Nr) return;
  deldef_point(G, y, i, j);
}

// dtbeta = drivercsi * Delta (evolve.f90:174-195), whole arrays incl. images
__global__ void k_dtbeta_init(long n, double csi, const double* __restrict__ dr, const double* __restrict__ dz,
                              double* __restrict__ tr, double* __restrict__ tz) {
  const long q = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  tr[q] = csi * dr[q];
  tz[q] = csi * dz[q];
}

// saveold for the fine-boundary strips (one thread per strip element)
__global__ void k_saveold(KGrid G, int nf, const double* __restrict__ y, double* __restrict__ rR,
                          double* __restrict__ zL, double* __restrict__ zR) {
  const int q = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int f = blockIdx.y;
  if (f >= nf) return;
  saveold_point(G, y, rR, zL, zR, f, q);
}

// finebound(box 0, level, dtw, .true.): one thread per fine strip point
template <bool SHIFT, bool SCALAR>
__global__ void __launch_bounds__(256) k_finebound(KGrid Gf, KGrid Gc, KPhys P, KTime T, int time0,
                                                   const FbPoint* __restrict__ table, long n,
                                                   const double* __restrict__ yc, double* __restrict__ yf,
                                                   const double* __restrict__ rR0, const double* __restrict__ rR1,
                                                   const double* __restrict__ zL0, const double* __restrict__ zL1,
                                                   const double* __restrict__ zR0, const double* __restrict__ zR1) {
  const long q = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  finebound_point<SHIFT, SCALAR>(Gf, Gc, P, T, time0, yc, yf, rR0, rR1, zL0, zL1, zR0, zR1, table[q],
                                 static_cast<int>(blockIdx.y));
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

// restrict(box 0, level, .true.): one thread per (coarse point, field)
__global__ void __launch_bounds__(256) k_restrict(KGrid Gf, KGrid Gc, int interporder,
                                                  const double* __restrict__ yf, double* __restrict__ yc) {
  const int jc = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + 1 - Gc.g;
  const int ic = static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y) + 1 - Gc.g;
  if (jc > Gc.Nz || ic > Gc.Nr) return;
  restrict_point(Gf, Gc, static_cast<int>(blockIdx.z), interporder, yf, yc, ic, jc);
}

// ------------------------------------------------------------- launchers
static dim3 level_grid(const KGrid& G, dim3 blk) {
  const int nj = G.Nz - G.jlo + 1, ni = G.Nr;
  return dim3((nj + blk.x - 1) / blk.x, (ni + blk.y - 1) / blk.y, 1);
}

static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    oa_set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return OA_ERR_HIP;
  }
  return OA_OK;
}

template <class K>
static int raise_lds_limit(K* kernel, size_t bytes) {
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          static_cast<int>(bytes)) != hipSuccess) {
    oa_set_error("hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed for an RHS kernel");
    return OA_ERR_HIP;
  }
  return OA_OK;
}

template <bool SH, bool SC, bool PM>
static int launch_rhs_t(const KGrid& G, const KPhys& P, const KStep& S, const double* yin, const double* yp,
                        double* acc, double* yout, double* gaux, hipStream_t st) {
  // the dynamic LDS of one block exceeds the 64 KiB default; two more
  // planes on a level's first substep (KStep::dtb_pre)
  const size_t per = sizeof(double) * kTileSize;
  const size_t lds = per * static_cast<size_t>(nslots(SH, SC) + (S.dtb_pre ? 2 : 0));
  static bool attr_set = false;
  if (!attr_set) {
    const int rc = raise_lds_limit(&k_rhs<SH, SC, PM>, per * static_cast<size_t>(nslots(SH, SC) + 2));
    if (rc != OA_OK) return rc;
    attr_set = true;
  }
  const dim3 blk(kTile, kTile, 1);
  const int nj = G.Nz - G.jlo + 1;
  const dim3 grd(static_cast<unsigned>((nj + kTile - 1) / kTile), static_cast<unsigned>((G.Nr + kTile - 1) / kTile), 1);
  hipLaunchKernelGGL((k_rhs<SH, SC, PM>), grd, blk, lds, st, G, P, S, yin, yp, acc, yout, gaux);
  return check_launch();
}

int launch_rhs(const KGrid& G, const KPhys& P, const KStep& S, bool shift, bool scalar, bool phimode,
               const double* yin, const double* yp, double* acc, double* yout, double* gaux, void* stream) {
  if (G.Nr < kTile || G.Nz - G.jlo + 1 < kTile) {
    oa_set_error("launch_rhs: a level needs at least 16 points in each direction");
    return OA_ERR_ARG;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int code = (shift ? 4 : 0) + (scalar ? 2 : 0) + (phimode ? 1 : 0);
  switch (code) {
    case 0: return launch_rhs_t<false, false, false>(G, P, S, yin, yp, acc, yout, gaux, st);
    case 1: return launch_rhs_t<false, false, true>(G, P, S, yin, yp, acc, yout, gaux, st);
    case 2: return launch_rhs_t<false, true, false>(G, P, S, yin, yp, acc, yout, gaux, st);
    case 3: return launch_rhs_t<false, true, true>(G, P, S, yin, yp, acc, yout, gaux, st);
    case 4: return launch_rhs_t<true, false, false>(G, P, S, yin, yp, acc, yout, gaux, st);
    case 5: return launch_rhs_t<true, false, true>(G, P, S, yin, yp, acc, yout, gaux, st);
    case 6: return launch_rhs_t<true, true, false>(G, P, S, yin, yp, acc, yout, gaux, st);
    default: return launch_rhs_t<true, true, true>(G, P, S, yin, yp, acc, yout, gaux, st);
  }
}

int launch_project(const KGrid& G, const KPhys& P, bool shift, bool scalar, bool time0, double* y, void* stream) {
  const dim3 blk(64, 4, 1);
  const dim3 grd = level_grid(G, blk);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int t0 = time0 ? 1 : 0;
  if (shift && scalar) hipLaunchKernelGGL((k_project<true, true>), grd, blk, 0, st, G, P, t0, y);
  else if (shift) hipLaunchKernelGGL((k_project<true, false>), grd, blk, 0, st, G, P, t0, y);
  else if (scalar) hipLaunchKernelGGL((k_project<false, true>), grd, blk, 0, st, G, P, t0, y);
  else hipLaunchKernelGGL((k_project<false, false>), grd, blk, 0, st, G, P, t0, y);
  return check_launch();
}

int launch_deldef(const KGrid& G, double* y, void* stream) {
  const dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_deldef, level_grid(G, blk), blk, 0, static_cast<hipStream_t>(stream), G, y);
  return check_launch();
}

int launch_dtbeta_init(const KGrid& G, double csi, bool scalar, double* y, void* stream) {
  (void)scalar;
  const long n = G.plane;
  const int blk = 256;
  const long pl = G.plane;
  hipLaunchKernelGGL(k_dtbeta_init, dim3(static_cast<unsigned>((n + blk - 1) / blk)), dim3(blk), 0,
                     static_cast<hipStream_t>(stream), n, csi, y + S_Delta_r * pl, y + S_Delta_z * pl,
                     y + shift_slot(SH_dtbeta_r) * pl, y + shift_slot(SH_dtbeta_z) * pl);
  return check_launch();
}

int launch_saveold(const KGrid& G, int nf, const double* y, double* rR, double* zL, double* zR, void* stream) {
  const int nrt = G.Nr + G.g, nzt = G.Nz + G.g;
  const int per = G.g * (nrt > nzt ? nrt : nzt);
  const int blk = 256;
  hipLaunchKernelGGL(k_saveold, dim3((per + blk - 1) / blk, nf), dim3(blk), 0, static_cast<hipStream_t>(stream),
                     G, nf, y, rR, zL, zR);
  return check_launch();
}

long fb_points(const KGrid& Gf, int depth) { return fb_count(Gf, depth); }
size_t fb_record_bytes() { return sizeof(FbPoint); }

int launch_fb_table(const KGrid& Gf, const KGrid& Gc, int depth, int interporder, FbPoint* table, long n,
                    void* stream) {
  hipLaunchKernelGGL(k_fb_table, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), Gf, Gc, depth, interporder, n, table);
  return check_launch();
}

int launch_finebound(const KGrid& Gf, const KGrid& Gc, const KPhys& P, const KTime& T, bool shift, bool scalar,
                     bool time0, const FbPoint* table, long n, const double* yc, double* yf, const double* rR0,
                     const double* rR1, const double* zL0, const double* zL1, const double* zR0, const double* zR1,
                     void* stream) {
  const int blk = 256;
  const dim3 grd(static_cast<unsigned>((n + blk - 1) / blk), static_cast<unsigned>(finebound_groups(shift, scalar)), 1);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int t0 = time0 ? 1 : 0;
#define LAUNCH_FB(SH, SC) \
  hipLaunchKernelGGL((k_finebound<SH, SC>), grd, dim3(blk), 0, st, Gf, Gc, P, T, t0, table, n, yc, yf, rR0, rR1, \
                     zL0, zL1, zR0, zR1)
  if (shift && scalar) LAUNCH_FB(true, true);
  else if (shift) LAUNCH_FB(true, false);
  else if (scalar) LAUNCH_FB(false, true);
  else LAUNCH_FB(false, false);
#undef LAUNCH_FB
  return check_launch();
}

int launch_restrict(const KGrid& Gf, const KGrid& Gc, int nf, int interporder, const double* yf, double* yc,
                    void* stream) {
  const dim3 blk(64, 4, 1);
  const dim3 grd((Gc.Nz + Gc.g + blk.x - 1) / blk.x, (Gc.Nr + Gc.g + blk.y - 1) / blk.y,
                 static_cast<unsigned>(nf));
  hipLaunchKernelGGL(k_restrict, grd, blk, 0, static_cast<hipStream_t>(stream), Gf, Gc, interporder, yf, yc);
  return check_launch();
}

}  // namespace oax
