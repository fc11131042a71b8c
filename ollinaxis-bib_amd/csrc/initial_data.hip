// Closed-form initial data computed on the device, one thread per point, in
// the level's slab.  It follows initial.f90:53-150 (Minkowski defaults), :189-225
// (Schwarzschild), :284-330 (Brill-Lindquist), :528-580 (factorhdet),
// :585-620 (trace removal), :627-745 (precollapsed lapse, gauss lapse and
// shift perturbations), idata_brillwave.f90:88-105,183-208 (Brill/Holz metric)
// and idata_scalarpulse.f90:60-103 (scalar pulse).  For Brill waves and the
// scalar pulse psi comes from the WaveElliptic solve of the Hamiltonian
// constraint (elliptic.hip), which needs the fields of a first pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "oa_internal.h"
#include "oaxis.h"
#define OA_HD __device__ __forceinline__
#define OA_HHD __host__ __device__ inline
#include "oaxis_device.h"

namespace {

__device__ double pulse_shell(double a0, double r0, double z0, double sr, double sz, double r, double z, double rr) {
  if (r0 == 0.0 && z0 == 0.0) return a0 * std::exp(-(rr * rr * rr * rr / (r * r * sr * sr + z * z * sz * sz)));
  const double c = (r * r * r0 + z * z * z0) / (rr * rr);
  const double w = rr * rr / (r * r * sr * sr + z * z * sz * sz);
  return a0 * (std::exp(-(rr - c) * (rr - c) * w) + std::exp(-(rr + c) * (rr + c) * w));
}

__device__ double torus(double a0, double r0, double z0, double sr, double sz, double r, double z, bool eqsym) {
  auto e = [](double x) { return std::exp(-x * x); };
  if (r0 == 0.0 && z0 == 0.0) return a0 * e(r / sr) * e(z / sz);
  if (eqsym) return a0 * (e((r - r0) / sr) + e((r + r0) / sr)) * (e((z - z0) / sz) + e((z + z0) / sz));
  return a0 * e((z - z0) / sz) * (e((r - r0) / sr) + e((r + r0) / sr));
}

}  // namespace

// the opaque handle is defined in oaxis_host.cpp; initial data only needs
// these accessors and the elliptic solve
namespace oax {
int handle_level_slab(oa_handle* h, int level, double** slab, KGrid* G, int* nf, bool* shift, bool* scalar,
                      const oa_params** p);
void handle_mark_dirty(oa_handle* h, int level);
void* handle_stream(oa_handle* h);
int handle_elliptic_exit(oa_handle* h);
void handle_elliptic_record(oa_handle* h, int steps0, int steps1, double residual);
struct EllSolver;
int ell_solve_begin(oa_handle* h, int stationary_exit, EllSolver** out);
int ell_solve_run(EllSolver* s, int* steps0, int* steps1, double* gres);
const double* ell_solution(const EllSolver* s, int level);
void ell_solve_end(EllSolver* s);
}  // namespace oax

namespace {

// One point (i, j) of the level: every slot of the state at that point.
// psi: the solved conformal factor of Brill-wave / scalar-pulse data on this
// level (null in the first pass, which only needs the scalar field and xi).
__global__ void __launch_bounds__(256) k_initial_data(oa_params p, oax::KGrid G, bool shift, bool scalar,
                                                      const double* __restrict__ psi, double* __restrict__ y) {
  using namespace oax;
  const int g = G.g;
  const int j = 1 - g + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = 1 - g + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
  if (j > G.Nz || i > G.Nr) return;
  auto at = [&](int s) -> double& { return y[static_cast<long>(s) * G.plane + IDX(G, i, j)]; };
  const double r = (static_cast<double>(i) - 0.5) * G.dr;
  const double z = (static_cast<double>(j) - G.zoff) * G.dz;
  const double rr = std::sqrt(r * r + z * z);
  // Minkowski defaults
  at(S_alpha) = 1.0;
  at(S_chi) = 1.0;
  at(S_phi) = 0.0;
  at(S_A) = 1.0;
  at(S_B) = 1.0;
  at(S_H) = 1.0;
  double ps = 1.0;
  if (p.idata == OA_IDATA_SCHWARZSCHILD) {
    ps = 1.0 + 0.5 * p.BH1mass / std::fabs(rr);
    if (p.ilapse == OA_ILAPSE_ISOTROPIC)
      at(S_alpha) = (1.0 - 0.5 * p.BH1mass / std::fabs(rr)) / (1.0 + 0.5 * p.BH1mass / std::fabs(rr));
  } else if (p.idata == OA_IDATA_BRILLLINDQUIST) {
    auto d = [&](double zc) { return std::sqrt(r * r + (z - zc) * (z - zc)); };
    auto dm = [&](double zc) { return std::sqrt(r * r + (z + zc) * (z + zc)); };
    if (p.N_BH == 1) {
      ps = 1.0 + 0.5 * p.BH1mass / d(p.BH1Z);
      if (p.eqsym && p.duplicate_bhs) ps += 0.5 * p.BH1mass / dm(p.BH1Z);
    } else {
      ps = 1.0 + 0.5 * (p.BH1mass / d(p.BH1Z) + p.BH2mass / d(p.BH2Z));
      if (p.eqsym && p.duplicate_bhs) ps += 0.5 * (p.BH1mass / dm(p.BH1Z) + p.BH2mass / dm(p.BH2Z));
    }
  }
  if (p.idata == OA_IDATA_SCHWARZSCHILD || p.idata == OA_IDATA_BRILLLINDQUIST) {
    at(S_phi) = std::log(ps);
    at(S_chi) = 1.0 / std::pow(ps, static_cast<double>(p.chipower));
  }
  if ((p.idata == OA_IDATA_BRILLWAVE || p.idata == OA_IDATA_SCALARPULSE) && psi) {
    // idata_brillwave.f90:183-197, idata_scalarpulse.f90:185-206
    const double sp = psi[IDX(G, i, j)];
    at(S_phi) = std::log(sp);
    at(S_chi) = 1.0 / std::pow(sp, static_cast<double>(p.chipower));
  }
  if (p.idata == OA_IDATA_BRILLWAVE) {
    // idata_brillwave.f90:88-105,198-208
    const double q = p.brill_a0 * r * r *
                     std::exp(-(r / p.brill_sr0) * (r / p.brill_sr0) - (z / p.brill_sz0) * (z / p.brill_sz0));
    at(S_A) = std::exp(2.0 * q);
    at(S_B) = at(S_A);
    at(S_Delta_r) = (1.0 - 1.0 / at(S_A)) / r;
  }
  // factorhdet (initial.f90:528-580)
  {
    const double rc = r * at(S_C);
    const double hdet = (at(S_A) * at(S_B) - rc * rc) * at(S_H);
    at(S_phi) = at(S_phi) + std::log(hdet) / 12.0;
    const double aux = std::pow(hdet, 1.0 / 3.0);
    at(S_A) /= aux;
    at(S_B) /= aux;
    at(S_C) /= aux;
    at(S_H) /= aux;
  }
  ps = std::exp(at(S_phi));
  if (p.ilapse == OA_ILAPSE_PSIMINUS2) at(S_alpha) = 1.0 / (ps * ps);
  if (p.ilapse == OA_ILAPSE_PSIMINUS4) at(S_alpha) = 1.0 / (ps * ps * ps * ps);
  if (p.lapsepert == OA_PERT_GAUSS) {
    if (p.ilapse == OA_ILAPSE_SHELL)
      at(S_alpha) += pulse_shell(p.lapse_a0, p.lapse_r0, p.lapse_z0, p.lapse_sr0, p.lapse_sz0, r, z, rr);
    else if (p.ilapse == OA_ILAPSE_TORUS)
      at(S_alpha) += torus(p.lapse_a0, p.lapse_r0, p.lapse_z0, p.lapse_sr0, p.lapse_sz0, r, z, p.eqsym);
  }
  if (shift && p.shiftpert == OA_PERT_GAUSS && p.ishift == OA_ISHIFT_SHELL) {
    double prof;
    const double sr = p.shift_sr0, sz = p.shift_sz0;
    if (p.shift_r0 == 0.0 && p.shift_z0 == 0.0) {
      prof = p.shift_a0 * std::exp(-(rr * rr * rr * rr / (r * r * sr * sr + z * z * sz * sz)));
    } else {
      const double c = (r * r * p.shift_r0 + z * z * p.shift_z0) / (rr * rr);
      const double w = rr * rr / (r * r * sr * sr + z * z * sz * sz);
      prof = p.shift_a0 * (std::exp(-(rr - c) * (rr - c) * w) - std::exp(-(rr + c) * (rr + c) * w));
    }
    at(shift_slot(SH_beta_r)) += (r / rr) * prof;
    at(shift_slot(SH_beta_z)) += (z / rr) * prof;
  }
  if (scalar && p.idata == OA_IDATA_SCALARPULSE) {
    double v;
    if (p.scalarpulse_type == OA_PULSE_SHELL) {
      if (p.scalar_r0 == 0.0 && p.scalar_z0 == 0.0)
        v = p.scalar_a0 * std::exp(-(r * r / (p.scalar_sr0 * p.scalar_sr0) + z * z / (p.scalar_sz0 * p.scalar_sz0)));
      else
        v = pulse_shell(p.scalar_a0, p.scalar_r0, p.scalar_z0, p.scalar_sr0, p.scalar_sz0, r, z, rr);
    } else if (p.scalarpulse_type == OA_PULSE_TORUS) {
      v = torus(p.scalar_a0, p.scalar_r0, p.scalar_z0, p.scalar_sr0, p.scalar_sz0, r, z, p.eqsym);
    } else {
      v = p.scalar_a0 * std::exp(-(r * r + (1.0 - p.scalar_sz0) * z * z));
    }
    at(scalar_slot(shift, SC_phi)) = v;
  }
}

// xi = grad(phi) of the scalar pulse with the 4th-order stencils and
// closures of diff1r/diff1z: xi_r for i in [1, Nr] with its axis images,
// xi_z for every i and j, replaced by its equator images with eqsym.
__global__ void __launch_bounds__(256) k_initial_xi(oax::KGrid G, int eqsym, int sp, int sxr, int sxz,
                                                    double* __restrict__ y) {
  using namespace oax;
  const int g = G.g;
  const int j = 1 - g + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = 1 - g + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
  if (j > G.Nz || i > G.Nr) return;
  const double* ph = y + static_cast<long>(sp) * G.plane;
  double* xr = y + static_cast<long>(sxr) * G.plane;
  double* xz = y + static_cast<long>(sxz) * G.plane;
  const double idr = 1.0 / G.dr, idz = 1.0 / G.dz;
  if (i >= 1) {
    auto F = [&](int ii) { return ph[IDX(G, ii, j)]; };
    double d;
    if (i <= G.Nr - 2) d = idr * (8.0 * (F(i + 1) - F(i - 1)) - (F(i + 2) - F(i - 2))) / 12.0;
    else if (i == G.Nr - 1) d = idr * (3.0 * F(i + 1) + 10.0 * F(i) - 18.0 * F(i - 1) + 6.0 * F(i - 2) - F(i - 3)) / 12.0;
    else d = idr * (25.0 * F(i) - 48.0 * F(i - 1) + 36.0 * F(i - 2) - 16.0 * F(i - 3) + 3.0 * F(i - 4)) / 12.0;
    xr[IDX(G, i, j)] = d;
    if (i <= g) xr[IDX(G, 1 - i, j)] = -d;
  }
  if (!(eqsym && j <= 0)) {
    auto F = [&](int jj) { return ph[IDX(G, i, jj)]; };
    double d;
    if (j >= 3 - g && j <= G.Nz - 2) d = idz * (8.0 * (F(j + 1) - F(j - 1)) - (F(j + 2) - F(j - 2))) / 12.0;
    else if (j == 2 - g) d = -idz * (3.0 * F(j - 1) + 10.0 * F(j) - 18.0 * F(j + 1) + 6.0 * F(j + 2) - F(j + 3)) / 12.0;
    else if (j == 1 - g) d = -idz * (25.0 * F(j) - 48.0 * F(j + 1) + 36.0 * F(j + 2) - 16.0 * F(j + 3) + 3.0 * F(j + 4)) / 12.0;
    else if (j == G.Nz - 1) d = idz * (3.0 * F(j + 1) + 10.0 * F(j) - 18.0 * F(j - 1) + 6.0 * F(j - 2) - F(j - 3)) / 12.0;
    else d = idz * (25.0 * F(j) - 48.0 * F(j - 1) + 36.0 * F(j - 2) - 16.0 * F(j - 3) + 3.0 * F(j - 4)) / 12.0;
    xz[IDX(G, i, j)] = d;
    if (eqsym && j >= 1 && j <= g) xz[IDX(G, i, 1 - j)] = -d;
  }
}

}  // namespace

namespace {

// every closed-form field of every level; psi[l] the solved conformal
// factor of level l, or null
int fill_levels(oa_handle* h, int nlev, const double* const* psi) {
  using namespace oax;
  hipStream_t st = static_cast<hipStream_t>(handle_stream(h));
  for (int l = 0; l < nlev; ++l) {
    double* slab = nullptr;
    KGrid G{};
    int nf = 0;
    bool shift = false, scalar = false;
    const oa_params* pp = nullptr;
    handle_level_slab(h, l, &slab, &G, &nf, &shift, &scalar, &pp);
    const oa_params& p = *pp;
    const int g = G.g;
    // the slots no formula sets are zero
    if (hipMemsetAsync(slab, 0, sizeof(double) * static_cast<size_t>(nf) * G.plane, st) != hipSuccess) {
      oa_set_error("oa_initial_data: clearing the state failed");
      return OA_ERR_HIP;
    }
    const dim3 blk(64, 4, 1);
    const dim3 grd((G.Nz + g + 63) / 64, (G.Nr + g + 3) / 4, 1);
    hipLaunchKernelGGL(k_initial_data, grd, blk, 0, st, p, G, shift, scalar, psi ? psi[l] : nullptr, slab);
    if (scalar && p.idata == OA_IDATA_SCALARPULSE)
      hipLaunchKernelGGL(k_initial_xi, grd, blk, 0, st, G, p.eqsym, scalar_slot(shift, SC_phi),
                         scalar_slot(shift, SC_xi_r), scalar_slot(shift, SC_xi_z), slab);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
      oa_set_error("oa_initial_data: initial-data kernel failed");
      return OA_ERR_HIP;
    }
    handle_mark_dirty(h, l);
  }
  return OA_OK;
}

}  // namespace

extern "C" int oa_initial_data(oa_handle* h) {
  using namespace oax;
  int nlev = 0;
  if (!h || oa_num_levels(h, &nlev) != OA_OK) return OA_ERR_ARG;
  const oa_params* pp = nullptr;
  {
    double* slab = nullptr;
    KGrid G{};
    int nf = 0;
    bool shift = false, scalar = false;
    handle_level_slab(h, 0, &slab, &G, &nf, &shift, &scalar, &pp);
  }
  // initial.f90:157-161: the state comes from the checkpoint directory
  if (pp->idata == OA_IDATA_CHECKPOINT) {
    oa_set_error("oa_initial_data: idata = checkpoint; read the state with oa_checkpoint_restart(h, checkpointfile)");
    return OA_ERR_STATE;
  }
  int rc = fill_levels(h, nlev, nullptr);
  if (rc != OA_OK) return rc;
  if (pp->idata != OA_IDATA_BRILLWAVE && pp->idata != OA_IDATA_SCALARPULSE) return OA_OK;
  // Brill waves and the scalar pulse: solve the Hamiltonian constraint for
  // psi on every level (idata_brillwave.f90:117-135,
  // idata_scalarpulse.f90:131-178), then fill the fields again with it
  EllSolver* es = nullptr;
  rc = ell_solve_begin(h, handle_elliptic_exit(h), &es);
  if (rc != OA_OK) return rc;
  int steps0 = 0, steps1 = 0;
  double res = 0.0;
  rc = ell_solve_run(es, &steps0, &steps1, &res);
  if (rc == OA_OK) {
    handle_elliptic_record(h, steps0, steps1, res);
    std::vector<const double*> psi(static_cast<size_t>(nlev));
    for (int l = 0; l < nlev; ++l) psi[static_cast<size_t>(l)] = ell_solution(es, l);
    rc = fill_levels(h, nlev, psi.data());
  }
  ell_solve_end(es);
  return rc;
}
