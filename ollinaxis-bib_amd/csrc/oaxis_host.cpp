// Host runtime of the evolution hot path: device residency of every level of
// box 0, the Berger-Oliger schedule of onestep.f90, host <-> device transfer
// in the reference layout, and the C ABI declared in include/oaxis.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "oa_internal.h"
#include "oaxis.h"

namespace {
thread_local std::string g_last_error;
}

void oa_set_error(const std::string& msg) { g_last_error = msg; }
extern "C" const char* oa_last_error(void) { return g_last_error.c_str(); }

#include "oa_state.h"

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

struct FieldName {
  const char* name;
  int slot;       // device slot, -1 = derived
  int symr, symz;
};

int slot_of(const oa_handle* h, const std::string& name) {
  static const char* base[] = {"alpha", "chi", "phi", "A", "B", "H", "C", "trK", "KTA", "KTB", "KTC",
                               "Delta_r", "Delta_z"};
  for (int s = 0; s < NBASE; ++s)
    if (name == base[s]) return s;
  if (h->shift) {
    static const char* sh[] = {"beta_r", "beta_z", "dtbeta_r", "dtbeta_z"};
    for (int s = 0; s < NSHIFT; ++s)
      if (name == sh[s]) return shift_slot(s);
  }
  if (h->scalar) {
    static const char* sc[] = {"scalar_phi", "scalar_xi_r", "scalar_xi_z", "scalar_pi"};
    for (int s = 0; s < NSCALAR; ++s)
      if (name == sc[s]) return scalar_slot(h->shift, s);
  }
  return -1;
}

// constraint array names (arrays.config) and their planes in Level::cons
int cons_slot(const std::string& name) {
  static const char* names[] = {"ham", "mom_r", "mom_z", "CDelta_r", "CDelta_z"};
  for (int c = 0; c < 5; ++c)
    if (name == names[c]) return c;
  return -1;
}

void fill_phys(oa_handle* h) {
  const oa_params& p = h->p;
  KPhys& P = h->phys;
  P.slicing = p.slicing;
  P.gauge_f = p.gauge_f;
  P.gauge_eta = p.gauge_eta;
  P.geodiss = p.geodiss;
  P.scalardiss = p.scalardiss;
  P.eta = p.eta;
  P.drivercsi = p.drivercsi;
  P.drivereta = p.drivereta;
  P.driveradv = p.driveradv;
  P.shiftadvect = p.shiftadvect;
  P.scalarpotential = p.scalarpotential;
  P.scalarmethod = p.scalarmethod;
  P.scalar_mass = p.scalar_mass;
  P.scalar_lambda = p.scalar_lambda;
  P.chimethod = p.chimethod;
  P.radiative = 0;
}

Level* level_of(oa_handle* h, int box, int level) {
  if (!h || box != 0 || level < 0 || level >= h->nlev) {
    oa_set_error("bad box/level (only box 0 exists; level out of range)");
    return nullptr;
  }
  return &h->L[level];
}

double* strip_ptr(oa_handle* h, Level& L, int tl, int piece) {
  // piece 0: rR, 1: zL, 2: zR
  double* base = L.strip[L.ring[tl]];
  const long nf = h->nf;
  if (piece == 0) return base;
  if (piece == 1) return base + nf * L.strip_r;
  return base + nf * L.strip_r + nf * L.strip_z;
}

}  // namespace

// ------------------------------------------------------------ step pieces
namespace oax {

int niter_of(const oa_handle* h) { return (h->p.integrator == OA_RK4) ? 4 : h->p.icniter; }

// dtw and the RK4 weight of internal iteration iter (onestep.f90:141-177)
void substep_weights(const oa_handle* h, int level, int iter, double* dtw, double* w) {
  const double dt = h->L[level].G.dt;
  *w = 0.0;
  if (h->p.integrator == OA_RK4) {
    switch (iter) {
      case 1: *dtw = 0.5 * dt; *w = 1.0 / 6.0; break;
      case 2: *dtw = 0.5 * dt; *w = 1.0 / 3.0; break;
      case 3: *dtw = dt; *w = 1.0 / 3.0; break;
      default: *dtw = dt; *w = 1.0 / 6.0; break;
    }
  } else {
    *dtw = (iter < h->p.icniter) ? 0.5 * dt : dt;
  }
}

int prepare_saveold(oa_handle* h, int level, SaveItem* it) {
  Level& L = h->L[level];
  // time levels shift: t-2 <- t-1 <- t0 <- new copy
  const int oldest = L.ring[2];
  L.ring[2] = L.ring[1];
  L.ring[1] = L.ring[0];
  L.ring[0] = oldest;
  *it = SaveItem{};
  it->G = L.G;
  it->nf = h->nf;
  it->y = L.slab[L.cur];
  it->rR = strip_ptr(h, L, 0, 0);
  it->zL = strip_ptr(h, L, 0, 1);
  it->zR = strip_ptr(h, L, 0, 2);
  L.step_a = L.cur;
  L.last_out = L.cur;
  L.in_step = true;
  return OA_OK;
}

int prepare_substep(oa_handle* h, int level, int iter, int niter, double dtw, double w, RhsItem* it,
                    bool* phimode) {
  Level& L = h->L[level];
  if (!L.in_step) {
    oa_set_error("oa_substep called before oa_saveold");
    return OA_ERR_STATE;
  }
  if (L.dirty) {
    oa_set_error("oa_substep on a level whose state is not projected (call oa_auxiliary / oa_initial_aux)");
    return OA_ERR_STATE;
  }
  *it = RhsItem{};
  RC(rhs_tiles(L.G, &it->nbi, &it->nbj));
  const int a = L.step_a;
  const int in = (iter == 1) ? a : L.last_out;
  int out;
  if (iter == niter && in != a) {
    out = a;
  } else {
    out = 0;
    while (out == a || out == in) ++out;
  }
  const bool time0 = (L.t == 0.0);
  *phimode = L.aux_phi;
  it->G = L.G;
  it->P = h->phys;
  it->P.radiative = (level == 0 && h->p.boundtype == OA_BOUND_RADIATIVE) ? 1 : 0;
  it->S = KStep{iter, niter, h->p.integrator == OA_RK4 ? 1 : 0, dtw, w, time0 ? 1 : 0,
                L.dtb_pre_valid ? L.dtb_pre : nullptr};
  it->yin = L.slab[in];
  it->yp = L.slab[a];
  it->acc = L.acc;
  it->yout = L.slab[out];
  it->gaux = L.gaux;
  L.aux_phi = time0;
  L.dtb_pre_valid = false;
  L.cons_valid = false;
  L.last_out = out;
  if (!time0 && h->p.chimethod) L.phi_valid = false;
  if (iter == niter) {
    L.cur = out;
    L.in_step = false;
  }
  return OA_OK;
}

bool prepare_finebound(oa_handle* h, int level, double dtw, FbItem* it) {
  if (level == 0) return false;
  Level& L = h->L[level];
  Level& Lc = h->L[level - 1];
  const double tp = Lc.t, t0 = L.t, tm1 = L.t1, tl = t0 + dtw;
  *it = FbItem{};
  KTime& T = it->T;
  T.same = (tp == tl) ? 1 : 0;
  T.depth = T.same ? L.G.g + 2 : L.G.g - 1;
  T.border = (tm1 == 0.0) ? 1 : 2;
  // coefficients exactly as bound_interp.inc forms them
  T.c_aux = (tl - t0) / (tp - t0);
  T.c_p = (tl - tp) / (t0 - tp);
  T.c0 = (tl - t0) * (tl - tm1) / ((tp - t0) * (tp - tm1));
  T.c1 = (tl - tp) * (tl - tm1) / ((t0 - tp) * (t0 - tm1));
  T.c2 = (tl - tp) * (tl - t0) / ((tm1 - tp) * (tm1 - t0));
  const int tab = T.same ? 0 : 1;
  it->Gf = L.G;
  it->Gc = Lc.G;
  it->P = h->phys;
  it->time0 = (L.t == 0.0) ? 1 : 0;
  it->table = L.fbtab[tab];
  it->n = L.fbn[tab];
  it->yc = Lc.slab[Lc.cur];
  it->yf = L.slab[L.in_step ? L.last_out : L.cur];
  it->rR0 = strip_ptr(h, L, 0, 0);
  it->rR1 = strip_ptr(h, L, 1, 0);
  it->zL0 = strip_ptr(h, L, 0, 1);
  it->zL1 = strip_ptr(h, L, 1, 1);
  it->zR0 = strip_ptr(h, L, 0, 2);
  it->zR1 = strip_ptr(h, L, 1, 2);
  return true;
}

void prepare_auxiliary(oa_handle* h, int level, ProjItem* pj, bool* proj, DelItem* dd, bool* del) {
  Level& L = h->L[level];
  const bool time0 = (L.t == 0.0);
  double* y = L.slab[L.in_step ? L.last_out : L.cur];
  L.dtb_pre_valid = false;   // derivatives are re-formed from the current state
  *proj = L.dirty;
  if (L.dirty) {
    *pj = ProjItem{};
    pj->G = L.G;
    pj->P = h->phys;
    pj->time0 = time0 ? 1 : 0;
    pj->y = y;
    L.dirty = false;
    L.aux_phi = time0;
    if (!time0 && h->p.chimethod) L.phi_valid = false;
  }
  *del = time0;
  if (time0) {
    *dd = DelItem{};
    dd->G = L.G;
    dd->y = y;
  }
}

// The coarse points a restriction may write: r0 in (rmin, rmax), z0 in
// (zmin, zmax), with the bounds of restrict_point (restrict.f90:88-162).
// The box adds a point of margin on each side and is clipped to the coarse
// level; the kernels still test every point against the exact bounds.
void restrict_box(const KGrid& Gf, const KGrid& Gc, int* ic0, int* nic, int* jc0, int* njc) {
  const int g = Gc.g;
  const double deltar = 2.0 * g * Gf.dr + 1.0e-10, deltaz = 2.0 * g * Gf.dz + 1.0e-10;
  const double rminf = (1 - g - 0.5) * Gf.dr, rmaxf = (Gf.Nr - 0.5) * Gf.dr;
  const double zminf = (1 - g - Gf.zoff) * Gf.dz, zmaxf = (Gf.Nz - Gf.zoff) * Gf.dz;
  const double rmin = (rminf < 0.0) ? 0.0 : rminf + deltar, rmax = rmaxf - deltar;
  const double zmin = (Gf.eqsym && zminf < 0.0) ? 0.0 : zminf + deltaz, zmax = zmaxf - deltaz;
  // coarse r_i = (i - 1/2) dr, z_j = (j - zoff) dz
  const int ilo = std::max(1 - g, static_cast<int>(std::floor(rmin / Gc.dr + 0.5)) - 1);
  const int ihi = std::min(Gc.Nr, static_cast<int>(std::ceil(rmax / Gc.dr + 0.5)) + 1);
  const int jlo = std::max(1 - g, static_cast<int>(std::floor(zmin / Gc.dz + Gc.zoff)) - 1);
  const int jhi = std::min(Gc.Nz, static_cast<int>(std::ceil(zmax / Gc.dz + Gc.zoff)) + 1);
  *ic0 = ilo;
  *nic = std::max(1, ihi - ilo + 1);
  *jc0 = jlo;
  *njc = std::max(1, jhi - jlo + 1);
  if (*ic0 + *nic - 1 > Gc.Nr) *nic = Gc.Nr - *ic0 + 1;
  if (*jc0 + *njc - 1 > Gc.Nz) *njc = Gc.Nz - *jc0 + 1;
}

void prepare_restrict(oa_handle* h, int level, RsItem* it) {
  Level& L = h->L[level];
  Level& Lc = h->L[level - 1];
  *it = RsItem{};
  it->Gf = L.G;
  it->Gc = Lc.G;
  it->nf = h->nf;
  it->interporder = h->p.interporder;
  it->yf = L.slab[L.cur];
  it->yc = Lc.slab[Lc.cur];
  restrict_box(it->Gf, it->Gc, &it->ic0, &it->nic, &it->jc0, &it->njc);
  Lc.dirty = true;
  Lc.cons_valid = false;
}

void finish_level_step(oa_handle* h, int level) {
  Level& L = h->L[level];
  L.t2 = L.t1;
  L.t1 = L.t;
  L.s = L.s + 1;
  L.t = L.t + L.G.dt;
  h->point_updates += static_cast<double>(h->p.Nrtotal) * static_cast<double>(h->p.Nztotal);
}

}  // namespace oax

namespace {

// The piecewise entry points: one launch each, through the handle's plan.
int do_saveold(oa_handle* h, int level) {
  SaveItem it;
  RC(prepare_saveold(h, level, &it));
  h->plan.begin();
  h->plan.add_one(K_SAVE, false, false, false, it);
  return h->plan.submit(h->stream);
}

int do_substep(oa_handle* h, int level, int iter, int niter, double dtw, double w) {
  RhsItem it;
  bool phimode = false;
  RC(prepare_substep(h, level, iter, niter, dtw, w, &it, &phimode));
  h->plan.begin();
  h->plan.add_one(K_RHS, h->shift, h->scalar, phimode, it);
  return h->plan.submit(h->stream);
}

int do_finebound(oa_handle* h, int level, double dtw) {
  FbItem it;
  if (!prepare_finebound(h, level, dtw, &it)) return OA_OK;
  h->plan.begin();
  h->plan.add_one(K_FB, h->shift, h->scalar, false, it);
  return h->plan.submit(h->stream);
}

int do_auxiliary(oa_handle* h, int level) {
  ProjItem pj;
  DelItem dd;
  bool proj = false, del = false;
  prepare_auxiliary(h, level, &pj, &proj, &dd, &del);
  h->plan.begin();
  if (proj) h->plan.add_one(K_PROJ, h->shift, h->scalar, false, pj);
  if (del) h->plan.add_one(K_DEL, false, false, false, dd);
  return h->plan.submit(h->stream);
}

int do_restrict(oa_handle* h, int level) {
  RsItem it;
  prepare_restrict(h, level, &it);
  h->plan.begin();
  h->plan.add_one(K_RS, false, false, false, it);
  return h->plan.submit(h->stream);
}

// one coarse step: onestep(0) (onestep.f90:78-469) planned as a whole and
// submitted with one transfer of its work items
int coarse_step(oa_handle* h) {
  h->plan.begin();
  RC(plan_level(h->plan, std::vector<oa_handle*>{h}, 0));
  return h->plan.submit(h->stream, true);
}

void fill_grid(KGrid& G, const oa_params& p, int level) {
  const int g = p.ghost;
  G.Nr = p.Nrtotal;
  G.Nz = p.Nztotal;
  G.g = g;
  const int nzt = G.Nz + g;
  G.ld = (nzt + 15) / 16 * 16;
  G.plane = static_cast<long>(G.Nr + g) * G.ld;
  G.eqsym = p.eqsym;
  G.jlo = p.eqsym ? 1 : 1 - g;
  const double scale = static_cast<double>(1 << level);
  G.dr = p.dr0 / scale;
  G.dz = p.dz0 / scale;
  G.idr = 1.0 / G.dr;
  G.idz = 1.0 / G.dz;
  G.zoff = p.eqsym ? 0.5 : 0.5 * static_cast<double>(G.Nz - g + 1);
}

}  // namespace

// ================================================================== C ABI
extern "C" int oa_create(const oa_params* pin, oa_handle** out) {
  if (!pin || !out) {
    oa_set_error("oa_create: null argument");
    return OA_ERR_ARG;
  }
  *out = nullptr;
  char msg[256];
  RC(oa_params_check_supported(pin, msg, sizeof msg));
  oa_handle* h = new oa_handle();
  h->p = *pin;
  oa_params& p = h->p;
  p.ghost = 3;
  if (!p.eqsym && ((p.Nztotal + p.ghost) % 2 == 0)) p.Nztotal += 1;   // setupboxes.f90:104-108
  h->shift = (p.shift != OA_SHIFT_NONE);
  h->scalar = (p.mattertype & OA_MATTER_SCALAR) != 0;
  h->nf = nslots(h->shift, h->scalar);
  h->nlev = p.Nl0 + 1;
  h->dt0 = p.dtfac * std::fmin(p.dr0, p.dz0);
  fill_phys(h);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    delete h;
    oa_set_error("no HIP device available: this library runs on the GPU only");
    return OA_ERR_HIP;
  }
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    oa_set_error("hipGetDevice failed");
    return OA_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev_begin) != hipSuccess || hipEventCreate(&h->ev_end) != hipSuccess) {
    delete h;
    oa_set_error("HIP stream/event creation failed");
    return OA_ERR_HIP;
  }
  h->stream = h->own_stream;
  if (h->plan.init() != OA_OK) {
    oa_destroy(h);
    oa_set_error("HIP event creation failed");
    return OA_ERR_HIP;
  }
  if (hipMalloc(reinterpret_cast<void**>(&h->norms_part), sizeof(double) * 6 * norms_blocks()) != hipSuccess) {
    oa_destroy(h);
    oa_set_error("device allocation failed");
    return OA_ERR_NOMEM;
  }
  h->L.resize(h->nlev);
  for (int l = 0; l < h->nlev; ++l) {
    Level& L = h->L[l];
    fill_grid(L.G, p, l);
    L.G.dt = h->dt0 / static_cast<double>(1 << l);   // setupboxes.f90:121-125
    const size_t bytes = sizeof(double) * static_cast<size_t>(h->nf) * L.G.plane;
    bool ok = true;
    for (int b = 0; b < 3; ++b) ok = ok && hipMalloc(&L.slab[b], bytes) == hipSuccess;
    ok = ok && hipMalloc(&L.acc, bytes) == hipSuccess;
    L.strip_r = static_cast<long>(L.G.g) * (L.G.Nz + L.G.g);
    L.strip_z = static_cast<long>(L.G.g) * (L.G.Nr + L.G.g);
    const size_t sbytes = sizeof(double) * static_cast<size_t>(h->nf) * (L.strip_r + 2 * L.strip_z);
    for (int b = 0; b < 3; ++b) ok = ok && hipMalloc(&L.strip[b], sbytes) == hipSuccess;
    if (h->shift) ok = ok && hipMalloc(&L.dtb_pre, sizeof(double) * 2 * static_cast<size_t>(L.G.plane)) == hipSuccess;
    ok = ok && hipMalloc(&L.gaux, sizeof(double) * NAUX * static_cast<size_t>(L.G.plane)) == hipSuccess;
    if (l > 0) {
      const int depth[2] = {L.G.g + 2, L.G.g - 1};
      for (int k = 0; k < 2; ++k) {
        L.fbn[k] = fb_points(L.G, depth[k]);
        ok = ok && hipMalloc(reinterpret_cast<void**>(&L.fbtab[k]), fb_record_bytes() * L.fbn[k]) == hipSuccess;
        ok = ok && launch_fb_table(L.G, h->L[l - 1].G, depth[k], p.interporder, L.fbtab[k], L.fbn[k], h->stream) ==
                       OA_OK;
      }
    }
    if (!ok) {
      oa_destroy(h);
      oa_set_error("device allocation failed");
      return OA_ERR_NOMEM;
    }
    for (int b = 0; b < 3; ++b) {
      (void)hipMemsetAsync(L.slab[b], 0, bytes, h->stream);
      (void)hipMemsetAsync(L.strip[b], 0, sbytes, h->stream);
    }
    (void)hipMemsetAsync(L.acc, 0, bytes, h->stream);
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    oa_destroy(h);
    oa_set_error("initial device clear failed");
    return OA_ERR_HIP;
  }
  *out = h;
  return OA_OK;
}

extern "C" int oa_destroy(oa_handle* h) {
  if (!h) return OA_OK;
  if (ah_pending(h)) {
    oa_set_error("oa_destroy: a begun horizon search of this spacetime has not ended (oa_ahfind_end)");
    return OA_ERR_STATE;
  }
  if (h->batch) (void)oa_batch_remove(h->batch, h);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (Level& L : h->L) {
    for (int b = 0; b < 3; ++b) {
      if (L.slab[b]) (void)hipFree(L.slab[b]);
      if (L.strip[b]) (void)hipFree(L.strip[b]);
    }
    if (L.acc) (void)hipFree(L.acc);
    if (L.dtb_pre) (void)hipFree(L.dtb_pre);
    if (L.gaux) (void)hipFree(L.gaux);
    for (int k = 0; k < 2; ++k)
      if (L.fbtab[k]) (void)hipFree(L.fbtab[k]);
    if (L.cons) (void)hipFree(L.cons);
  }
  ah_release(h);
  h->plan.release();
  if (h->norms_part) (void)hipFree(h->norms_part);
  if (h->ev_begin) (void)hipEventDestroy(h->ev_begin);
  if (h->ev_end) (void)hipEventDestroy(h->ev_end);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
  return OA_OK;
}

extern "C" int oa_num_levels(const oa_handle* h, int* n) {
  if (!h || !n) return OA_ERR_ARG;
  *n = h->nlev;
  return OA_OK;
}

extern "C" int oa_level_info(const oa_handle* h, int box, int level, int* Nr, int* Nz, int* ghost, double* dr,
                             double* dz, double* dt) {
  const Level* L = level_of(const_cast<oa_handle*>(h), box, level);
  if (!L) return OA_ERR_ARG;
  if (Nr) *Nr = L->G.Nr;
  if (Nz) *Nz = L->G.Nz;
  if (ghost) *ghost = L->G.g;
  if (dr) *dr = L->G.dr;
  if (dz) *dz = L->G.dz;
  if (dt) *dt = L->G.dt;
  return OA_OK;
}

extern "C" int oa_upload_field(oa_handle* h, int box, int level, const char* name, const double* src, int ld) {
  Level* L = level_of(h, box, level);
  if (!L || !name || !src) return OA_ERR_ARG;
  const int slot = slot_of(h, name);
  if (slot < 0) {
    oa_set_error(std::string("oa_upload_field: no evolved array named ") + name + " in this configuration");
    return OA_ERR_ARG;
  }
  const KGrid& G = L->G;
  const int nrt = G.Nr + G.g, nzt = G.Nz + G.g;
  if (ld < nrt) {
    oa_set_error("oa_upload_field: leading dimension smaller than Nr+ghost");
    return OA_ERR_ARG;
  }
  std::vector<double> plane(static_cast<size_t>(G.plane), 0.0);
  for (int jj = 0; jj < nzt; ++jj)
    for (int ii = 0; ii < nrt; ++ii) plane[static_cast<size_t>(ii) * G.ld + jj] = src[static_cast<size_t>(jj) * ld + ii];
  HIPCHK(hipMemcpyAsync(L->slab[L->cur] + static_cast<long>(slot) * G.plane, plane.data(),
                        sizeof(double) * plane.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  L->dirty = true;
  L->cons_valid = false;
  if (slot == S_phi) L->phi_valid = true;
  return OA_OK;
}

extern "C" int oa_download_field(oa_handle* h, int box, int level, const char* name, double* dst, int ld) {
  Level* L = level_of(h, box, level);
  if (!L || !name || !dst) return OA_ERR_ARG;
  const KGrid& G = L->G;
  const int nrt = G.Nr + G.g, nzt = G.Nz + G.g;
  if (ld < nrt) {
    oa_set_error("oa_download_field: leading dimension smaller than Nr+ghost");
    return OA_ERR_ARG;
  }
  const std::string nm(name);
  auto fetch = [&](int slot, std::vector<double>& buf) -> int {
    buf.resize(static_cast<size_t>(G.plane));
    double* y = L->slab[L->in_step ? L->last_out : L->cur];
    HIPCHK(hipMemcpyAsync(buf.data(), y + static_cast<long>(slot) * G.plane, sizeof(double) * buf.size(),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return OA_OK;
  };
  auto emit = [&](auto fn) {
    for (int jj = 0; jj < nzt; ++jj)
      for (int ii = 0; ii < nrt; ++ii) dst[static_cast<size_t>(jj) * ld + ii] = fn(static_cast<size_t>(ii) * G.ld + jj, ii, jj);
  };
  const int slot = slot_of(h, nm);
  if (slot >= 0 && !(slot == S_phi && !L->phi_valid)) {
    std::vector<double> b;
    RC(fetch(slot, b));
    emit([&](size_t q, int, int) { return b[q]; });
    return OA_OK;
  }
  if (nm == "phi" || nm == "psi" || nm == "psi4") {
    std::vector<double> c;
    if (nm == "phi" || !L->phi_valid) {
      RC(fetch(S_chi, c));
      if (nm == "phi") emit([&](size_t q, int, int) { return -0.5 * std::log(c[q]); });
      else if (nm == "psi") emit([&](size_t q, int, int) { return std::exp(-0.5 * std::log(c[q])); });
      else emit([&](size_t q, int, int) { return 1.0 / (c[q] * c[q]); });
    } else {
      RC(fetch(S_phi, c));
      if (nm == "psi") emit([&](size_t q, int, int) { return std::exp(c[q]); });
      else emit([&](size_t q, int, int) { const double e = std::exp(c[q]); return e * e * e * e; });
    }
    return OA_OK;
  }
  if (nm == "KTH" || nm == "lambda" || nm == "Alambda") {
    std::vector<double> A, B, H, C, KA, KB, KC;
    RC(fetch(S_A, A)); RC(fetch(S_B, B)); RC(fetch(S_H, H)); RC(fetch(S_C, C));
    RC(fetch(S_KTA, KA)); RC(fetch(S_KTB, KB)); RC(fetch(S_KTC, KC));
    emit([&](size_t q, int ii, int) {
      const double r = (static_cast<double>(ii + 1 - G.g) - 0.5) * G.dr;
      const double gA = B[q] * H[q], gB = A[q] * H[q], gC = -C[q] * H[q];
      const double kth = -(gA * KA[q] + gB * KB[q] + 2.0 * r * r * gC * KC[q]) * H[q];
      if (nm == "KTH") return kth;
      if (nm == "lambda") return (A[q] - H[q]) / (r * r);
      return (KA[q] - kth) / (r * r);
    });
    return OA_OK;
  }
  if (cons_slot(nm) >= 0) {
    if (!L->cons_valid) {
      oa_set_error(std::string("oa_download_field: ") + name + " needs oa_constraints on the current state first");
      return OA_ERR_STATE;
    }
    std::vector<double> b(static_cast<size_t>(G.plane));
    HIPCHK(hipMemcpyAsync(b.data(), L->cons + static_cast<long>(cons_slot(nm)) * G.plane, sizeof(double) * b.size(),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    emit([&](size_t q, int, int) { return b[q]; });
    return OA_OK;
  }
  if (nm == "dtalpha" || nm == "dtphi" || nm == "z4theta") {
    emit([&](size_t, int, int) { return 0.0; });
    return OA_OK;
  }
  oa_set_error(std::string("oa_download_field: unknown array ") + name);
  return OA_ERR_ARG;
}

extern "C" int oa_constraints(oa_handle* h, int box, int level) {
  Level* L = level_of(h, box, level);
  if (!L) return OA_ERR_ARG;
  if (L->dirty) {
    oa_set_error("oa_constraints: state not projected (call oa_auxiliary / oa_initial_aux)");
    return OA_ERR_STATE;
  }
  if (!L->cons) {
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMalloc(reinterpret_cast<void**>(&L->cons), sizeof(double) * 5 * static_cast<size_t>(L->G.plane)));
    HIPCHK(hipMemsetAsync(L->cons, 0, sizeof(double) * 5 * static_cast<size_t>(L->G.plane), h->stream));
  }
  KStep S{1, 1, 0, 0.0, 0.0, L->t == 0.0 ? 1 : 0, nullptr};
  const double* y = L->slab[L->in_step ? L->last_out : L->cur];
  RC(launch_constraints(L->G, h->phys, S, L->aux_phi, h->scalar, scalar_slot(h->shift, SC_phi), y, L->gaux,
                        L->cons, h->stream));
  L->cons_valid = true;
  return OA_OK;
}

namespace {
// the plane oa_norms0d reduces for `name`: an evolved array whose stored
// values are current, or a constraint computed on the current state
int norms_plane(oa_handle* h, Level* L, const char* caller, const char* name, const double** out) {
  const int slot = slot_of(h, name);
  const int cslot = cons_slot(name);
  if (slot >= 0 && !(slot == S_phi && !L->phi_valid)) {
    *out = L->slab[L->in_step ? L->last_out : L->cur] + static_cast<long>(slot) * L->G.plane;
    return OA_OK;
  }
  if (cslot >= 0 && L->cons_valid) {
    *out = L->cons + static_cast<long>(cslot) * L->G.plane;
    return OA_OK;
  }
  oa_set_error(std::string(caller) + ": " + name +
               " is not a stored evolved array of this configuration (constraints need oa_constraints first)");
  return OA_ERR_ARG;
}
}  // namespace

extern "C" int oa_norms0d(oa_handle* h, int box, int level, const char* name, double* out) {
  Level* L = level_of(h, box, level);
  if (!L || !name || !out) return OA_ERR_ARG;
  const double* plane = nullptr;
  RC(norms_plane(h, L, "oa_norms0d", name, &plane));
  RC(launch_norms_partial(L->G, plane, h->norms_part, h->stream));
  const int nb = norms_blocks();
  std::vector<double> part(static_cast<size_t>(6 * nb));
  HIPCHK(hipMemcpyAsync(part.data(), h->norms_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double vmax = -1.0e300, vmin = 1.0e300, vabs = 0.0, sq = 0.0, cnt = 0.0, nan = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double* o = &part[static_cast<size_t>(6 * b)];
    vmax = std::fmax(vmax, o[0]);
    vmin = std::fmin(vmin, o[1]);
    vabs = std::fmax(vabs, o[2]);
    sq += o[3];
    cnt += o[4];
    nan = std::fmax(nan, o[5]);
  }
  if (nan != 0.0) {
    oa_set_error(std::string("oa_norms0d: found NaN in ") + name);   // save0Ddata.f90:230-237 dies here
    return OA_ERR_STATE;
  }
  out[0] = vmax;
  out[1] = vmin;
  out[2] = vabs;
  out[3] = (cnt > 0.0) ? std::sqrt(sq / cnt) : 0.0;
  return OA_OK;
}

extern "C" int oa_norms0d_noalloc(oa_handle* h, int box, int level, int n, const char* const* names,
                                  double* out_dev) {
  Level* L = level_of(h, box, level);
  if (!L || !names || !out_dev) {
    if (L) oa_set_error("oa_norms0d_noalloc: null argument");
    return OA_ERR_ARG;
  }
  if (n < 1 || n > norms_max_fields()) {
    oa_set_error("oa_norms0d_noalloc: " + std::to_string(n) + " arrays requested; one call takes 1 to " +
                 std::to_string(norms_max_fields()));
    return OA_ERR_ARG;
  }
  const double* planes[64];
  for (int k = 0; k < n; ++k) {
    if (!names[k]) {
      oa_set_error("oa_norms0d_noalloc: null array name");
      return OA_ERR_ARG;
    }
    RC(norms_plane(h, L, "oa_norms0d_noalloc", names[k], &planes[k]));
  }
  return launch_norms_fields_noalloc(h->device, L->G, n, planes, out_dev, h->stream);
}

extern "C" int oa_set_elliptic_exit(oa_handle* h, int stationary) {
  if (!h) return OA_ERR_ARG;
  h->ell_stationary = stationary ? 1 : 0;
  return OA_OK;
}

extern "C" int oa_elliptic_stats(const oa_handle* h, int* steps0, int* steps1, double* residual) {
  if (!h) return OA_ERR_ARG;
  if (steps0) *steps0 = h->ell_steps[0];
  if (steps1) *steps1 = h->ell_steps[1];
  if (residual) *residual = h->ell_res;
  return OA_OK;
}

extern "C" int oa_params_footprint(const oa_params* pin, double* bytes) {
  if (!pin || !bytes) return OA_ERR_ARG;
  oa_params p = *pin;
  p.ghost = 3;
  if (!p.eqsym && ((p.Nztotal + p.ghost) % 2 == 0)) p.Nztotal += 1;
  const bool shift = (p.shift != OA_SHIFT_NONE), scalar = (p.mattertype & OA_MATTER_SCALAR) != 0;
  const double nf = nslots(shift, scalar);
  double total = sizeof(double) * 6.0 * norms_blocks();
  for (int l = 0; l <= p.Nl0; ++l) {
    KGrid G;
    fill_grid(G, p, l);
    const double plane = static_cast<double>(G.plane) * sizeof(double);
    const double strips = static_cast<double>(G.g) * ((G.Nz + G.g) + 2.0 * (G.Nr + G.g)) * sizeof(double);
    total += 4.0 * nf * plane;              // three state slabs and the RK4 accumulator
    total += 3.0 * nf * strips;             // fine-boundary time levels
    total += NAUX * plane;                  // geometric auxiliaries
    if (shift) total += 2.0 * plane;        // pre-initialisation dtbeta
    if (l > 0)
      total += static_cast<double>(fb_record_bytes()) *
               static_cast<double>(fb_points(G, G.g + 2) + fb_points(G, G.g - 1));
    if (p.ahfind) total += ah_record_doubles() * plane;   // horizon finder: point records
  }
  if (p.ahfind) total += ah_fixed_bytes(p.N_ah < 1 ? 1 : (p.N_ah > 3 ? 3 : p.N_ah), p.ahmethod);
  *bytes = total;
  return OA_OK;
}

// ------------------------------------------------------------ checkpoint
// Snapshot layout: a header (magic, version, level count, field count, and
// per level the plane and strip sizes that a matching handle must have),
// then per level the clocks, the flags and the device data: the current
// state slab, the three fine-boundary time levels in time order and, with
// the Gamma-driver, the pre-initialisation dtbeta planes.
namespace {

constexpr uint64_t kCkptMagic = 0x3154504b43584f41ull;   // "AOXCKPT1"

struct CkptLevel {
  double t, t1, t2;
  int32_t s, dirty, phi_valid, aux_phi, dtb_pre_valid, pad;
};

size_t ckpt_level_bytes(const oa_handle* h, const Level& L) {
  const size_t slab = sizeof(double) * static_cast<size_t>(h->nf) * L.G.plane;
  const size_t strips = sizeof(double) * static_cast<size_t>(h->nf) * (L.strip_r + 2 * L.strip_z);
  const size_t pre = h->shift ? sizeof(double) * 2 * static_cast<size_t>(L.G.plane) : 0;
  return sizeof(CkptLevel) + slab + 3 * strips + pre;
}

size_t ckpt_header_bytes(const oa_handle* h) { return sizeof(uint64_t) * (4 + 3 * static_cast<size_t>(h->nlev)); }

int ckpt_check(const oa_handle* h, const char* who) {
  for (const Level& L : h->L)
    if (L.in_step) {
      oa_set_error(std::string(who) + ": a level is inside a step (checkpoint between coarse steps)");
      return OA_ERR_STATE;
    }
  return OA_OK;
}

}  // namespace

extern "C" int oa_checkpoint_size(const oa_handle* h, size_t* bytes) {
  if (!h || !bytes) return OA_ERR_ARG;
  size_t n = ckpt_header_bytes(h);
  for (const Level& L : h->L) n += ckpt_level_bytes(h, L);
  *bytes = n;
  return OA_OK;
}

extern "C" int oa_checkpoint_save(oa_handle* h, void* buf, size_t bytes) {
  if (!h || !buf) return OA_ERR_ARG;
  RC(ckpt_check(h, "oa_checkpoint_save"));
  size_t need = 0;
  RC(oa_checkpoint_size(h, &need));
  if (bytes < need) {
    oa_set_error("oa_checkpoint_save: buffer smaller than oa_checkpoint_size");
    return OA_ERR_ARG;
  }
  char* p = static_cast<char*>(buf);
  uint64_t* hd = reinterpret_cast<uint64_t*>(p);
  hd[0] = kCkptMagic;
  hd[1] = 1;
  hd[2] = static_cast<uint64_t>(h->nlev);
  hd[3] = static_cast<uint64_t>(h->nf);
  for (int l = 0; l < h->nlev; ++l) {
    const Level& L = h->L[l];
    hd[4 + 3 * l] = static_cast<uint64_t>(L.G.plane);
    hd[5 + 3 * l] = static_cast<uint64_t>(L.strip_r);
    hd[6 + 3 * l] = static_cast<uint64_t>(L.strip_z);
  }
  p += ckpt_header_bytes(h);
  for (Level& L : h->L) {
    CkptLevel c{L.t, L.t1, L.t2, L.s, L.dirty ? 1 : 0, L.phi_valid ? 1 : 0, L.aux_phi ? 1 : 0,
                L.dtb_pre_valid ? 1 : 0, 0};
    std::memcpy(p, &c, sizeof c);
    p += sizeof c;
    const size_t slab = sizeof(double) * static_cast<size_t>(h->nf) * L.G.plane;
    HIPCHK(hipMemcpyAsync(p, L.slab[L.cur], slab, hipMemcpyDeviceToHost, h->stream));
    p += slab;
    const size_t strips = sizeof(double) * static_cast<size_t>(h->nf) * (L.strip_r + 2 * L.strip_z);
    for (int tl = 0; tl < 3; ++tl) {
      HIPCHK(hipMemcpyAsync(p, L.strip[L.ring[tl]], strips, hipMemcpyDeviceToHost, h->stream));
      p += strips;
    }
    if (h->shift) {
      const size_t pre = sizeof(double) * 2 * static_cast<size_t>(L.G.plane);
      HIPCHK(hipMemcpyAsync(p, L.dtb_pre, pre, hipMemcpyDeviceToHost, h->stream));
      p += pre;
    }
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return OA_OK;
}

extern "C" int oa_checkpoint_load(oa_handle* h, const void* buf, size_t bytes) {
  if (!h || !buf) return OA_ERR_ARG;
  RC(ckpt_check(h, "oa_checkpoint_load"));
  size_t need = 0;
  RC(oa_checkpoint_size(h, &need));
  const char* p = static_cast<const char*>(buf);
  const uint64_t* hd = reinterpret_cast<const uint64_t*>(p);
  bool ok = bytes >= need && hd[0] == kCkptMagic && hd[1] == 1 && hd[2] == static_cast<uint64_t>(h->nlev) &&
            hd[3] == static_cast<uint64_t>(h->nf);
  for (int l = 0; ok && l < h->nlev; ++l) {
    const Level& L = h->L[l];
    ok = hd[4 + 3 * l] == static_cast<uint64_t>(L.G.plane) && hd[5 + 3 * l] == static_cast<uint64_t>(L.strip_r) &&
         hd[6 + 3 * l] == static_cast<uint64_t>(L.strip_z);
  }
  if (!ok) {
    oa_set_error("oa_checkpoint_load: the snapshot does not belong to a spacetime with these parameters");
    return OA_ERR_ARG;
  }
  p += ckpt_header_bytes(h);
  for (Level& L : h->L) {
    CkptLevel c;
    std::memcpy(&c, p, sizeof c);
    p += sizeof c;
    const size_t slab = sizeof(double) * static_cast<size_t>(h->nf) * L.G.plane;
    HIPCHK(hipMemcpyAsync(L.slab[L.cur], p, slab, hipMemcpyHostToDevice, h->stream));
    p += slab;
    const size_t strips = sizeof(double) * static_cast<size_t>(h->nf) * (L.strip_r + 2 * L.strip_z);
    for (int tl = 0; tl < 3; ++tl) {
      HIPCHK(hipMemcpyAsync(L.strip[L.ring[tl]], p, strips, hipMemcpyHostToDevice, h->stream));
      p += strips;
    }
    if (h->shift) {
      const size_t pre = sizeof(double) * 2 * static_cast<size_t>(L.G.plane);
      HIPCHK(hipMemcpyAsync(L.dtb_pre, p, pre, hipMemcpyHostToDevice, h->stream));
      p += pre;
    }
    L.t = c.t;
    L.t1 = c.t1;
    L.t2 = c.t2;
    L.s = c.s;
    L.dirty = c.dirty != 0;
    L.phi_valid = c.phi_valid != 0;
    L.aux_phi = c.aux_phi != 0;
    L.dtb_pre_valid = c.dtb_pre_valid != 0;
    L.cons_valid = false;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return OA_OK;
}

// ------------------------------------------------------------ reference checkpoint
// checkpointrestart (src/base/checkpoint.f90:338-690): the state a
// single-process reference run wrote with checkpointsave (:2-328) into a
// directory -- checkparam (grid and symmetry settings), checklist (the par
// file's name, then one array name per line), per array and level the
// <name>_b0l<level>.2D dump (save2Dvariable: "#Time = t", then r, z, value
// for j = 1-g..Nz, i = 1-g..Nr, 9 significant digits) and four strip files
// <name>_bound_{rL,rR,zL,zR}_b0l<level> ("t t1 t2", a blank line, then the
// three time levels of every strip point, 16 digits).  The arrays this
// spacetime evolves are read; the others the reference lists (dtalpha, KTH,
// lambda, hdet, ...) are not formed on this path and are skipped, as are
// the rL strips (box 0 touches the axis).  The clocks come from the strip
// headers and the step counters restart at 0, as the reference's do.
namespace {

bool read_values(std::FILE* f, double* v, int n) {
  for (int k = 0; k < n; ++k)
    if (std::fscanf(f, "%lf", &v[k]) != 1) return false;   // list-directed: line breaks anywhere
  return true;
}

struct TextFile {
  std::FILE* f = nullptr;
  explicit TextFile(const std::string& path) : f(std::fopen(path.c_str(), "r")) {}
  ~TextFile() {
    if (f) std::fclose(f);
  }
};

int ckref_fail(const std::string& what) {
  oa_set_error("oa_checkpoint_restart: " + what);
  return OA_ERR_ARG;
}

// checkparam: the settings a restart must share (:400-530)
int ckref_params(const oa_handle* h, const std::string& dir) {
  TextFile in(dir + "/checkparam");
  if (!in.f) return ckref_fail("checkpoint directory \"" + dir + "\" has no checkparam");
  oa_params q;
  oa_params_default(&q);
  int nproctot = 1;
  char line[1024];
  while (std::fgets(line, sizeof line, in.f)) {
    std::string l(line);
    const size_t eq = l.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string x) {
      const size_t a = x.find_first_not_of(" \t\r\n"), b = x.find_last_not_of(" \t\r\n");
      return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
    };
    const std::string key = trim(l.substr(0, eq)), val = trim(l.substr(eq + 1));
    if (key == "nproctot") nproctot = std::atoi(val.c_str());
    else if (oa_params_set(&q, key.c_str(), val.c_str()) != OA_OK)
      return ckref_fail("checkparam: cannot read " + key + " = " + val);
  }
  const oa_params& p = h->p;
  if (nproctot != 1)
    return ckref_fail("the checkpoint was written by " + std::to_string(nproctot) +
                      " processes; only single-process checkpoints are read");
  if ((q.eqsym != 0) != (p.eqsym != 0))
    return ckref_fail("When restarting from a checkpoint the equatorial symmetry flag must be the same.");
  if ((q.angmom != 0) != (p.angmom != 0))
    return ckref_fail("When restarting from a checkpoint the angular momentum flag must be the same.");
  if (q.dr0 != p.dr0 || q.dz0 != p.dz0)
    return ckref_fail("When restarting from a checkpoint the grid spacing must be the same.");
  if (q.Nrtotal != p.Nrtotal || q.Nztotal != p.Nztotal)
    return ckref_fail("When restarting from a checkpoint the grid size must be the same.");
  if (q.Nb != p.Nb) return ckref_fail("When restarting from a checkpoint the number of boxes must be the same.");
  if (q.Nl0 != p.Nl0 || q.Nl1 != p.Nl1 || q.Nl2 != p.Nl2 || q.Nl3 != p.Nl3)
    return ckref_fail("When restarting from a checkpoint the number of levels must be the same.");
  if (q.mattertype != p.mattertype)
    return ckref_fail("When restarting from a checkpoint the type of matter must be the same.");
  return OA_OK;
}

}  // namespace

namespace {

// Fortran ESw.d of one value, right-justified in w columns (the exponent
// loses its E above 99, as the edit descriptor without Ee writes it)
std::string fortran_es(double x, int w, int d) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*E", d, x);
  std::string s(buf);
  const size_t e = s.find('E');
  std::string mant = s.substr(0, e);
  const int ev = std::atoi(s.c_str() + e + 1);
  char ex[16];
  if (ev >= -99 && ev <= 99) std::snprintf(ex, sizeof ex, "E%c%02d", ev < 0 ? '-' : '+', ev < 0 ? -ev : ev);
  else std::snprintf(ex, sizeof ex, "%c%03d", ev < 0 ? '-' : '+', ev < 0 ? -ev : ev);
  std::string out = mant + ex;
  if (static_cast<int>(out.size()) > w) return std::string(static_cast<size_t>(w), '*');
  return std::string(static_cast<size_t>(w) - out.size(), ' ') + out;
}

// list-directed output of one "key = value" line of checkparam (:180-205)
std::string ld_int(const char* key, int v) {
  char b[128];
  std::snprintf(b, sizeof b, " %s = %12d\n", key, v);
  return b;
}
std::string ld_real(const char* key, double v) {
  return std::string(" ") + key + " = " + fortran_es(v, 25, 16) + "\n";
}
std::string ld_logical(const char* key, int v) { return std::string(" ") + key + " =  " + (v ? "T" : "F") + "\n"; }

int ckref_write(const std::string& path, const std::string& text) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return ckref_fail("cannot write " + path);
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  std::fclose(f);
  return ok ? OA_OK : ckref_fail("cannot write " + path);
}

}  // namespace

// checkpointsave (checkpoint.f90:2-328) for a single process into the
// directory dir (made if missing; the reference names it
// <directory>/checkpoint_t=<t>): checkparam, checklist (parfile, then r, z,
// rr and the arrays this spacetime evolves), every listed array's .2D dump
// (save2Dvariable with status replace: 9 significant digits, + 1e-30) and,
// except for the coordinates, its four _bound_ strip files with the level's
// clocks (16 digits).  The rL strips, which box 0 never fills, hold zeros.
extern "C" int oa_checkpoint_write(oa_handle* h, const char* dir, const char* parfile) {
  if (!h || !dir) return OA_ERR_ARG;
  RC(ckpt_check(h, "oa_checkpoint_write"));
  const std::string d(dir);
  for (size_t k = 1; k <= d.size(); ++k)                 // mkdir -p
    if (k == d.size() || d[k] == '/') (void)::mkdir(d.substr(0, k).c_str(), 0755);
  const oa_params& p = h->p;
  std::string cp;
  cp += ld_int("nproctot", 1);
  cp += ld_logical("eqsym", p.eqsym);
  cp += ld_logical("angmom", p.angmom);
  cp += ld_real("dr0", p.dr0);
  cp += ld_real("dz0", p.dz0);
  cp += ld_int("Nrtotal", p.Nrtotal);
  cp += ld_int("Nztotal", p.Nztotal);
  cp += ld_int("Nb", p.Nb);
  cp += ld_int("Nl0", p.Nl0);
  cp += ld_int("Nl1", p.Nl1);
  cp += ld_int("Nl2", p.Nl2);
  cp += ld_int("Nl3", p.Nl3);
  for (const char* k : {"Nrbox1", "Nzbox1", "Nrbox2", "Nzbox2", "Nrbox3", "Nzbox3"}) cp += ld_int(k, 0);
  for (const char* k : {"rbox1", "zbox1", "rbox2", "zbox2", "rbox3", "zbox3"}) cp += ld_real(k, 0.0);
  cp += std::string(" mattertype = ") + ((p.mattertype & OA_MATTER_SCALAR) ? "scalar" : "vacuum") + "\n";
  RC(ckref_write(d + "/checkparam", cp));
  static const char* coords[] = {"r", "z", "rr"};
  std::vector<std::string> names(coords, coords + 3);
  {
    static const char* cand[] = {"alpha",    "chi",      "phi",         "A",           "B",           "H",
                                 "C",        "trK",      "KTA",         "KTB",         "KTC",         "Delta_r",
                                 "Delta_z",  "beta_r",   "beta_z",      "dtbeta_r",    "dtbeta_z",    "scalar_phi",
                                 "scalar_xi_r", "scalar_xi_z", "scalar_pi"};
    std::vector<std::pair<int, std::string>> ev;
    for (const char* c : cand)
      if (slot_of(h, c) >= 0) ev.push_back({slot_of(h, c), c});
    std::sort(ev.begin(), ev.end());
    for (auto& e : ev) names.push_back(e.second);
  }
  std::string cl = std::string(" ") + (parfile ? parfile : "checkpoint.par") + "\n";
  for (const std::string& n : names) cl += " " + n + "\n";
  RC(ckref_write(d + "/checklist", cl));

  for (int l = 0; l < h->nlev; ++l) {
    Level& L = h->L[l];
    const KGrid& G = L.G;
    const int g = G.g, nrt = G.Nr + g, nzt = G.Nz + g;
    const std::string filen = "_b0l" + std::to_string(l);
    const size_t sset = static_cast<size_t>(h->nf) * (L.strip_r + 2 * L.strip_z);
    std::vector<double> strips[3];
    for (int tl = 0; tl < 3; ++tl) {
      strips[tl].resize(sset);
      HIPCHK(hipMemcpyAsync(strips[tl].data(), L.strip[L.ring[tl]], sizeof(double) * sset, hipMemcpyDeviceToHost,
                            h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> plane(static_cast<size_t>(G.plane));
    for (const std::string& n : names) {
      const int slot = slot_of(h, n);
      if (slot >= 0) RC(oa_download_field(h, 0, l, n.c_str(), plane.data(), nrt));   // phi as the reference has it
      std::string out = "#Time = " + fortran_es(L.t, 23, 15) + "\n";
      for (int jj = 0; jj < nzt; ++jj) {
        const double z = (static_cast<double>(jj + 1 - g) - G.zoff) * G.dz;
        for (int ii = 0; ii < nrt; ++ii) {
          const double r = (static_cast<double>(ii + 1 - g) - 0.5) * G.dr;
          double v;
          if (slot >= 0) v = plane[static_cast<size_t>(jj) * nrt + ii];
          else if (n == "r") v = r;
          else if (n == "z") v = z;
          else v = std::sqrt(r * r + z * z);
          out += fortran_es(r, 16, 8) + fortran_es(z, 16, 8) + fortran_es(v + 1.0e-30, 16, 8) + "\n";
        }
        out += "\n";
      }
      out += "\n";
      RC(ckref_write(d + "/" + n + filen + ".2D", out));
      if (slot < 0) continue;
      const std::string head = fortran_es(L.t, 23, 15) + fortran_es(L.t1, 23, 15) + fortran_es(L.t2, 23, 15) + "\n\n";
      const char* piece_name[4] = {"_bound_rL", "_bound_rR", "_bound_zL", "_bound_zR"};
      for (int pc = 0; pc < 4; ++pc) {
        std::string sv = head;
        const int piece = pc - 1;                      // -1: rL (zeros)
        const long per = piece <= 0 ? L.strip_r : L.strip_z;
        const long off = piece <= 0 ? 0 : (piece == 1 ? h->nf * L.strip_r : h->nf * L.strip_r + h->nf * L.strip_z);
        const int outer = piece <= 0 ? g : nrt, inner = piece <= 0 ? nzt : g;
        for (int a = 0; a < outer; ++a)
          for (int b = 0; b < inner; ++b) {
            const long q = piece <= 0 ? static_cast<long>(a) * nzt + b : static_cast<long>(b) * nrt + a;
            for (int tl = 0; tl < 3; ++tl)
              sv += fortran_es(piece < 0 ? 0.0 : strips[tl][static_cast<size_t>(off + slot * per + q)], 23, 15);
            sv += "\n";
          }
        RC(ckref_write(d + "/" + n + piece_name[pc] + filen, sv));
      }
    }
  }
  return OA_OK;
}

extern "C" int oa_checkpoint_restart(oa_handle* h, const char* dir) {
  if (!h || !dir) return OA_ERR_ARG;
  RC(ckpt_check(h, "oa_checkpoint_restart"));
  if (ah_pending(h)) {
    oa_set_error("oa_checkpoint_restart: a begun horizon search of this spacetime has not ended");
    return OA_ERR_STATE;
  }
  const std::string d(dir);
  RC(ckref_params(h, d));
  // a restarted run is a new run of the reference: its horizon finder starts
  // afresh (ahfound, ack and the first-call flags are not checkpointed)
  ah_release(h);
  std::vector<std::string> names;                                  // checklist (:553-565)
  {
    TextFile in(d + "/checklist");
    if (!in.f) return ckref_fail("checkpoint directory \"" + d + "\" has no checklist");
    char line[1024];
    bool first = true;
    while (std::fgets(line, sizeof line, in.f)) {
      char word[1024];
      if (std::sscanf(line, "%1023s", word) != 1) continue;
      if (first) {                                                   // the par file's name
        first = false;
        continue;
      }
      names.push_back(word);
    }
  }
  std::vector<bool> seen(static_cast<size_t>(h->nf), false);
  for (const std::string& n : names) {
    const int slot = slot_of(h, n);
    if (slot >= 0) seen[static_cast<size_t>(slot)] = true;
  }
  for (int sl = 0; sl < h->nf; ++sl)
    if (!seen[static_cast<size_t>(sl)]) return ckref_fail("the checklist lacks an array this spacetime evolves");

  for (int l = 0; l < h->nlev; ++l) {
    Level& L = h->L[l];
    const KGrid& G = L.G;
    const int g = G.g, nrt = G.Nr + g, nzt = G.Nz + g;
    const std::string filen = "_b0l" + std::to_string(l);
    std::vector<double> plane(static_cast<size_t>(G.plane));
    const size_t sset = static_cast<size_t>(h->nf) * (L.strip_r + 2 * L.strip_z);
    std::vector<double> strips[3];
    for (auto& v : strips) v.assign(sset, 0.0);
    double clk[3] = {L.t, L.t1, L.t2};
    for (const std::string& n : names) {
      const int slot = slot_of(h, n);
      if (slot < 0) continue;
      {                                                              // read2Ddata (read2Ddata.f90:38-66)
        const std::string path = d + "/" + n + filen + ".2D";
        TextFile in(path);
        if (!in.f) return ckref_fail("cannot open " + path);
        char line[1024];
        const long start = std::ftell(in.f);
        bool header = false;
        if (std::fgets(line, sizeof line, in.f)) {
          const char* c = line;
          while (*c == ' ' || *c == '\t') ++c;
          header = *c == '#';
        }
        if (!header) std::fseek(in.f, start, SEEK_SET);
        std::fill(plane.begin(), plane.end(), 0.0);
        for (int jj = 0; jj < nzt; ++jj)
          for (int ii = 0; ii < nrt; ++ii) {
            double v[3];
            if (!read_values(in.f, v, 3)) return ckref_fail(path + ": fewer values than the level holds");
            plane[static_cast<size_t>(ii) * G.ld + jj] = v[2];
          }
        HIPCHK(hipMemcpyAsync(L.slab[L.cur] + static_cast<long>(slot) * G.plane, plane.data(),
                              sizeof(double) * plane.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
      }
      // the strips: rR (piece 0), zL (1), zR (2); the file loops of :596-650
      const char* piece_name[3] = {"_bound_rR", "_bound_zL", "_bound_zR"};
      for (int piece = 0; piece < 3; ++piece) {
        const std::string path = d + "/" + n + piece_name[piece] + filen;
        TextFile in(path);
        if (!in.f) return ckref_fail("cannot open " + path);
        if (!read_values(in.f, clk, 3)) return ckref_fail(path + ": no clocks");
        const long per = piece == 0 ? L.strip_r : L.strip_z;
        const long off = piece == 0 ? 0 : (piece == 1 ? h->nf * L.strip_r : h->nf * L.strip_r + h->nf * L.strip_z);
        const int outer = piece == 0 ? g : nrt, inner = piece == 0 ? nzt : g;
        for (int a = 0; a < outer; ++a)
          for (int b = 0; b < inner; ++b) {
            double v[3];
            if (!read_values(in.f, v, 3)) return ckref_fail(path + ": fewer values than the strip holds");
            // rR(m, k): m = a, k = b + 1 - g; zL / zR(i, n): i = a + 1 - g, n = b
            const long q = piece == 0 ? static_cast<long>(a) * nzt + b : static_cast<long>(b) * nrt + a;
            for (int tl = 0; tl < 3; ++tl) strips[tl][static_cast<size_t>(off + slot * per + q)] = v[tl];
          }
      }
    }
    for (int tl = 0; tl < 3; ++tl)
      HIPCHK(hipMemcpyAsync(L.strip[L.ring[tl]], strips[tl].data(), sizeof(double) * sset, hipMemcpyHostToDevice,
                            h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    L.t = clk[0];
    L.t1 = clk[1];
    L.t2 = clk[2];
    L.s = 0;                                                         // evolve.f90: s = 0 before initial
    L.dirty = true;
    L.phi_valid = true;
    L.cons_valid = false;
    L.dtb_pre_valid = false;
  }
  return OA_OK;
}

extern "C" int oa_get_time(const oa_handle* h, int box, int level, double* t, double* t1, double* t2, int* s) {
  const Level* L = level_of(const_cast<oa_handle*>(h), box, level);
  if (!L) return OA_ERR_ARG;
  if (t) *t = L->t;
  if (t1) *t1 = L->t1;
  if (t2) *t2 = L->t2;
  if (s) *s = L->s;
  return OA_OK;
}

extern "C" int oa_set_time(oa_handle* h, int box, int level, double t, double t1, double t2, int s) {
  Level* L = level_of(h, box, level);
  if (!L) return OA_ERR_ARG;
  L->t = t; L->t1 = t1; L->t2 = t2; L->s = s;
  return OA_OK;
}

extern "C" int oa_saveold(oa_handle* h, int box, int level) {
  if (!level_of(h, box, level)) return OA_ERR_ARG;
  return do_saveold(h, level);
}

extern "C" int oa_substep(oa_handle* h, int box, int level, int iter, int niter, double dtw, double weight) {
  if (!level_of(h, box, level)) return OA_ERR_ARG;
  if (iter < 1 || iter > niter) {
    oa_set_error("oa_substep: iter must lie in 1..niter");
    return OA_ERR_ARG;
  }
  return do_substep(h, level, iter, niter, dtw, weight);
}

extern "C" int oa_finebound(oa_handle* h, int box, int level, double dtw) {
  if (!level_of(h, box, level)) return OA_ERR_ARG;
  return do_finebound(h, level, dtw);
}

extern "C" int oa_auxiliary(oa_handle* h, int box, int level) {
  if (!level_of(h, box, level)) return OA_ERR_ARG;
  return do_auxiliary(h, level);
}

extern "C" int oa_restrict(oa_handle* h, int box, int level) {
  if (!level_of(h, box, level)) return OA_ERR_ARG;
  if (level == 0) {
    oa_set_error("oa_restrict: level 0 has no coarser level");
    return OA_ERR_ARG;
  }
  return do_restrict(h, level);
}

extern "C" int oa_advance_time(oa_handle* h, int box, int level) {
  if (!level_of(h, box, level)) return OA_ERR_ARG;
  finish_level_step(h, level);   // the clocks, and the point updates as oa_coarse_step counts them
  return OA_OK;
}

extern "C" int oa_initial_aux(oa_handle* h) {
  if (!h) return OA_ERR_ARG;
  for (int l = 0; l < h->nlev; ++l) {
    h->L[l].dirty = true;
    RC(do_auxiliary(h, l));
  }
  if (h->p.shift == OA_SHIFT_GD2 && !h->p.driverD0)
    for (int l = 0; l < h->nlev; ++l) {
      Level& L = h->L[l];
      // evolve.f90:174-195 sets dtbeta = drivercsi * Delta after the auxiliary
      // pass that differentiated dtbeta: keep the planes it differentiated
      const size_t bytes = sizeof(double) * 2 * static_cast<size_t>(L.G.plane);
      HIPCHK(hipMemcpyAsync(L.dtb_pre, L.slab[L.cur] + static_cast<long>(shift_slot(SH_dtbeta_r)) * L.G.plane, bytes,
                            hipMemcpyDeviceToDevice, h->stream));
      L.dtb_pre_valid = true;
      RC(launch_dtbeta_init(L.G, h->p.drivercsi, h->scalar, L.slab[L.cur], h->stream));
    }
  HIPCHK(hipStreamSynchronize(h->stream));
  return OA_OK;
}

int refuse_member(const oa_handle* h, const char* what) {
  oa_set_error(std::string(what) + ": the spacetime is a member of a batch; step it with oa_batch_coarse_step");
  (void)h;
  return OA_ERR_STATE;
}

extern "C" int oa_coarse_step(oa_handle* h) {
  if (!h) return OA_ERR_ARG;
  if (h->batch) return refuse_member(h, "oa_coarse_step");
  HIPCHK(hipEventRecord(h->ev_begin, h->stream));
  RC(coarse_step(h));
  HIPCHK(hipEventRecord(h->ev_end, h->stream));
  HIPCHK(hipEventSynchronize(h->ev_end));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev_begin, h->ev_end));
  h->last_ms = ms;
  return OA_OK;
}

extern "C" int oa_evolve(oa_handle* h, int nsteps) {
  if (!h || nsteps < 0) return OA_ERR_ARG;
  if (h->batch) return refuse_member(h, "oa_evolve");
  for (Level& L : h->L)
    if (L.dirty) {
      oa_set_error("oa_evolve: state not projected; call oa_initial_aux after uploading");
      return OA_ERR_STATE;
    }
  h->plan.reset_timing();
  HIPCHK(hipEventRecord(h->ev_begin, h->stream));
  for (int s = 0; s < nsteps; ++s) RC(coarse_step(h));
  HIPCHK(hipEventRecord(h->ev_end, h->stream));
  HIPCHK(hipEventSynchronize(h->ev_end));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev_begin, h->ev_end));
  h->last_ms = ms;
  if (h->plan.kernel_timing && h->plan.kev_used > 0) {
    double sum = 0.0;
    RC(h->plan.rhs_stats(&sum, &h->rhs_launches, nullptr));
    h->rhs_avg_ms = sum / static_cast<double>(h->rhs_launches);
  }
  return OA_OK;
}

extern "C" int oa_sync(oa_handle* h) {
  if (!h) return OA_ERR_ARG;
  HIPCHK(hipStreamSynchronize(h->stream));
  return OA_OK;
}

extern "C" int oa_last_elapsed_ms(const oa_handle* h, double* ms) {
  if (!h || !ms) return OA_ERR_ARG;
  *ms = h->last_ms;
  return OA_OK;
}

extern "C" int oa_set_graphs(oa_handle* h, int on) {
  if (!h) return OA_ERR_ARG;
  h->plan.graphs = (on != 0);
  h->plan.tuner.enabled = false;     // the caller's choice stands
  return OA_OK;
}

extern "C" int oa_schedule(const oa_handle* h, int* graphs, int* xcd_remap, int* tuned, double* step_ms) {
  if (!h) return OA_ERR_ARG;
  const Plan& pl = h->plan;
  if (graphs) *graphs = pl.graphs ? 1 : 0;
  if (xcd_remap) *xcd_remap = pl.xcd_remap ? 1 : 0;
  if (tuned) *tuned = pl.tuner.done ? 1 : 0;
  if (step_ms)
    for (int c = 0; c < Plan::kTuneCands; ++c) step_ms[c] = pl.tuner.ms[c];
  return OA_OK;
}

extern "C" int oa_set_kernel_timing(oa_handle* h, int on) {
  if (!h) return OA_ERR_ARG;
  h->plan.kernel_timing = (on != 0);
  return OA_OK;
}

extern "C" int oa_rhs_kernel_stats(const oa_handle* h, double* avg_ms, long* launches) {
  if (!h) return OA_ERR_ARG;
  if (avg_ms) *avg_ms = h->rhs_avg_ms;
  if (launches) *launches = h->rhs_launches;
  return OA_OK;
}

extern "C" int oa_point_updates(const oa_handle* h, double* updates) {
  if (!h || !updates) return OA_ERR_ARG;
  *updates = h->point_updates;
  return OA_OK;
}

// accessors used by the host-side initial-data generator
namespace oax {
int handle_level_slab(oa_handle* h, int level, double** slab, KGrid* G, int* nf, bool* shift, bool* scalar,
                      const oa_params** p) {
  Level& L = h->L[level];
  *slab = L.slab[L.cur];
  *G = L.G;
  *nf = h->nf;
  *shift = h->shift;
  *scalar = h->scalar;
  *p = &h->p;
  return OA_OK;
}
void handle_mark_dirty(oa_handle* h, int level) {
  h->L[level].dirty = true;
  h->L[level].phi_valid = true;
}
void* handle_stream(oa_handle* h) { return h->stream; }
int handle_elliptic_exit(oa_handle* h) { return h->ell_stationary; }
void handle_elliptic_record(oa_handle* h, int steps0, int steps1, double residual) {
  h->ell_steps[0] = steps0;
  h->ell_steps[1] = steps1;
  h->ell_res = residual;
}
}  // namespace oax
