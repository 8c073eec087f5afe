// fz_capi.cpp -- extern "C" entry points of libfenzi (declared in include/fenzi.h).
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "fz_engine.h"

using namespace fz;

namespace fz {
const char* last_error();
}

struct fz_system {
  fz_ctx* ctx;
  System sys;
  std::unique_ptr<Engine> eng;
};

struct fz_batch {
  fz_ctx* ctx;
  std::vector<fz_system*> members;
  std::unique_ptr<Engine> eng;
};

#define FZ_GUARD_BEGIN try {
#define FZ_GUARD_END                                                                 \
  }                                                                                  \
  catch (const std::bad_alloc&) {                                                    \
    return set_error(FZ_ERR_NOMEM, "host allocation failed");                        \
  }                                                                                  \
  catch (...) {                                                                      \
    return set_error(FZ_ERR_ARG, "unexpected exception");                            \
  }

extern "C" {

const char* fz_last_error(void) { return fz::last_error(); }

int fz_ctx_create(int device, fz_ctx** out) {
  if (!out) return set_error(FZ_ERR_ARG, "null output");
  int n = 0;
  FZ_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_error(FZ_ERR_ARG, "device %d not present (%d devices)", device, n);
  FZ_HIP(hipSetDevice(device));
  FZ_TRY(noalloc_init(device));
  *out = new fz_ctx{device};
  return FZ_OK;
}

void fz_ctx_destroy(fz_ctx* ctx) { delete ctx; }

int fz_system_create(fz_ctx* ctx, const fz_topology* topo, const fz_params* par, int precision, fz_system** out) {
  FZ_GUARD_BEGIN
  if (!ctx || !topo || !par || !out) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(ctx->device));
  auto* s = new fz_system();
  s->ctx = ctx;
  int r = s->sys.init(topo, par, precision, ctx->device);
  if (r != FZ_OK) { delete s; return r; }
  s->eng.reset(new Engine(ctx->device, {&s->sys}));
  r = s->eng->prepare();
  if (r != FZ_OK) { delete s; return r; }
  *out = s;
  return FZ_OK;
  FZ_GUARD_END
}

void fz_system_destroy(fz_system* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  s->eng.reset();
  delete s;
}

int fz_set_state(fz_system* s, const double* xyz, const double* vel) {
  FZ_GUARD_BEGIN
  if (!s || !xyz) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->sys.upload_positions(xyz));
  FZ_TRY(s->sys.upload_velocities(vel));
  FZ_TRY(s->eng->forces(true));
  return s->eng->npt_reset_virial();
  FZ_GUARD_END
}

int fz_get_state(fz_system* s, double* xyz, double* vel) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->eng->sync());
  return s->sys.download_state(xyz, vel, nullptr);
  FZ_GUARD_END
}

int fz_set_fixed_positions(fz_system* s, const uint64_t* u) {
  FZ_GUARD_BEGIN
  if (!s || !u) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->sys.upload_fixed(u));
  FZ_TRY(s->eng->forces(true));
  return s->eng->npt_reset_virial();
  FZ_GUARD_END
}

int fz_set_barostat(fz_system* s, double pref_atm, double piston_mass) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->eng->sync());
  FZ_TRY(s->sys.set_barostat(pref_atm, piston_mass));
  FZ_TRY(s->eng->forces(true));
  return s->eng->npt_reset_virial();
  FZ_GUARD_END
}

int fz_get_barostat(fz_system* s, double* box, double* state) {
  FZ_GUARD_BEGIN
  if (!s || !box) return set_error(FZ_ERR_ARG, "null argument");
  for (int k = 0; k < 3; k++) box[k] = s->sys.dv.L[k];
  if (state) {
    const NptState& st = s->sys.npt_host;
    const bool on = s->sys.dv.npt != 0;
    state[0] = s->sys.dv.L[0] * s->sys.dv.L[1] * s->sys.dv.L[2];
    state[1] = on ? st.Pi : 0.0;
    state[2] = on ? st.Wtr : 0.0;
  }
  return FZ_OK;
  FZ_GUARD_END
}

int fz_get_fixed_positions(fz_system* s, uint64_t* u) {
  FZ_GUARD_BEGIN
  if (!s || !u) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->eng->sync());
  return s->sys.download_state(nullptr, nullptr, u);
  FZ_GUARD_END
}

int fz_reassign_velocities(fz_system* s, double T, uint64_t seed, uint64_t counter) {
  FZ_GUARD_BEGIN
  if (!s || !(T >= 0)) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->maxwell(T, seed, counter);
  FZ_GUARD_END
}

int fz_set_fixed_atoms(fz_system* s, const uint8_t* fixed) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->sys.set_fixed(fixed);
  FZ_GUARD_END
}

int fz_set_com_restraint(fz_system* s, const int32_t* group, const double dist[3], double fc) {
  FZ_GUARD_BEGIN
  if (!s || (group && !dist)) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->sys.set_com_restraint(group, dist, fc);
  FZ_GUARD_END
}

int fz_compute_forces(fz_system* s, double* f, fz_energies* e) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->eng->forces(true));
  if (f) {
    const SysDev& d = s->sys.dv;
    const int n = d.n_atoms, ns = d.n_slots_cap;
    std::vector<int> slot(n);
    std::vector<unsigned long long> F((size_t)3 * ns), B((size_t)3 * n), G((size_t)3 * n);
    FZ_HIP(hipMemcpy(slot.data(), d.atom_slot, n * sizeof(int), hipMemcpyDeviceToHost));
    FZ_HIP(hipMemcpy(F.data(), d.Fs, F.size() * 8, hipMemcpyDeviceToHost));
    FZ_HIP(hipMemcpy(B.data(), d.Fa, B.size() * 8, hipMemcpyDeviceToHost));
    FZ_HIP(hipMemcpy(G.data(), d.Fg, G.size() * 8, hipMemcpyDeviceToHost));
    // pair (slot order) + bonded + gather (atom order), as the integrator adds them
    for (int i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) {
        const unsigned long long p = F[(size_t)slot[i] * 3 + k], b = B[(size_t)k * n + i], g = G[(size_t)k * n + i];
        double x;
        if (s->sys.mixed) {
          x = (double)(long long)(p + b + g) / 4294967296.0;
        } else {
          double xp, xb, xg;
          memcpy(&xp, &p, 8); memcpy(&xb, &b, 8); memcpy(&xg, &g, 8);
          x = (xp + xb) + xg;
        }
        f[3 * i + k] = x;
      }
  }
  if (e) FZ_TRY(s->sys.energies(e));
  return FZ_OK;
  FZ_GUARD_END
}

int fz_step(fz_system* s, int nsteps) {
  FZ_GUARD_BEGIN
  if (!s || nsteps < 0) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->steps(nsteps);
  FZ_GUARD_END
}

int fz_set_neighbor_capacity(fz_system* s, int64_t pairs_cap, int32_t clusters_cap) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->set_capacity(0, pairs_cap, clusters_cap);
  FZ_GUARD_END
}

int fz_phase_times(fz_system* s, int nsteps, double* ms) {
  FZ_GUARD_BEGIN
  if (!s || nsteps <= 0 || !ms) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->phase_times(nsteps, ms);
  FZ_GUARD_END
}

int fz_pair_kernel_time(fz_system* s, int reps, double* ms) {
  FZ_GUARD_BEGIN
  if (!s || reps <= 0 || !ms) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->pair_kernel_time(reps, ms);
  FZ_GUARD_END
}

int fz_autotune(fz_system* s, int nsteps, double* out) {
  FZ_GUARD_BEGIN
  if (!s || nsteps <= 0 || !out) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->autotune(nsteps, out);
  FZ_GUARD_END
}

int fz_set_launch(fz_system* s, int nl, int grid, int pme_path, int32_t* out) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->set_launch(nl, grid, pme_path, out);
  FZ_GUARD_END
}

int fz_batch_set_launch(fz_batch* b, int nl, int grid, int pme_path, int32_t* out) {
  FZ_GUARD_BEGIN
  if (!b) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(b->ctx->device));
  return b->eng->set_launch(nl, grid, pme_path, out);
  FZ_GUARD_END
}

int fz_batch_phase_times(fz_batch* b, int nsteps, double* ms) {
  FZ_GUARD_BEGIN
  if (!b || nsteps <= 0 || !ms) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(b->ctx->device));
  return b->eng->phase_times(nsteps, ms);
  FZ_GUARD_END
}

int fz_batch_autotune(fz_batch* b, int nsteps, double* out) {
  FZ_GUARD_BEGIN
  if (!b || nsteps <= 0 || !out) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(b->ctx->device));
  return b->eng->autotune(nsteps, out);
  FZ_GUARD_END
}

int fz_get_energies(fz_system* s, fz_energies* e) {
  if (!s || !e) return set_error(FZ_ERR_ARG, "null argument");
  return s->sys.energies(e);
}

int fz_system_info(fz_system* s, int64_t* n_atoms, int64_t* device_bytes, int32_t* precision) {
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  if (n_atoms) *n_atoms = s->sys.topo.n;
  if (device_bytes) *device_bytes = (int64_t)s->sys.bytes;
  if (precision) *precision = s->sys.mixed ? FZ_MIXED : FZ_FP64;
  return FZ_OK;
}

int fz_get_cluster_layout(fz_system* s, int32_t* ncx_ncy, int32_t* atom_column, int32_t* slot_atom,
                          int64_t slot_cap, int64_t* n_slots) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->eng->sync());
  const SysDev& d = s->sys.dv;
  int ncl = 0;
  FZ_HIP(hipMemcpy(&ncl, d.n_clusters, sizeof(int), hipMemcpyDeviceToHost));
  if (ncx_ncy) { ncx_ncy[0] = d.ncx; ncx_ncy[1] = d.ncy; }
  if (n_slots) *n_slots = (int64_t)ncl * CLUSTER;
  if (atom_column) FZ_HIP(hipMemcpy(atom_column, d.atom_col, d.n_atoms * sizeof(int), hipMemcpyDeviceToHost));
  if (slot_atom) {
    if (slot_cap < (int64_t)ncl * CLUSTER) return set_error(FZ_ERR_ARG, "slot_cap too small (%lld needed)", (long long)ncl * CLUSTER);
    FZ_HIP(hipMemcpy(slot_atom, d.slot_atom, (size_t)ncl * CLUSTER * sizeof(int), hipMemcpyDeviceToHost));
  }
  return FZ_OK;
  FZ_GUARD_END
}

int64_t fz_get_cluster_pairs(fz_system* s, int32_t* pairs, uint64_t* masks, int64_t cap) {
  FZ_GUARD_BEGIN
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  FZ_TRY(s->eng->sync());
  const SysDev& d = s->sys.dv;
  long long np = 0;
  FZ_HIP(hipMemcpy(&np, d.n_pairs, sizeof(long long), hipMemcpyDeviceToHost));
  if (!pairs || !masks || cap < np) return np;
  std::vector<int> I(np), J(np);
  FZ_HIP(hipMemcpy(I.data(), d.pl_i, np * sizeof(int), hipMemcpyDeviceToHost));
  FZ_HIP(hipMemcpy(J.data(), d.pl_j, np * sizeof(int), hipMemcpyDeviceToHost));
  FZ_HIP(hipMemcpy(masks, d.pl_mask, np * sizeof(uint64_t), hipMemcpyDeviceToHost));
  for (long long k = 0; k < np; k++) { pairs[2 * k] = I[k]; pairs[2 * k + 1] = J[k]; }
  return np;
  FZ_GUARD_END
}

int fz_pme_convolve(fz_ctx* ctx, const int32_t K[3], const double box[3], double beta, int precision, int mode,
                    const double* q, double* phi, double* energy) {
  FZ_GUARD_BEGIN
  if (!ctx || !K || !box || !q || !phi) return set_error(FZ_ERR_ARG, "null argument");
  return pme_convolve(ctx->device, K, box, beta, precision == FZ_MIXED ? 1 : 0, mode, q, phi, energy);
  FZ_GUARD_END
}

int fz_pme_convolve_noalloc(fz_ctx* ctx, const int32_t K[3], const double box[3], double beta, int precision,
                            int mode, const double* q_dev, double* phi_dev, double* energy_dev, void* stream) {
  FZ_GUARD_BEGIN
  if (!ctx || !K || !box || !q_dev || !phi_dev) return set_error(FZ_ERR_ARG, "null argument");
  return pme_convolve_noalloc(ctx->device, K, box, beta, precision == FZ_MIXED ? 1 : 0, mode, q_dev, phi_dev,
                              energy_dev, (hipStream_t)stream);
  FZ_GUARD_END
}

int64_t fz_pme_convolve_noalloc_bytes(void) { return (int64_t)noalloc_bytes(); }

int fz_batch_create(fz_ctx* ctx, fz_system* const* systems, int n, fz_batch** out) {
  FZ_GUARD_BEGIN
  if (!ctx || !systems || n <= 0 || !out) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(ctx->device));
  auto* b = new fz_batch();
  b->ctx = ctx;
  std::vector<System*> ss;
  for (int i = 0; i < n; i++) {
    if (!systems[i] || systems[i]->ctx->device != ctx->device) { delete b; return set_error(FZ_ERR_ARG, "system %d invalid", i); }
    b->members.push_back(systems[i]);
    ss.push_back(&systems[i]->sys);
  }
  b->eng.reset(new Engine(ctx->device, ss));
  int r = b->eng->prepare();
  if (r != FZ_OK) { delete b; return r; }
  *out = b;
  return FZ_OK;
  FZ_GUARD_END
}

int fz_batch_step(fz_batch* b, int nsteps) {
  FZ_GUARD_BEGIN
  if (!b || nsteps < 0) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(b->ctx->device));
  return b->eng->steps(nsteps);
  FZ_GUARD_END
}

void fz_batch_destroy(fz_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  b->eng.reset();
  delete b;
}

int fz_minimize(fz_system* s, int nsteps, double max_step) {
  FZ_GUARD_BEGIN
  if (!s || nsteps < 0 || !(max_step > 0)) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->minimize(nsteps, max_step);
  FZ_GUARD_END
}

int fz_batch_minimize(fz_batch* b, int nsteps, double max_step) {
  FZ_GUARD_BEGIN
  if (!b || nsteps < 0 || !(max_step > 0)) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(b->ctx->device));
  return b->eng->minimize(nsteps, max_step);
  FZ_GUARD_END
}

static bool min_params_ok(const fz_min_params* p) {
  if (!p) return false;
  if (p->method == 0) return p->max_step > 0;
  return (p->method == 1 || p->method == 2) && p->cgfac >= 0 && p->sdfac >= 0 && p->cgrate >= 0 && p->sdrate >= 0;
}

int fz_minimize_ex(fz_system* s, int nsteps, const fz_min_params* p) {
  FZ_GUARD_BEGIN
  if (!s || nsteps < 0 || !min_params_ok(p)) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(s->ctx->device));
  return s->eng->minimize_ref(nsteps, *p);
  FZ_GUARD_END
}

int fz_batch_minimize_ex(fz_batch* b, int nsteps, const fz_min_params* p) {
  FZ_GUARD_BEGIN
  if (!b || nsteps < 0 || !min_params_ok(p)) return set_error(FZ_ERR_ARG, "invalid argument");
  FZ_HIP(hipSetDevice(b->ctx->device));
  return b->eng->minimize_ref(nsteps, *p);
  FZ_GUARD_END
}

}  // extern "C"
