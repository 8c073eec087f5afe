// fz_server.cpp -- continuous-batching MD job server over device-resident systems.
//
// The reference runs one system per process (main, src/fenzi.cu:112-630).
// Here many jobs share one GPU: each job is a System kept in HBM from
// admission to completion; every poll round advances all minimising jobs as
// one batch (one launch per phase over all of them) and all dynamics jobs as
// another, retires finished jobs (freeing their memory) and admits queued
// jobs while the estimated footprint fits the HBM budget and a slot is free.
// A job's arithmetic does not depend on which other jobs share its batch
// (per-system work items, order-independent fixed-point sums), so a job's
// result is the same alone and in the mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "fz_engine.h"

using namespace fz;

namespace {

struct OwnedTopo {
  std::vector<double> charge, mass, lj_eps, lj_sig2, lj14_eps, lj14_sig2, bond_k, bond_r0, angle_k, angle_theta0, ub_k,
      ub_s0, dihedral_k, dihedral_delta, improper_k, improper_psi0, constraint_length;
  std::vector<int32_t> type, bond_atoms, angle_atoms, dihedral_atoms, dihedral_n, improper_atoms, constraint_atoms;
  fz_topology t{};

  template <class T>
  static void cp(std::vector<T>& v, const T* p, size_t n) { v.assign(p ? p : nullptr, p ? p + n : nullptr); if (v.empty()) v.push_back(T()); }

  void copy(const fz_topology& s) {
    const size_t n = s.n_atoms, T2 = (size_t)s.n_types * s.n_types;
    cp(charge, s.charge, n); cp(mass, s.mass, n); cp(type, s.type, n);
    cp(lj_eps, s.lj_eps, T2); cp(lj_sig2, s.lj_sig2, T2); cp(lj14_eps, s.lj14_eps, T2); cp(lj14_sig2, s.lj14_sig2, T2);
    cp(bond_atoms, s.bond_atoms, 2 * (size_t)s.n_bonds); cp(bond_k, s.bond_k, s.n_bonds); cp(bond_r0, s.bond_r0, s.n_bonds);
    cp(angle_atoms, s.angle_atoms, 3 * (size_t)s.n_angles); cp(angle_k, s.angle_k, s.n_angles);
    cp(angle_theta0, s.angle_theta0, s.n_angles); cp(ub_k, s.ub_k, s.n_angles); cp(ub_s0, s.ub_s0, s.n_angles);
    cp(dihedral_atoms, s.dihedral_atoms, 4 * (size_t)s.n_dihedrals); cp(dihedral_k, s.dihedral_k, s.n_dihedrals);
    cp(dihedral_n, s.dihedral_n, s.n_dihedrals); cp(dihedral_delta, s.dihedral_delta, s.n_dihedrals);
    cp(improper_atoms, s.improper_atoms, 4 * (size_t)s.n_impropers); cp(improper_k, s.improper_k, s.n_impropers);
    cp(improper_psi0, s.improper_psi0, s.n_impropers);
    cp(constraint_atoms, s.constraint_atoms, 2 * (size_t)s.n_constraints);
    cp(constraint_length, s.constraint_length, s.n_constraints);
    t = s;
    t.charge = charge.data(); t.mass = mass.data(); t.type = type.data();
    t.lj_eps = lj_eps.data(); t.lj_sig2 = lj_sig2.data(); t.lj14_eps = lj14_eps.data(); t.lj14_sig2 = lj14_sig2.data();
    t.bond_atoms = bond_atoms.data(); t.bond_k = bond_k.data(); t.bond_r0 = bond_r0.data();
    t.angle_atoms = angle_atoms.data(); t.angle_k = angle_k.data(); t.angle_theta0 = angle_theta0.data();
    t.ub_k = ub_k.data(); t.ub_s0 = ub_s0.data();
    t.dihedral_atoms = dihedral_atoms.data(); t.dihedral_k = dihedral_k.data(); t.dihedral_n = dihedral_n.data();
    t.dihedral_delta = dihedral_delta.data();
    t.improper_atoms = improper_atoms.data(); t.improper_k = improper_k.data(); t.improper_psi0 = improper_psi0.data();
    t.constraint_atoms = constraint_atoms.data(); t.constraint_length = constraint_length.data();
  }
};

struct Job {
  int64_t id = 0;
  int state = FZ_JOB_QUEUED;
  OwnedTopo topo;
  fz_params par{};
  std::vector<double> xyz;
  int min_steps = 0, md_steps = 0, reassign_every = 0;
  double T = 300.0;
  uint64_t seed = 0, counter = 0;
  int min_done = 0, md_done = 0;
  int64_t est_bytes = 0, bytes = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  std::unique_ptr<System> sys;
  std::unique_ptr<Engine> solo;
  std::vector<double> rx, rv;
  std::vector<uint64_t> ru;
  fz_energies e{};
};

int64_t estimate_bytes(const fz_topology* t, const fz_params* p, int precision) {
  return footprint_bytes(t, p, precision);
}

}  // namespace

struct fz_server {
  int device = 0;
  fz_server_config cfg{};
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::deque<int64_t> queue;
  int64_t next_id = 1;
  int64_t used_bytes = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::unique_ptr<Engine> min_eng, md_eng;
  std::vector<int64_t> min_members, md_members;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

  void finish(Job& j, int state) {
    j.state = state;
    j.t_done = now();
    if (j.sys && state == FZ_JOB_DONE) {
      const size_t n = (size_t)j.sys->topo.n;
      j.rx.resize(3 * n); j.rv.resize(3 * n); j.ru.resize(3 * n);
      j.sys->download_state(j.rx.data(), j.rv.data(), j.ru.data());
      j.sys->energies(&j.e);
      j.e.step = j.md_done;
    }
    used_bytes -= j.est_bytes;
    j.solo.reset();
    j.sys.reset();
  }

  int admit() {
    while (!queue.empty()) {
      Job& j = *jobs[queue.front()];
      int active = 0;
      for (auto& kv : jobs)
        if (kv.second->state == FZ_JOB_MINIMIZING || kv.second->state == FZ_JOB_DYNAMICS) active++;
      if (active >= cfg.max_active) break;
      if (active > 0 && used_bytes + j.est_bytes > cfg.hbm_budget_bytes) break;
      queue.pop_front();
      j.sys.reset(new System());
      int r = j.sys->init(&j.topo.t, &j.par, cfg.precision, device);
      if (r == FZ_OK) r = j.sys->upload_positions(j.xyz.data());
      if (r == FZ_OK) r = j.sys->upload_velocities(nullptr);
      if (r == FZ_OK) {
        j.solo.reset(new Engine(device, {j.sys.get()}));
        r = j.solo->prepare();
      }
      used_bytes += j.est_bytes;
      j.t_admit = now();
      if (r != FZ_OK) { finish(j, FZ_JOB_FAILED); continue; }
      j.bytes = (int64_t)j.sys->bytes;
      j.xyz.clear();
      j.xyz.shrink_to_fit();
      if (j.min_steps > 0) j.state = FZ_JOB_MINIMIZING;
      else if (start_dynamics(j) != FZ_OK) finish(j, FZ_JOB_FAILED);
    }
    return FZ_OK;
  }

  int start_dynamics(Job& j) {
    j.state = FZ_JOB_DYNAMICS;
    if (j.md_steps == 0) { finish(j, FZ_JOB_DONE); return FZ_OK; }
    return j.solo->maxwell(j.T, j.seed, j.counter++);
  }

  // engine over the jobs in state st (rebuilt when the membership changes)
  Engine* group(int st, std::unique_ptr<Engine>& eng, std::vector<int64_t>& members) {
    std::vector<int64_t> ids;
    std::vector<System*> ss;
    for (auto& kv : jobs)
      if (kv.second->state == st) { ids.push_back(kv.first); ss.push_back(kv.second->sys.get()); }
    if (ids.empty()) { eng.reset(); members.clear(); return nullptr; }
    if (ids != members || !eng) {
      eng.reset(new Engine(device, ss));
      members = ids;
      if (eng->prepare() != FZ_OK) { eng.reset(); members.clear(); return nullptr; }
    }
    return eng.get();
  }

  // a failing batch call is isolated by retrying its members one at a time
  template <class F>
  void run_isolated(std::vector<int64_t> ids, F&& call_one) {
    for (int64_t id : ids) {
      Job& j = *jobs[id];
      if (call_one(j) != FZ_OK) finish(j, FZ_JOB_FAILED);
    }
  }

  int poll() {
    FZ_HIP(hipSetDevice(device));
    admit();
    const int chunk = std::max(1, cfg.chunk_steps);
    // The minimisation round (a host loop that reads each iteration's energies)
    // runs on a second host thread while this one runs the dynamics round: the
    // two batches have their own engines and streams, so they overlap on the
    // device.  Job bookkeeping happens afterwards, here.  Jobs that finish
    // minimising join the dynamics batch in the next round.
    Engine* me = group(FZ_JOB_MINIMIZING, min_eng, min_members);
    Engine* de = group(FZ_JOB_DYNAMICS, md_eng, md_members);
    int n_min = chunk, n_md = chunk;
    for (int64_t id : min_members) n_min = std::min(n_min, jobs[id]->min_steps - jobs[id]->min_done);
    // the dynamics round ends at the first job's completion or velocity
    // reassignment, so every job sees its own schedule exactly
    for (int64_t id : md_members) {
      Job& j = *jobs[id];
      n_md = std::min(n_md, j.md_steps - j.md_done);
      if (j.reassign_every > 0) n_md = std::min(n_md, j.reassign_every - j.md_done % j.reassign_every);
    }
    const std::vector<int64_t> min_ids = min_members, md_ids = md_members;
    int rc_min = FZ_OK, rc_md = FZ_OK;
    std::thread th;
    if (me)
      th = std::thread([&] {
        if (hipSetDevice(device) != hipSuccess) { rc_min = FZ_ERR_HIP; return; }
        rc_min = me->minimize(n_min, cfg.min_max_step);
      });
    if (de) rc_md = de->steps(n_md);
    if (th.joinable()) th.join();
    if (me) {
      if (rc_min != FZ_OK) {
        min_eng.reset(); min_members.clear();
        run_isolated(min_ids, [&](Job& j) { return j.solo->minimize(n_min, cfg.min_max_step); });
      }
      for (int64_t id : min_ids) {
        Job& j = *jobs[id];
        if (j.state != FZ_JOB_MINIMIZING) continue;
        j.min_done += n_min;
        if (j.min_done >= j.min_steps && start_dynamics(j) != FZ_OK) finish(j, FZ_JOB_FAILED);
      }
    }
    if (de) {
      const int n = n_md;
      std::vector<int64_t> ids = md_ids;
      if (rc_md != FZ_OK) {
        md_eng.reset(); md_members.clear();
        run_isolated(ids, [&](Job& j) { return j.solo->steps(n); });
      }
      for (int64_t id : ids) {
        Job& j = *jobs[id];
        if (j.state != FZ_JOB_DYNAMICS) continue;
        j.md_done += n;
        if (j.md_done >= j.md_steps) finish(j, FZ_JOB_DONE);
        else if (j.reassign_every > 0 && j.md_done % j.reassign_every == 0 &&
                 j.solo->maxwell(j.T, j.seed, j.counter++) != FZ_OK)
          finish(j, FZ_JOB_FAILED);
      }
    }
    admit();
    return FZ_OK;
  }
};

extern "C" {

int64_t fz_server_estimate_bytes(const fz_topology* topo, const fz_params* par, int32_t precision) {
  if (!topo || !par) return set_error(FZ_ERR_ARG, "null argument");
  return estimate_bytes(topo, par, precision);
}

int fz_server_create(fz_ctx* ctx, const fz_server_config* cfg, fz_server** out) {
  if (!ctx || !cfg || !out) return set_error(FZ_ERR_ARG, "null argument");
  if (!(cfg->hbm_budget_bytes > 0) || cfg->max_active <= 0) return set_error(FZ_ERR_ARG, "invalid server configuration");
  auto* s = new fz_server();
  s->device = ctx->device;
  s->cfg = *cfg;
  if (s->cfg.chunk_steps <= 0) s->cfg.chunk_steps = 100;
  if (!(s->cfg.min_max_step > 0)) s->cfg.min_max_step = 0.05;
  *out = s;
  return FZ_OK;
}

int fz_server_submit(fz_server* s, const fz_topology* topo, const fz_params* par, const double* xyz, int32_t min_steps,
                     int32_t md_steps, double temperature, int32_t reassign_every, uint64_t seed, int64_t* job_id) {
  try {
    if (!s || !topo || !par || !xyz || min_steps < 0 || md_steps < 0) return set_error(FZ_ERR_ARG, "invalid argument");
    std::unique_ptr<Job> j(new Job());
    j->id = s->next_id++;
    j->topo.copy(*topo);
    j->par = *par;
    j->xyz.assign(xyz, xyz + 3 * (size_t)topo->n_atoms);
    j->min_steps = min_steps;
    j->md_steps = md_steps;
    j->reassign_every = reassign_every;
    j->T = temperature;
    j->seed = seed;
    j->est_bytes = estimate_bytes(topo, par, s->cfg.precision);
    j->t_submit = s->now();
    if (j->est_bytes > s->cfg.hbm_budget_bytes)
      return set_error(FZ_ERR_BUDGET, "job needs about %lld bytes, more than the budget", (long long)j->est_bytes);
    if (job_id) *job_id = j->id;
    s->queue.push_back(j->id);
    s->jobs[j->id] = std::move(j);
    return FZ_OK;
  } catch (...) {
    return set_error(FZ_ERR_NOMEM, "host allocation failed");
  }
}

int fz_server_poll(fz_server* s, int32_t* n_active, int32_t* n_queued, int32_t* n_done) {
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  int r;
  try {
    r = s->poll();
  } catch (...) {
    return set_error(FZ_ERR_NOMEM, "host allocation failed");
  }
  int a = 0, q = 0, d = 0;
  for (auto& kv : s->jobs) {
    int st = kv.second->state;
    if (st == FZ_JOB_QUEUED) q++;
    else if (st == FZ_JOB_MINIMIZING || st == FZ_JOB_DYNAMICS) a++;
    else d++;
  }
  if (n_active) *n_active = a;
  if (n_queued) *n_queued = q;
  if (n_done) *n_done = d;
  return r;
}

int fz_server_job_info(fz_server* s, int64_t id, fz_job_info* info) {
  if (!s || !info) return set_error(FZ_ERR_ARG, "null argument");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return set_error(FZ_ERR_ARG, "unknown job %lld", (long long)id);
  const Job& j = *it->second;
  info->id = j.id;
  info->state = j.state;
  info->n_atoms = j.topo.t.n_atoms;
  info->steps_done = j.md_done;
  info->min_done = j.min_done;
  info->device_bytes = j.bytes;
  info->t_submit = j.t_submit;
  info->t_admit = j.t_admit;
  info->t_done = j.t_done;
  return FZ_OK;
}

int fz_server_job_result(fz_server* s, int64_t id, double* xyz, double* vel, uint64_t* fixed, fz_energies* e) {
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return set_error(FZ_ERR_ARG, "unknown job %lld", (long long)id);
  const Job& j = *it->second;
  if (j.state != FZ_JOB_DONE) return set_error(FZ_ERR_STATE, "job %lld is not finished", (long long)id);
  if (xyz) std::copy(j.rx.begin(), j.rx.end(), xyz);
  if (vel) std::copy(j.rv.begin(), j.rv.end(), vel);
  if (fixed) std::copy(j.ru.begin(), j.ru.end(), fixed);
  if (e) *e = j.e;
  return FZ_OK;
}

int fz_server_release(fz_server* s, int64_t id) {
  if (!s) return set_error(FZ_ERR_ARG, "null argument");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return set_error(FZ_ERR_ARG, "unknown job %lld", (long long)id);
  if (it->second->state != FZ_JOB_DONE && it->second->state != FZ_JOB_FAILED)
    return set_error(FZ_ERR_STATE, "job %lld is still running", (long long)id);
  s->jobs.erase(it);
  return FZ_OK;
}

void fz_server_destroy(fz_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  s->min_eng.reset();
  s->md_eng.reset();
  delete s;
}

}  // extern "C"
