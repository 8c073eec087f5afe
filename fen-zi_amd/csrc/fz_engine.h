// fz_engine.h -- host runtime: systems resident in HBM and the engine that
// advances one or many of them with one launch per phase per step.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fenzi.h"
#include "fz_launch.h"
#include "fz_sys.h"

namespace fz {

int set_error(int code, const char* fmt, ...);

#define FZ_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::fz::set_error(FZ_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define FZ_TRY(call)             \
  do {                           \
    int r_ = (call);             \
    if (r_ != FZ_OK) return r_;  \
  } while (0)

// Owned copy of the processed topology.
struct HostTopo {
  int n = 0, ntypes = 0;
  std::vector<double> q, mass, minv;
  std::vector<int> type;
  std::vector<double> lj_eps, lj_sig2, lj14_eps, lj14_sig2;
  std::vector<int2> bond_ij; std::vector<double2> bond_p;
  std::vector<int4> angle_ijk; std::vector<double4> angle_p;
  std::vector<int4> dihe_ijkl; std::vector<double2> dihe_p; std::vector<int> dihe_n;
  std::vector<int4> impr_ijkl; std::vector<double> impr_k;
  std::vector<int2> ew_ij, p14_ij;
  std::vector<unsigned long long> excl_bits;
  int use_excl_list = 0;
  std::vector<int> excl_off, excl_list;
  std::vector<int4> ccl_atoms, ccl_cons; std::vector<double4> ccl_d2;
  std::vector<int> free_atoms;  // atoms in no constraint cluster, ascending
  int n_constraints = 0;
  std::vector<int2> cons_ab; std::vector<double> cons_len;  // the constraint list as given
  std::vector<uint8_t> fixed;  // consfix: atoms held in place
  int n_fixed = 0;
  int n_cons_active = 0;       // constraints with a mobile atom (the ones SHAKE solves)
  double q2sum = 0.0;
};

struct System {
  int device = 0;
  int mixed = 0;
  fz_params par{};
  HostTopo topo;
  SysDev dv{};
  // device allocations with their sizes: the neighbour-list block (replaced
  // whenever the column grid or the capacities change) apart from the rest
  std::vector<std::pair<void*, size_t>> allocs, nb_allocs;
  bool alloc_nb = false;  // dalloc adds to nb_allocs (inside alloc_neighbor)
  size_t bytes = 0;       // sum of the live allocations
  int64_t step = 0;
  int gen = 0;  // bumped whenever device buffers are reallocated
  double self_energy = 0.0;
  double min_step = 0.0;  // adaptive minimiser step carried between calls (0 = start at the maximum)
  bool need_rebuild = true;
  // spread path that last wrote the PME accumulation grid: 0 atomic (the grid is
  // left zero), 1 bricks (the grid holds that spread); allocated zero
  int grid_mode = 0;
  // Snapshot for restoring a call that failed or overflowed: 3n positions,
  // 3n velocities (as bits), 3n positions of the last list build, then the
  // status flags (F_NSLOT ints); step and minimiser step on the host.
  unsigned long long* snap = nullptr;
  int64_t snap_step = 0;
  double snap_min_step = 0.0;
  // host mirror of the last read accumulators
  unsigned long long acc_host[A_NSLOT] = {0};
  int flags_host[F_NSLOT] = {0};
  ~System();
  int init(const fz_topology* t, const fz_params* p, int precision, int device);
  int alloc_device();
  int alloc_neighbor(long long pairs_cap, int clusters_cap);
  int alloc_neighbor_block(long long pairs_cap, int clusters_cap);
  int set_column_grid(int ncx, int ncy, double width_factor);
  void free_device();
  void free_neighbor();
  void dfree(void* p);  // one allocation of allocs (no-op for null)
  template <class T> int dalloc(T** p, size_t count, bool zero = true);
  int upload_positions(const double* xyz);
  int upload_fixed(const uint64_t* u);
  int upload_velocities(const double* v);
  int upload_mdyn();
  int set_fixed(const uint8_t* fixed);
  int set_com_restraint(const int32_t* group, const double dist[3], double fc);
  // Andersen barostat: reference pressure (atm) and piston mass (amu/A^4);
  // mass <= 0 turns it off (the box stays where it is).  The box at this call
  // is the reference box L0 (scale 1).
  int set_barostat(double pref_atm, double mass);
  NptState npt_host{};  // host copy of the barostat state, read after every step call
  // the host descriptor's box fields (and par.box) from the state
  void npt_sync_host();
  int download_state(double* xyz, double* v, uint64_t* u);
  int energies(fz_energies* e);
};

// Launch composition over a set of systems.
class Engine {
 public:
  Engine(int device, std::vector<System*> sys);
  ~Engine();
  int prepare();                       // work tables, descriptors, streams
  // stand-alone evaluation at current positions; lite: pair energies summed in FP32
  // (the minimisers' accept/reject tests), otherwise in FP64
  int forces(bool ev, bool lite = false);
  // n MD steps (energies on the last).  On any error every member is back in
  // the state it had before the call (positions, velocities, list, counters).
  int steps(int n);
  int maxwell(double T, uint64_t seed, uint64_t counter);
  // adaptive steepest descent, velocities zeroed; restores the members on error
  int minimize(int nsteps, double max_step);
  // the reference's CG -> SD schedule; restores the members on error
  int minimize_ref(int nsteps, const fz_min_params& p);
  // Neighbour capacities of member s (pair-list entries, clusters; 0 keeps
  // one): its list is rebuilt at the positions of its last build, so the state
  // and the list it holds are unchanged.  Overflow grows them again.
  int set_capacity(int s, long long pairs_cap, int clusters_cap);
  // nsteps eager (non-graph) steps with HIP events between phases; mean ms per
  // phase per step into out[FZ_NPHASE] (order: see fz_phase_times in fenzi.h)
  int phase_times(int nsteps, double* out);
  // mean time of the force-only pair kernel over reps back-to-back launches
  int pair_kernel_time(int reps, double* ms);
  // Times graph-replayed steps over candidate column widths, FFT lines per
  // workgroup and persistent pair-kernel grids (coordinate descent), keeps the
  // fastest, and restores the systems' state.  out: {width factor, fft lines,
  // pair grid, best ms per step}.
  int autotune(int nsteps, double* out);
  // FFT lines, pair grid (0 default, < 0 keep), PME path (FZ_PME_*, < 0 keep);
  // out (nullable): the three in effect
  int set_launch(int nl, int grid, int pme_path, int32_t* out);
  int read_status();                   // flags + accumulators (+ barostat state) of every system -> host
  // barostat members: the piston's virial trace from the last stand-alone
  // evaluation (after a change of positions outside dynamics)
  int npt_reset_virial();
  int sync();
  hipStream_t stream() const { return L_.stream; }
  Launch& launch() { return L_; }
  std::vector<System*>& systems() { return sys_; }

 private:
  int device_;
  std::vector<System*> sys_;
  Launch L_;
  SysDev* d_sys_ = nullptr;
  std::vector<int*> d_pre_;
  hipGraphExec_t g_plain_ = nullptr, g_ev_ = nullptr, g_npt_ = nullptr;
  // The same steps captured GRAPH_STEPS at a time: the host cost of launching a
  // graph whose streams fork and join is a fixed ~60 us plus ~3 us per node, about
  // as much as a small system's whole step, so steps are launched in blocks.
  static constexpr int GRAPH_STEPS = 8;
  hipGraphExec_t g_plain_k_ = nullptr, g_npt_k_ = nullptr;
  // The force evaluation forks before the list build (ev_early_): the build
  // and the pair kernel on the main stream, bonded terms and the PME chain on
  // two side streams (default priority), joined before the second half kick.
  // The side branches write atom-ordered force sums and the pair kernel
  // slot-ordered ones; energies and virials go into fixed-point (mixed) or
  // FP64 accumulators by atomic adds, so the branches need no other ordering.
  // In dynamics a third side stream runs the pair kernel over the systems
  // without a pending rebuild (ev_split_: their work items are laid out).
  static constexpr int NSIDE = 3;
  hipStream_t side_[NSIDE] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_early_ = nullptr, ev_split_ = nullptr, ev_join_[NSIDE] = {nullptr, nullptr, nullptr};
  bool prepared_ = false;
  std::vector<int> gens_;
  int tuned_nl_ = 0, tuned_grid_ = 0;  // autotuned launch shapes (0 = defaults)
  int tuned_planes_ = -1;              // PME plane kernels: -1 by size, 0 off, 1 on
  int ensure();  // (re)prepare when a member's buffers moved since the last prepare
  int time_steps(int nsteps, double* ms);
  int make_work(Work& w, const std::vector<int>& blocks);
  int capture(bool ev, hipGraphExec_t* out, int nsteps = 1);
  void record_step(bool ev);
  void record_npt_step();
  int steps_npt(int n);
  int capture_npt(hipGraphExec_t* out, int nsteps = 1);
  void record_forces(bool ev, bool lite = false, bool check = true);
  int grow_and_reset(int s, int bits);
  int realloc_neighbor(int s, long long pairs, int clus);
  // device copies of every member's positions, velocities, build positions and
  // flags (stream-ordered), host counters
  int snapshot(bool clear_shake = false);
  // every member back to its snapshot; grow[k] != 0 enlarges member k's
  // neighbour capacity (overflow bits) first
  int restore(const std::vector<int>* grow = nullptr);
  int minimize_sd(int nsteps, double max_step);
  int minimize_cgsd(int nsteps, const fz_min_params& p);
  void drop_graphs();
  int upload_desc();
};

// Column grid and initial neighbour capacities for n atoms under par
// (explicit ncx, ncy > 0 kept); width factor f scales the column width.
struct NbShape {
  int ncx, ncy, ncol, nzb, clusters_cap;
  long long pairs_cap;
};
NbShape nb_shape(const fz_params& par, int n, int ncx, int ncy, double f);
// device bytes of one neighbour block
size_t nb_bytes(int ncol, int nzb, int mixed, long long pairs_cap, int clusters_cap);
// upper bound of one system's device bytes (the serving admission figure)
int64_t footprint_bytes(const fz_topology* t, const fz_params* p, int precision);

int pme_convolve(int device, const int32_t K[3], const double box[3], double beta, int mixed, int mode,
                 const double* q, double* phi, double* energy);

// The convolution over buffers carved out of a file-scope device array
// (fz_noalloc.hip): noalloc_init takes that array's device address on
// `device` (once per device, that device current); pme_convolve_noalloc
// enqueues the whole call on `stream` (caller's device grids q, phi and the
// nullable device scalar energy) and allocates nothing.
int noalloc_init(int device);
size_t noalloc_bytes();
int pme_convolve_noalloc(int device, const int32_t K[3], const double box[3], double beta, int mixed, int mode,
                         const double* q, double* phi, double* energy, hipStream_t stream);

}  // namespace fz

// device context handle of the C ABI
struct fz_ctx {
  int device;
};
