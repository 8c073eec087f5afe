// fz_launch.h -- host-side launch description shared by the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "fz_sys.h"

namespace fz {

// A static work decomposition: pre[s] = first block of system s, total blocks.
struct Work {
  int* pre = nullptr;  // device, nsys + 1 entries
  int total = 0;
};

// pair-search waves per I cluster of a system: its buffer parts, at most the launch's cap
__host__ __device__ inline int search_split(int parts, int cap) { return parts < cap ? parts : cap; }

// LDS of the PME plane kernels: two buffers of K0 K1 complex values of csz
// bytes and the x and y twiddles
inline int plane_lds_bytes(int K0, int K1, int csz) { return (2 * K0 * K1 + K0 + K1) * csz; }
// the plane kernels are used up to this many atoms per system (each of a
// system's K2 workgroups reads every atom's z position) and this much LDS
constexpr int PLANE_MAX_ATOMS = 65536, PLANE_MAX_LDS = 65536;
// launches of fewer planes than this use workgroups of 1,024 threads, others 256
// PME gather: four lanes per atom in launches of up to this many 256-atom blocks
constexpr int GATHER_QUAD_BLOCKS = 512;
constexpr int PLANE_WIDE_BELOW = 512;

struct Launch {
  hipStream_t stream = nullptr;
  const SysDev* d_sys = nullptr;  // device array of nsys descriptors
  int nsys = 0;
  int mixed = 0;                  // 0 = FP64 form, 1 = mixed form
  Work w_atoms;   // 256 atoms per block
  Work w_atoms64; // 64 atoms per block
  Work w_slots;   // 256 slots per block
  Work w_cols;    // one block per column
  Work w_clus;    // 4 clusters (waves) per block
  Work w_search;  // 4 search waves per block, search_split(parts, search_cap) waves per cluster
  Work w_terms;   // 256 bonded terms per block
  Work w_ccl;     // 128 constraint clusters per block
  Work w_step;    // 128 per block: constraint clusters, then free atoms
  Work w_fftx;    // x-line groups
  Work w_ffty;    // y-line groups
  Work w_fftz;    // z-line groups
  Work w_bricks;  // one block per PME brick
  Work w_planes;  // one block per PME grid z plane
  int* chunk_pre = nullptr;  // device: pair-kernel chunks per system (prefix)
  int* chunk_pre_a = nullptr;  // device: the same for the systems without a pending rebuild
  int pair_grid = 0;         // persistent pair-kernel blocks
  int fft_lds = 0;           // dynamic LDS bytes for the FFT kernels
  int fft_nl = 16;           // lines per FFT workgroup
  int lj_sep = 0;            // every system follows the LJ combination rule
  int com_any = 0;           // some system has a COM restraint
  int spread_bricks = 1;     // PME spread by bricks (1) or by global atomics (0)
  int pme_planes = 0;        // PME spread and x, y transforms by the plane kernels
  int plane_lds = 0;         // their dynamic LDS bytes
  int search_cap = 1;  // cap on the pair-search waves per I cluster (1, 2 or 4)
  int npt = 0;         // every system has the Andersen barostat on
};

// check: run the displacement test first (the fused step kernel does it otherwise)
void launch_nblist(const Launch& L, bool check);
void launch_gather_sorted(const Launch& L);
void launch_chunk_prefix(const Launch& L);  // pair-kernel work items of every system, rebuild flags cleared
// pair-kernel work items of the systems without (built = false, into
// chunk_pre_a) or with (built = true, into chunk_pre; flags cleared) a pending rebuild
void launch_chunk_split(const Launch& L, bool built);
// 0 forces, 1 + FP32 energies, 2 + FP64 energies; work items cp (null: chunk_pre);
// fold: the energy / virial copies are folded after it (else launch_ev_fold
// later); spend: the kernel clears a single system's rebuild decision
void launch_pair(const Launch& L, int ev, const int* cp = nullptr, bool fold = true, bool spend = false);
void launch_ev_fold(const Launch& L);
void launch_bonded(const Launch& L, bool ev);
void launch_pme_fft_only(const Launch& L, int conv_mode);
void launch_pme_gtable(const SysDev& d, hipStream_t st);  // d.pme_g from d's box tables (host descriptor)
void launch_pme_part(const Launch& L, int part, bool ev);  // 0 spread, 1 FFT + convolution, 2 gather
// kick, drift, SHAKE, rebuild test; finish: the previous step's kick + RATTLE first
void launch_step_first(const Launch& L, bool finish = false);
void launch_step_second(const Launch& L);  // kick, RATTLE
void launch_shake(const Launch& L);
void launch_rattle(const Launch& L);
void launch_kinetic(const Launch& L);
void launch_maxwell(const Launch& L, double T, unsigned long long seed, unsigned long long counter);
void launch_zero_acc(const Launch& L);
void launch_clear_shake_flag(const Launch& L);
// every member's positions, velocities, build positions and flags into its
// snapshot (clear_shake: the SHAKE-failure flag cleared first)
void launch_snapshot(const Launch& L, bool clear_shake);
void launch_zero_forces(const Launch& L);
void launch_restraint(const Launch& L);  // COM restraint forces (systems with com_on)
// Andersen barostat step phases (every system of the launch has the barostat on)
void launch_npt_a(const Launch& L);
void launch_npt_b(const Launch& L);
void launch_npt_c(const Launch& L);
void launch_npt_d(const Launch& L);
void launch_npt_intra(const Launch& L);  // the clusters' internal virial (after the forces)
void launch_npt_g(const Launch& L);
// box fields and PME tables from the state (after a restore); build_scale: at the
// box of the last list build
void launch_npt_apply(const Launch& L, bool build_scale);
void launch_npt_mark(const Launch& L);  // a list build since: build scale = current box
void launch_npt_wtr(const Launch& L);   // piston virial from a stand-alone evaluation

}  // namespace fz
