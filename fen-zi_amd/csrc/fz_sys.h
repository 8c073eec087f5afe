// fz_sys.h -- per-system device descriptor shared by the host engine and the
// gfx950 kernels.  One SysDev describes one periodic system resident in HBM;
// a batch is an array of them plus per-launch block-prefix tables, so every
// kernel runs unchanged on one system or on many.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fz {

constexpr int CLUSTER = 8;          // atoms per cluster (8 x 8 pair tile = one wave64)
constexpr int PAIR_CHUNK = 32;      // cluster-pair entries per pair-kernel work item
constexpr int SEARCH_TMP_CAP = 512;  // max J clusters per I cluster in the search buffer
constexpr int MAX_COLUMN_ATOMS = 2048; // LDS sort capacity per column
constexpr int MAX_FFT = 256;        // largest PME grid dimension

// Energy / virial accumulator slots (int64 fixed point, scale 2^32).
enum AccSlot {
  A_BOND = 0, A_ANGLE, A_UB, A_IMPR, A_DIHED, A_VDW, A_ELEC, A_PME, A_EWEXCL,
  A_KIN, A_VIR /* 9 slots */, A_NPAIR = A_VIR + 9, A_MOM /* 3 slots */, A_FMAX = A_MOM + 3,
  A_COM /* 6 slots: sum m x of group 0, then group 1 */, A_NTILE = A_COM + 6, A_NSLOT
};

// The pair kernels' energy / virial sums of one evaluation go first to
// EV_SPREAD copies of EV_W words (one 128-byte line each; a wave adds into copy
// wave & (EV_SPREAD - 1)), and a fold kernel adds the copies into acc: several
// thousand waves adding into the few lines of acc serialise there.  Words: vdw,
// elec, virial xx yy zz xy xz yz, pair count, tile count.
constexpr int EV_SPREAD = 64, EV_W = 16, EV_NV = 10;

// Device-side status flags.
enum FlagSlot { F_REBUILD = 0, F_OVERFLOW, F_SHAKE_FAIL, F_NBUILD, F_NSLOT };

enum OverflowBits {
  OV_SLOTS = 1, OV_COLUMN = 2, OV_SEARCH = 4, OV_PAIRS = 8,
  OV_BOX = 16  // barostat: the box shrank below NPT_MIN_SCALE of its size at the barostat's start
};

// Andersen barostat: the smallest box scale the column search ranges are sized for
constexpr double NPT_MIN_SCALE = 0.85;

// Andersen barostat state of one system (device memory).  The box is
// ell * L0, V = ell^3 V0; Pi = M dV/dt; Wtr the molecular virial at the current
// positions (tr sum r F minus each constraint cluster's internal part); ell_b
// the box scale at the last list build (F_NBUILD then nbuild).  ell0, ell1, a1,
// a2, V1, ke1 and wintra carry one step's intermediate values between its
// kernels.
struct NptState {
  double V, Pi, ell, ell_b, Wtr;
  double V1, ell0, ell1, a1, a2;
  unsigned long long ke1;     // molecular kinetic energy after the first kick, fixed point 2^28
  unsigned long long wintra;  // sum over clusters of (r_i - R_c) . F_i, fixed point 2^28
  int nbuild;
  int pad;
};

struct SysDev {
  // ---- sizes -----------------------------------------------------------
  int n_atoms;
  int n_types;
  int n_slots_cap;       // cluster slots (8 per cluster) capacity
  int n_clusters_cap;
  long long n_pairs_cap; // cluster-pair list capacity
  int ncx, ncy, ncol;    // column grid
  int kx_range, ky_range; // column search half-widths
  unsigned long long colw_x, colw_y; // floor(2^64 / ncx), floor(2^64 / ncy)
  // z windows of the pair search: z is cut into nzb bins (bin(z) = mulhi(u_z, nzb));
  // per column, zt_hi[c * (nzb + 1) + b] = first cluster k (within the column) whose
  // upper z bound has bin >= b, zt_lo likewise for the lower bound (b = nzb: count)
  int nzb;
  unsigned long long zwin;  // pair-list distance in z fixed-point units, rounded up
  float rl2f;               // FP32 atom-pair list threshold: rl^2 (1 + 1e-6)
  int K[3];              // PME grid
  int Kxh;               // K[0]/2 + 1
  int use_excl_list;     // 1: exclusions by sorted list, 0: 64-bit window bit vectors
  int n_ccl;             // constraint clusters
  int n_free;            // atoms in no constraint cluster
  unsigned long long spread_mult;  // atom visiting order of the atomic PME spread: odd, coprime to n_atoms
  // bonded term counts (flat index order: bonds, angles, dihedrals, impropers, ew pairs, 1-4 pairs)
  int n_bonds, n_angles, n_diheds, n_imprs, n_ew, n_p14;
  // ---- geometry / method parameters ----------------------------------
  double L[3];
  double sc[3];          // L * 2^-64: fixed -> Angstrom
  double isc[3];         // 2^64 / L
  float sc32[3];         // L * 2^-32 (32-bit top words)
  double rc2, ron2, rl2, skin2;   // cutoff^2, switch-on^2, pair-list^2, (0.5*(rl-rc))^2
  double beta;
  double sw_inv;         // 1 / (rc^2 - ron^2)^3
  double dt;             // AKMA time step (TU * dt_ps)
  double shake_tol;
  double pme_pre;        // CC / (pi V)
  double pi2_beta2;      // pi^2 / beta^2
  // ---- per-atom state (atom id order) --------------------------------
  unsigned long long *ux, *uy, *uz;   // fixed-point fractional positions
  unsigned long long *rx, *ry, *rz;   // positions at the last list build
  unsigned long long *ox, *oy, *oz;   // positions before the drift (SHAKE reference)
  double *vx, *vy, *vz;
  double *q, *minv;
  // inverse mass used by the integrator, SHAKE/RATTLE and the minimisers: minv
  // for mobile atoms, 0 for fixed atoms (consfix), which therefore keep zero
  // velocity and never move
  double *mdyn;
  int n_mobile;
  int *type;
  // ---- harmonic COM restraint between two atom groups (consharm) ------
  // Every evaluation, the atoms of group 1 receive fc * (com_d + c0 - c1)
  // (minimum image), c_g the mass-weighted mean of group g's positions taken in
  // [-L/2, L/2).  Group sums of m x go through A_COM (fixed point, 2^28).
  int com_on;
  int *com_grp;          // [n_atoms] 0, 1 or -1
  double com_d[3], com_fc, com_im[2];  // offset, force constant, 1 / group masses
  // ---- cluster-sorted arrays (slot order) ----------------------------
  int *slot_atom;        // [n_slots_cap], -1 = padding
  int *atom_slot;        // [n_atoms]
  void *xs;              // packed positions+charge per slot (int4 / longlong4)
  int *ts;               // type per slot
  void *Fs;              // pair-kernel (and restraint) force sums, [n_slots_cap][3]: a slot's x, y, z adjacent
  void *Fa;              // bonded force sums, 3 x n_atoms, atom order (cleared by the consumer)
  void *Fg;              // PME gather forces, 3 x n_atoms, atom order (stored by every evaluation)
  // ---- binning / search --------------------------------------------------
  int *atom_col;         // [n_atoms]
  int *col_count, *col_fill, *col_clu0; // [ncol], [ncol], [ncol + 1]
  int *zt_hi, *zt_lo;                   // [ncol * (nzb + 1)] z-window tables
  int *bin_tmp;          // [n_slots_cap] column-scattered atoms
  int *clu_col;          // [n_clusters_cap]
  unsigned long long *bb; // [6 * n_clusters_cap] lo/hi per axis
  unsigned long long *su; // [3 * n_slots_cap] slot-ordered positions at the last build (x, y, z blocks)
  int *n_clusters;       // [0] clusters of the last build
  unsigned *tick;        // [1] last-block ticket of the binning count
  int *search_cnt;       // [n_clusters_cap * search_parts + 1] kept entries per (I, part)
  int search_parts;      // 1, 2 or 4: search waves per I cluster the buffers are sized for
  int *search_tmp;       // [n_clusters_cap * search_parts * SEARCH_TMP_CAP] J of each kept entry
  unsigned long long *search_mtmp; // [same] lane mask of each kept entry
  int *tile_cnt;         // [ceil(n_clusters_cap / 64)] kept entries per 64 consecutive I clusters
  int *pl_i, *pl_j;      // [n_pairs_cap]
  unsigned long long *pl_mask; // [n_pairs_cap]
  long long *n_pairs;    // device scalar
  // ---- exclusions ---------------------------------------------------------
  unsigned long long *excl_bits; // [n_atoms]: bit (32 + j - i) set if j excluded, |j-i| <= 31
  int *excl_off, *excl_list;     // CSR (used when use_excl_list)
  // ---- bonded terms -------------------------------------------------------
  int2 *bond_ij;   double2 *bond_p;       // (k, r0)
  int4 *angle_ijk; double4 *angle_p;      // (k, theta0, k_ub, s0)
  int4 *dihe_ijkl; double2 *dihe_p; int *dihe_n; // (k, delta)
  int4 *impr_ijkl; double *impr_k;
  int2 *ew_ij;                            // Ewald-corrected exclusions (1-2, 1-3)
  int2 *p14_ij;                           // 1-4 pairs
  // ---- LJ tables --------------------------------------------------------------
  float2 *lj32;   double2 *lj64;          // (eps, sigma^2) per type pair
  // Mixed form: when every pair entry follows the combination rule
  // eps_ij = sqrt(eps_ii eps_jj), sigma_ij = (sigma_ii + sigma_jj) / 2 (no NBFIX),
  // lj_sep = 1 and lj_ah[t] = (sqrt(eps_tt), sigma_tt / 2); the pair kernel then
  // forms the pair parameters from per-slot values instead of a table lookup.
  int lj_sep;
  int mixed;             // precision form: slot layout of xs (int4 mixed, Slot64 FP64)
  float2 *lj_ah;                          // [n_types]
  float2 *lp;                             // [n_slots_cap] per slot: lj_ah[type] (lj_sep) or (type bits, 0)
  double2 *lj14;                          // 1-4 (eps, sigma^2)
  // ---- PME --------------------------------------------------------------------
  void *Qacc;      // K0*K1*K2 accumulators (double or fixed int64)
  void *Qc;        // complex half grid [K2][K1][Kxh] (float2 / double2)
  void *phi;       // real potential grid [K2][K1][K0] (float / double)
  double *pme_bsp; // per-axis |b(m)|^2, [K0 + K1 + K2]
  double *pme_ex;  // per-axis exp(-pi^2 m_d^2 / beta^2), [K0 + K1 + K2]
  double *pme_m;   // per-axis signed m_d (1/A), [K0 + K1 + K2]
  double *pme_g;   // G(m) over the half spectrum [K2][K1][Kxh] for the creation box; null once a barostat runs
  void *tw;        // twiddles e^{-2 pi i k / K_d}, [K0 + K1 + K2] complex
  // Charge spreading by grid bricks of 8^3 points: atoms are counted and
  // ordered by the brick holding the first point of their 4^3 stencil, each
  // brick's stencils are summed in LDS over its (8+3)^3 block, and the blocks
  // overlapping a grid point are added into Qacc.
  int nbr[3];      // bricks per axis, ceil(K_d / 8)
  int n_bricks;
  int *br_cnt;     // [n_bricks] atoms per brick (counted each step, cleared by the scan)
  int *br_off;     // [n_bricks + 1] first record of each brick
  int2 *br_atom;   // [n_atoms] (brick, rank in brick) of each atom, brick -1: none
  void *br_rec;    // [n_atoms] 16 reals per atom: stencil weights x, y, z (4 each), q, block offset
  void *br_part;   // [n_bricks * 1331] brick blocks (int64 fixed point / double)
  // ---- constraints --------------------------------------------------------
  int4 *ccl_atoms;    // up to 4 atoms (-1 = none)
  int *free_atoms;    // [n_free] atoms in no constraint cluster
  int4 *ccl_cons;     // x,y,z: constraint c as (local a | local b << 8); w: count
  double4 *ccl_d2;    // squared lengths (up to 3)
  // ---- accumulators / flags ---------------------------------------------
  unsigned long long *acc;  // [A_NSLOT] int64 fixed point (A_FMAX: bits of max |F|^2)
  unsigned long long *acc_ev;  // [EV_SPREAD * EV_W] pair energy / virial copies, zero between evaluations
  int *flags;               // [F_NSLOT]
  double *mstep;            // minimiser: [0] largest atom displacement of the next move (A) or the
                            // conjugate-gradient / steepest-descent factor, [1] schedule method
  double *mH, *mF;          // [3 * n_atoms] conjugate-gradient direction and previous force (x, y, z blocks)
  // snapshot of a call's starting state: blocks of n_atoms words ux, uy, uz,
  // vx, vy, vz (bits), rx, ry, rz, then the F_NSLOT status flags
  unsigned long long *snap;
  // ---- Andersen barostat (npt != 0) ------------------------------------
  int npt;
  double npt_pref;       // reference pressure, kcal/mol/A^3
  double npt_mass;       // piston mass, amu/A^4
  double npt_V0, npt_L0[3];
  double rl, rc;         // pair-list distance and cutoff (A)
  NptState *nst;         // [2]: live state, snapshot
};

// Box-dependent descriptor fields of a barostat system at box scale ell:
// box, fixed-point scales, the pair search's z window, the pair-list half-skin
// and the PME prefactor.  A pair beyond rl at the list's build scale ell_b is
// beyond rl ell / ell_b now, so the list holds every pair within rc while no
// atom has moved more than half of rl ell / ell_b - rc (the rebuild test).
// Host and device run the same arithmetic (no contraction), so both copies of
// the descriptor agree bit for bit; at ell = ell_b = 1 the fields are those of
// the system set-up.
__host__ __device__ inline void npt_box_fields(SysDev& d, double ell, double ell_b) {
#pragma clang fp contract(off)
  const double two64 = 18446744073709551616.0;
  for (int k = 0; k < 3; k++) {
    d.L[k] = ell * d.npt_L0[k];
    d.sc[k] = d.L[k] / two64;
    d.isc[k] = two64 / d.L[k];
    d.sc32[k] = (float)(d.L[k] / 4294967296.0);
  }
  d.zwin = (unsigned long long)(d.rl / d.sc[2] * (1.0 + 1e-9)) + 64ull;
  const double hs = 0.5 * (d.rl * (ell / ell_b) - d.rc);
  d.skin2 = hs > 0.0 ? hs * hs : 0.0;
  d.pme_pre = 332.0716 / (3.14159265358979323846 * (d.L[0] * d.L[1] * d.L[2]));
}

}  // namespace fz
