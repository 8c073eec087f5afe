/*
 * fenzi.h -- C ABI of libfenzi, an MI355X (gfx950) implementation of FEN ZI's
 * per-step force-and-integrate path.
 *
 * The reference has no library interface: its hot path is a set of
 * `void f(void)` host functions over global state inside one nvcc translation
 * unit.  Each entry point below names the reference function(s) it replaces
 * (paths relative to the reference source tree).
 *
 * Conventions
 *   - Units are AKMA: Angstrom, kcal/mol, amu, e.  Velocities are in Angstrom
 *     per AKMA time unit (1/20.4548 ps), as the reference stores them
 *     (src/mdcuda.cu:27-92).  Time steps are given in ps (`timestep` keyword).
 *   - Arrays of vectors are [n][3] row-major (x0 y0 z0 x1 ...).
 *   - Every call returns 0 on success or a negative FZ_ERR_* code; the message
 *     of the last failure on the calling thread is fz_last_error().  Nothing
 *     calls exit().
 *   - Handles own their device memory; caller arrays are copied, never kept.
 *     Calls on one handle must not overlap; different handles may be used from
 *     different threads.
 */
#ifndef FENZI_H
#define FENZI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FZ_OK = 0,
  FZ_ERR_ARG = -1,      /* invalid argument or unsupported input */
  FZ_ERR_HIP = -2,      /* HIP runtime error */
  FZ_ERR_CAPACITY = -3, /* device capacity exceeded and could not be grown */
  FZ_ERR_PARSE = -4,    /* input deck parse error */
  FZ_ERR_NOMEM = -5,    /* out of device or host memory */
  FZ_ERR_SHAKE = -6,    /* SHAKE did not converge */
  FZ_ERR_STATE = -7,    /* call not valid in the handle's current state */
  FZ_ERR_BUDGET = -8    /* serving: job exceeds the HBM budget */
};

typedef enum { FZ_FP64 = 0, FZ_MIXED = 1 } fz_precision;

typedef struct fz_ctx fz_ctx;
typedef struct fz_system fz_system;
typedef struct fz_batch fz_batch;
typedef struct fz_server fz_server;
typedef struct fz_deck fz_deck;

/* Parsed topology (the products of ReadPsf_CHARMM + ReadPrm_CHARMM,
 * src/prmread.cu:1442,2181).  Pair tables are already combined:
 * eps_ij = 4 sqrt(eps_i eps_j), sigma_ij = sigma_i/2 + sigma_j/2 (or NBFIX),
 * E_LJ = eps_ij ((sigma_ij/r)^12 - (sigma_ij/r)^6).  Energies: bond
 * k (r - r0)^2; angle k (theta - theta0)^2 (radians); Urey-Bradley
 * k_ub (S - S0)^2; dihedral k (1 + cos(n phi - delta)) (one entry per term);
 * improper k psi^2 (psi0 kept for reference but unused, src/bonded_forces.cu:196).
 * Exclusions are derived: 1-2 from bonds, 1-3 from angle ends, 1-4 from
 * dihedral ends (src/prmread.cu:2928-3062). */
typedef struct fz_topology {
  int32_t n_atoms, n_types;
  const double *charge, *mass;
  const int32_t *type;
  const double *lj_eps, *lj_sig2, *lj14_eps, *lj14_sig2; /* [n_types * n_types] */
  int32_t n_bonds;       const int32_t *bond_atoms;       const double *bond_k, *bond_r0;
  int32_t n_angles;      const int32_t *angle_atoms;      const double *angle_k, *angle_theta0, *ub_k, *ub_s0;
  int32_t n_dihedrals;   const int32_t *dihedral_atoms;   const double *dihedral_k; const int32_t *dihedral_n; const double *dihedral_delta;
  int32_t n_impropers;   const int32_t *improper_atoms;   const double *improper_k, *improper_psi0;
  int32_t n_constraints; const int32_t *constraint_atoms; const double *constraint_length;
} fz_topology;

/* Run parameters (the `.in` keywords of ReadInput, src/prmread.cu:491-913). */
typedef struct fz_params {
  double box[3];          /* cellbasisvector1/2/3 (orthorhombic) */
  double cutoff;          /* cutoff */
  double cuton;           /* cuton (LJ switch start) */
  double pairlist;        /* pairlistdist */
  double beta;            /* KAPPa (Ewald splitting, 1/A) */
  int32_t pme_grid[3];    /* pmegridsizex/y/z (2,3,5-smooth, <= 256) */
  int32_t pme_order;      /* must be 4 */
  double dt_ps;           /* timestep (ps) */
  double shake_tol;       /* shake on <tol>: max |r^2 - d^2| / (2 d^2) */
  int32_t column_grid[2]; /* binning columns (x, y); 0 = chosen automatically */
} fz_params;

typedef struct fz_energies {
  int64_t step;
  double temperature;     /* 2 KE / (RC (3N - 6 - N_constraints)) (src/evalprops.cu:221-225) */
  double bond, angle, urey_bradley, improper, dihedral;
  double vdw;             /* switched LJ, including 1-4 pairs */
  double elec;            /* real-space erfc Coulomb, including 1-4 pairs */
  double pme;             /* reciprocal-space energy */
  double ewald_excl;      /* -erf correction of 1-2 / 1-3 pairs */
  double ewald_self;      /* -CC beta / sqrt(pi) sum q^2 */
  double potential;       /* sum of the terms above */
  double kinetic;
  double virial[9];       /* sum r_ij (x) F_ij + reciprocal tensor; after fz_step also the constraint
                             virial of the last step's RATTLE, (2/dt) sum_c lambda_c r_c (x) r_c
                             (src/shake.cu:300-306, there halved and diagonal only) */
  int64_t pairs_within_cutoff; /* non-excluded pairs inside the cutoff, this evaluation */
  int64_t list_builds;    /* neighbour-list builds so far */
  double volume;          /* box volume (A^3) */
  double pressure;        /* (2 KE + tr virial) / 3V in atm (reduce_virial, src/cell_based_kernels.cu:586-589) */
  int64_t pair_tiles;     /* 8x8 cluster-pair tiles the pair kernel evaluated, this evaluation (one per
                             list entry); pairs_within_cutoff / (64 pair_tiles) is the in-cutoff lane fraction */
} fz_energies;

/* ---- errors / context ------------------------------------------------- */
const char *fz_last_error(void);
/* Replaces parseArgs' cudaSetDevice (src/fenzi.cu:59-110). */
int fz_ctx_create(int device, fz_ctx **out);
void fz_ctx_destroy(fz_ctx *ctx);

/* ---- one system ------------------------------------------------------------ */
/* Replaces InitHostMem + InitMem + InitDeviceConstants + GenExcl_minimal_new +
 * build_constraints (src/prmread.cu:86,1303,2928; src/mem.cu:49;
 * src/cell_based_kernels.cu:1831). */
int fz_system_create(fz_ctx *ctx, const fz_topology *topo, const fz_params *par, int precision, fz_system **out);
void fz_system_destroy(fz_system *sys);
/* Positions (any image; wrapped on input) and velocities (nullable = zero).
 * Replaces the coordinate upload of InitMem and LoadCheckpoint's H2D copies
 * (src/fenzi.cu:481-501).  Recomputes forces. */
int fz_set_state(fz_system *sys, const double *xyz, const double *vel);
/* Positions wrapped to [-L/2, L/2), velocities (either nullable). */
int fz_get_state(fz_system *sys, double *xyz, double *vel);
/* Exact 64-bit fixed-point fractional positions u = frac(x / L) * 2^64. */
int fz_set_fixed_positions(fz_system *sys, const uint64_t *u);
int fz_get_fixed_positions(fz_system *sys, uint64_t *u);
/* Maxwell velocities at T (Philox4x32-10 stream (seed, counter)), total
 * momentum removed, constraint-projected.  Replaces InitVelocity
 * (src/init.cu:19) and its time(NULL) seeding. */
int fz_reassign_velocities(fz_system *sys, double temperature, uint64_t seed, uint64_t counter);
/* Fixed atoms (consfix; src/prmread.cu:699-722,1665-1728, mobility v.w of the
 * half kick src/mdcuda.cu:70-80): fixed[i] != 0 holds atom i in place (zero
 * velocity; not moved by dynamics, SHAKE/RATTLE or the minimisers).
 * Constraints between fixed atoms are dropped; the temperature counts
 * 3 N_mobile - 6 - N_constraints degrees of freedom (src/evalprops.cu:222-225).
 * fixed = NULL releases every atom. */
int fz_set_fixed_atoms(fz_system *sys, const uint8_t *fixed);
/* Harmonic centre-of-mass restraint (consharm / consharmdist / consharmfc,
 * src/prmread.cu:674-690; reduce_COM src/cell_based_kernels.cu:713-866,
 * restraint_force src/mdcuda.cu:481-516): group[i] in {0, 1, -1}.  Each
 * evaluation the atoms of group 1 receive fc * (dist + c0 - c1), minimum image,
 * c_g the mass-weighted mean of group g's positions in [-L/2, L/2); group 0
 * receives no force and no energy is defined (as in the reference).
 * group = NULL or fc = 0 removes the restraint. */
int fz_set_com_restraint(fz_system *sys, const int32_t *group, const double dist[3], double fc);
/* Forces (nullable, [n][3]) and energies (nullable) at the current positions.
 * Replaces ComputeAccelGPU (src/mdcuda.cu:557) + EvalProps' reductions
 * (src/evalprops.cu:17). */
int fz_compute_forces(fz_system *sys, double *forces, fz_energies *e);
/* nsteps of velocity Verlet (+ SHAKE/RATTLE); neighbour rebuilds decided on
 * the device.  A capacity overflow is recovered inside the call (restore to the
 * call's start, grow, replay).  On an error (SHAKE failure, capacity that cannot
 * grow) the system is back in the state it had before the call.  Replaces
 * SingleStep (src/dynamics.cu:623). */
int fz_step(fz_system *sys, int nsteps);
/* Andersen barostat, isotropic (pmass / pref keywords, src/prmread.cu:664-670;
 * the piston of reduce_virial, src/cell_based_kernels.cu:403-600, whose NPT
 * step the reference leaves unfinished, src/npt_kernels.cu:26-33): the volume
 * V follows V'' = (P - pref) / M with P = (2 KE + tr W) / 3V, W the virial with
 * its SHAKE/RATTLE part; positions scale with the box.  pref in atm, piston
 * mass M in amu/A^4; mass <= 0 turns the barostat off and keeps the current
 * box.  The box at the call is the reference box (scale 1); fz_step fails
 * (FZ_ERR_STATE, state restored) if it shrinks below 0.85 of it.  Thermostat:
 * the caller's velocity reassignment (fz_reassign_velocities). */
int fz_set_barostat(fz_system *sys, double pref_atm, double piston_mass);
/* Current box [3] and piston state (nullable) [3]: volume, piston momentum
 * M dV/dt, virial trace at the current positions. */
int fz_get_barostat(fz_system *sys, double *box, double *state);
/* Energies of the last evaluated step (temperature from current velocities). */
int fz_get_energies(fz_system *sys, fz_energies *e);
/* Adaptive steepest descent, nsteps iterations: the atom with the largest force
 * moves max_step (A), an energy increase is undone and halves the step, a
 * decrease grows it by 1.2 (capped); SHAKE after every move; velocities zeroed.
 * Replaces MinimizeStep (src/dynamics.cu:491) with a step rule of this build. */
int fz_minimize(fz_system *sys, int nsteps, double max_step);

/* The reference's minimisation schedule (MinimizeStep, src/dynamics.cu:491-620;
 * ConjugatedGradient / SteepestDescent, src/mdcuda.cu:126-330; factor ramp,
 * src/fenzi.cu:402-415): per-atom Fletcher-Reeves conjugate gradient
 * r += (F + gamma H_prev) cgfac, gamma = |F|^2/|F_prev|^2; on an energy
 * increase the move is undone and conjugate gradient switches to steepest
 * descent (r += F sdfac), and an increase under steepest descent ends the
 * minimisation; after every 100 accepted moves the active factor is multiplied
 * by rate * 100.  The .in keywords conjgradfac, steepdesfac, conjgradrate,
 * steepdesrate (src/prmread.cu:547-561).  This build also applies SHAKE after
 * each move.  method: 0 adaptive steepest descent (fz_minimize, max_step),
 * 1 the reference schedule from conjugate gradient, 2 from steepest descent. */
typedef struct fz_min_params {
  int32_t method;
  double max_step;
  double cgfac, sdfac, cgrate, sdrate; /* reference defaults 1e-10, 1e-4, 0.1, 0.01 */
} fz_min_params;
int fz_minimize_ex(fz_system *sys, int nsteps, const fz_min_params *p);
int fz_system_info(fz_system *sys, int64_t *n_atoms, int64_t *device_bytes, int32_t *precision);
/* Neighbour-list capacities: pair-list entries and clusters (0 keeps the
 * current value).  The list is rebuilt at the positions of its last build, so
 * the state and the list the system holds do not change; a later build that
 * does not fit grows them again (restore to the start of the fz_step call and
 * replay).  Replaces the compile-time MAXNB / CELL_ATOMS capacities of the
 * reference (src/_config/fenzi_conf_options:20, src/configure:89-137) and its
 * exit on overflow (Check_Static_Parameters, src/dynamics.cu:52-234). */
int fz_set_neighbor_capacity(fz_system *sys, int64_t pairs_cap, int32_t clusters_cap);
/* Per-phase device time: runs nsteps steps without graph capture, with HIP
 * events between phases on the system's stream; mean ms per step per phase in
 * ms[FZ_NPHASE], phases: 0 first kick + drift + SHAKE + rebuild test, 1
 * neighbour rebuild (early exits when none is due), 2 sorted-coordinate gather
 * (run here every step; replayed dynamics steps keep the copies in phase 0 and
 * the rebuild), 3 pair kernel, 4 bonded + exclusion
 * terms, 5 PME spread (on the plane path with the x and y transforms), 6 PME
 * FFT + convolution (plane path: z pass and inverse planes), 7 PME force gather,
 * 8 second kick + RATTLE.  Replaces the reference's PROFILING clock() buckets
 * (src/globals.h:796-823). */
#define FZ_NPHASE 9
int fz_phase_times(fz_system *sys, int nsteps, double *ms);
/* Mean time (ms) of the force-only pair kernel over reps launches back to back
 * on the system's stream between two HIP events (the per-phase events above
 * add each launch's gap to a single kernel); the forces are evaluated again
 * afterwards.  For the pair kernel's roofline; the reference has only the
 * PROFILING clock() buckets (src/globals.h:796-823). */
int fz_pair_kernel_time(fz_system *sys, int reps, double *ms);
/* Autotuner: times graph-replayed steps for candidate column widths (atoms per
 * binning column and cluster compactness), FFT lines per workgroup and
 * persistent pair-kernel grid sizes, keeps the fastest and restores the state.
 * out[4] = {column width factor, FFT lines, pair grid blocks, ms per step}.
 * Replaces the compile-time CELL_BLOCKSIZE / CELL_ATOMS presets of the
 * reference's configure step (src/configure:89-137). */
int fz_autotune(fz_system *sys, int nsteps, double *out);
/* Launch shapes set directly (fz_autotune chooses the first two, and the
 * column width, which this call keeps): FFT lines per workgroup and persistent
 * pair-kernel blocks (0 = default, < 0 = keep), PME path (< 0 = keep):
 * FZ_PME_AUTO, the plane kernels for systems of up to 65,536 atoms whose grid
 * plane fits in 64 KiB of LDS and grid atomics or bricks otherwise;
 * FZ_PME_GRID; FZ_PME_PLANES (FZ_ERR_ARG where a plane does not fit).  All
 * paths give the same grid sums (bitwise in the mixed form).  in_effect
 * (nullable, 3 entries): FFT lines, pair blocks, PME path (FZ_PME_GRID or
 * FZ_PME_PLANES) in use.  Replaces the same compile-time presets
 * (src/configure:89-137). */
#define FZ_PME_AUTO 0
#define FZ_PME_GRID 1
#define FZ_PME_PLANES 2
int fz_set_launch(fz_system *sys, int fft_lines, int pair_grid, int pme_path, int32_t *in_effect);

/* ---- inspection (parity tests) ------------------------------------------------ */
/* Column grid and cluster layout of the last neighbour build: column of each
 * atom, atom of each slot (-1 = padding).  Replaces the cell lists of
 * CellBuild (src/cell_based_kernels.cu:198). */
int fz_get_cluster_layout(fz_system *sys, int32_t *ncx_ncy, int32_t *atom_column, int32_t *slot_atom,
                          int64_t slot_cap, int64_t *n_slots);
/* Cluster-pair list (I, J) and 64-bit tile masks; returns the number of pairs
 * (negative FZ_ERR_* on failure).  Replaces nblistd (src/cell_based_kernels.cu:1048). */
int64_t fz_get_cluster_pairs(fz_system *sys, int32_t *pairs, uint64_t *masks, int64_t cap);

/* PME convolution test entry: real grid q[K2][K1][K0] -> phi (unnormalised
 * inverse DFT of G * DFT(q)); mode 0 = PME kernel G for (box, beta), mode 1 = G = 1.
 * Returns the reciprocal energy 1/2 sum G |DFT(q)|^2 in *energy (mode 0). */
int fz_pme_convolve(fz_ctx *ctx, const int32_t K[3], const double box[3], double beta, int precision,
                    int mode, const double *q, double *phi, double *energy);

/* fz_pme_convolve over device memory, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  q_dev and phi_dev are the caller's
 * device grids of K0*K1*K2 doubles, energy_dev a nullable device double; the
 * whole call is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and nothing is synchronised.  Every buffer beyond them lives in one static
 * device array of fz_pme_convolve_noalloc_bytes() bytes per GPU, held until
 * the process ends; grids above 2^18 points (64^3) are rejected (FZ_ERR_ARG).
 * Calls that overlap on one GPU share that array and race.  The GPU's context
 * (fz_ctx_create) must exist. */
int fz_pme_convolve_noalloc(fz_ctx *ctx, const int32_t K[3], const double box[3], double beta, int precision,
                            int mode, const double *q_dev, double *phi_dev, double *energy_dev, void *stream);
int64_t fz_pme_convolve_noalloc_bytes(void);

/* ---- batch: many systems per launch ---------------------------------------------- */
/* Systems stay owned by their handles; the batch advances them in place with
 * one launch per phase per step (hipGraph-captured).  fz_batch_step,
 * fz_batch_minimize and fz_batch_minimize_ex leave every member in the state
 * it had before the call when they fail. */
int fz_batch_create(fz_ctx *ctx, fz_system *const *systems, int n, fz_batch **out);
int fz_batch_step(fz_batch *b, int nsteps);
/* Steepest descent on every member (per-system step control), one launch per phase. */
int fz_batch_minimize(fz_batch *b, int nsteps, double max_step);
/* fz_minimize_ex on every member (each member keeps its own schedule state). */
int fz_batch_minimize_ex(fz_batch *b, int nsteps, const fz_min_params *p);
/* fz_autotune over the whole batch (one setting for all members). */
int fz_batch_autotune(fz_batch *b, int nsteps, double *out);
/* fz_phase_times over the whole batch: each phase is one launch over every member. */
int fz_batch_phase_times(fz_batch *b, int nsteps, double *ms);
/* fz_set_launch for the whole batch. */
int fz_batch_set_launch(fz_batch *b, int fft_lines, int pair_grid, int pme_path, int32_t *in_effect);
void fz_batch_destroy(fz_batch *b);

/* ---- serving: continuous batching of MD jobs ------------------------------------------ */
/* The reference runs one system per process (main, src/fenzi.cu:112); the
 * server keeps many jobs resident in HBM.  Each job: steepest-descent
 * minimisation, Maxwell velocities, NVT dynamics with velocity reassignment
 * every reassign_every steps.  Jobs are admitted while the estimated HBM
 * footprint of the active set stays within the budget and fewer than
 * max_active are running; every fz_server_poll advances all minimising jobs
 * as one batch and all dynamics jobs as one batch, then retires finished jobs
 * and admits queued ones, so arrivals join running work without waiting for
 * it to drain.  A job that fails is marked failed and removed; the others
 * continue. */
typedef struct fz_server_config {
  double hbm_budget_bytes;  /* admission budget for resident job state */
  int32_t max_active;       /* at most this many jobs on the device */
  int32_t precision;        /* FZ_MIXED or FZ_FP64 */
  int32_t chunk_steps;      /* steps (or minimiser iterations) per poll round */
  double min_max_step;      /* steepest-descent step cap (A) */
} fz_server_config;

enum { FZ_JOB_QUEUED = 0, FZ_JOB_MINIMIZING = 1, FZ_JOB_DYNAMICS = 2, FZ_JOB_DONE = 3, FZ_JOB_FAILED = -1 };

typedef struct fz_job_info {
  int64_t id;
  int32_t state;
  int64_t n_atoms, steps_done, min_done, device_bytes;
  double t_submit, t_admit, t_done; /* seconds since the server was created */
} fz_job_info;

int fz_server_create(fz_ctx *ctx, const fz_server_config *cfg, fz_server **out);
/* Copies the topology, parameters and coordinates; returns the job id. */
int fz_server_submit(fz_server *srv, const fz_topology *topo, const fz_params *par, const double *xyz,
                     int32_t min_steps, int32_t md_steps, double temperature, int32_t reassign_every,
                     uint64_t seed, int64_t *job_id);
/* One round: admit, advance, retire.  Counts after the round (nullable). */
int fz_server_poll(fz_server *srv, int32_t *n_active, int32_t *n_queued, int32_t *n_done);
int fz_server_job_info(fz_server *srv, int64_t job_id, fz_job_info *info);
/* Final state of a finished job (any pointer nullable). */
int fz_server_job_result(fz_server *srv, int64_t job_id, double *xyz, double *vel, uint64_t *fixed, fz_energies *e);
/* Releases a finished job's host-side result. */
int fz_server_release(fz_server *srv, int64_t job_id);
/* Estimated device bytes of one job (the admission figure). */
int64_t fz_server_estimate_bytes(const fz_topology *topo, const fz_params *par, int32_t precision);
void fz_server_destroy(fz_server *srv);

/* ---- checkpoint and trajectory files ------------------------------------------------ */
/* Restart file in the reference's layouts (SaveCheckpointBinary/Ascii and
 * LoadCheckpointBinary/Ascii, src/prmwrite.cu:31-252): atom count, step, seed,
 * box, zero box velocities, then per atom position and velocity, single
 * precision.  binary: 1 binary, 0 text. */
int fz_restart_write(const char *path, int32_t binary, int64_t n_atoms, int64_t step, double seed,
                     const double *box, const double *xyz, const double *vel);
/* Fails (FZ_ERR_PARSE) on a short file or a different atom count.  Outputs nullable. */
int fz_restart_read(const char *path, int32_t binary, int64_t n_atoms, int64_t *step, double *seed,
                    double *box, double *xyz, double *vel);
/* CHARMM DCD trajectory (charmm_trj_write_header/titles/natoms/frame,
 * src/io/charmmio.c:50-175, Fortran records as ffwrite, src/io/ffile.c:155-203). */
typedef struct fz_dcd fz_dcd;
int fz_dcd_open(const char *path, int64_t n_atoms, int64_t nframes, int64_t first_step, int64_t every,
                int64_t nsteps, double dt_ps, fz_dcd **out);
int fz_dcd_write_frame(fz_dcd *f, const double *xyz, const double *box);
int fz_dcd_close(fz_dcd *f);

/* ---- input deck (.in / .psf / .prm / .rtf / .crd) --------------------------------- */
/* Replaces ReadInput, ReadCrd/ReadPdb/ReadXYZ, ReadPsf_CHARMM, ReadPrm_CHARMM
 * (src/prmread.cu:491,331,410,285,1442,2181). */
int fz_deck_load(const char *in_path, fz_deck **out);
void fz_deck_destroy(fz_deck *deck);
const fz_topology *fz_deck_topology(const fz_deck *deck);
const fz_params *fz_deck_params(const fz_deck *deck);
const double *fz_deck_coordinates(const fz_deck *deck);
/* Run settings of the deck: temperature, dynasteps, minimizesteps,
 * tempresetTiming, outputTiming, seed, shake flag, printtostdout. */
typedef struct fz_run_settings {
  double temperature;
  int64_t dynasteps, minimizesteps, startstep, tempreset_every, output_every;
  uint64_t seed;
  int32_t shake, print_to_stdout;
  char output_prefix[256];
  double cgfac, sdfac, cgrate, sdrate; /* conjgradfac, steepdesfac, conjgradrate, steepdesrate */
  int64_t checkpoint_every;            /* checkpointTiming (default 100000) */
  int64_t traj_every;                  /* trajfrequencyTiming (default outputTiming) */
  int32_t print_level;                 /* printlevel: 0 none, 1 energies + checkpoints, 2 + trajectory */
  int32_t checkpoint_binary;           /* checkpointformat: 1 binary (default), 0 ascii */
  char file_prefix[256];               /* outputtrajectory without its extension: <prefix>.rst, <prefix>_dev<d>.dcd */
  double pmass;                        /* barostat piston mass (amu/A^4; 0 = fixed box) */
  double pref;                         /* barostat reference pressure (atm, default 1) */
} fz_run_settings;
int fz_deck_settings(const fz_deck *deck, fz_run_settings *out);
/* Section counts of the PSF as read: bonds, angles, dihedrals, impropers. */
int fz_deck_psf_counts(const fz_deck *deck, int64_t *counts);
/* The deck's COM restraint: group[n_atoms] from the CRD segment names of the
 * consharm keyword (0 first, 1 second, -1 other), dist (consharmdist, default
 * 10 10 20) and fc (consharmfc, default 0.01).  Returns 1 if consharm is set,
 * 0 if not (group, dist and fc untouched), negative on error. */
int fz_deck_com_restraint(const fz_deck *deck, int32_t *group, double *dist, double *fc);
/* The deck's fixed atoms: fixed[n_atoms] = 1 for atoms whose PSF segment is
 * listed by consfix.  Returns the number of fixed atoms. */
int fz_deck_fixed_atoms(const fz_deck *deck, uint8_t *fixed);

/* Energy table in the reference's format (print_output_headers / EvalProps,
 * src/evalprops.cu:248-267,311-343).  Return the number of characters written. */
int fz_format_energy_header(char *buf, size_t len);
int fz_format_energy_row(const fz_energies *e, double dt_ps, char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
