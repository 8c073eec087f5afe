/*
 * oracle.h -- FP64 CPU restatement of FEN ZI's per-step force-and-integrate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it.  The product
 * (libfenzi) never links or calls it.
 *
 * Positions are passed as 64-bit fixed-point fractional coordinates
 * u = frac(x / L) * 2^64 (uint64, AoS [N][3]); a displacement is
 * (double)(int64_t)(u_j - u_i) * L / 2^64, which is the minimum image for any
 * pair closer than L/2.  Velocities are in A per AKMA time unit (as the
 * reference stores them), forces in kcal/mol/A, energies in kcal/mol.
 */
#ifndef FZ_ORACLE_H
#define FZ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Energy slots (same order as the reference's energy table columns). */
enum {
  OR_E_BOND = 0, OR_E_ANGLE, OR_E_UB, OR_E_IMPR, OR_E_DIHED,
  OR_E_VDW, OR_E_ELEC, OR_E_PME, OR_E_EWEXCL, OR_E_EWSELF, OR_NE
};

/* Term selection flags for or_forces. */
enum {
  OR_REAL = 1,        /* real-space LJ (switched) + erfc Coulomb over non-excluded pairs */
  OR_EXCL = 2,        /* Ewald exclusion correction (1-2, 1-3) and 1-4 pairs */
  OR_BONDED = 4,      /* bonds, angles, Urey-Bradley, dihedrals, impropers */
  OR_PME = 8,         /* smooth PME reciprocal space (order 4) */
  OR_EWALD = 16,      /* direct Ewald k-sum reciprocal space (replaces OR_PME) */
  OR_SELF = 32,       /* Ewald self energy */
  OR_RESTRAINT = 64,  /* COM restraint forces (when set by or_set_com_restraint) */
  OR_ALL = OR_REAL | OR_EXCL | OR_BONDED | OR_PME | OR_SELF | OR_RESTRAINT
};

typedef struct or_topo {
  int32_t n_atoms, n_types;
  const double *charge, *mass;
  const int32_t *type;
  const double *lj_eps, *lj_sig2, *lj14_eps, *lj14_sig2; /* [n_types * n_types] */
  int32_t n_bonds;      const int32_t *bond_atoms;      const double *bond_k, *bond_r0;
  int32_t n_angles;     const int32_t *angle_atoms;     const double *angle_k, *angle_theta0, *ub_k, *ub_s0;
  int32_t n_dihedrals;  const int32_t *dihedral_atoms;  const double *dihedral_k; const int32_t *dihedral_n; const double *dihedral_delta;
  int32_t n_impropers;  const int32_t *improper_atoms;  const double *improper_k, *improper_psi0;
  int32_t n_constraints; const int32_t *constraint_atoms; const double *constraint_length;
} or_topo;

typedef struct or_params {
  double box[3];
  double cutoff, cuton, pairlist, beta;
  int32_t K[3];
  double dt_ps;
  double shake_tol;
  double ewald_kmax_tol; /* truncation of the direct k-sum: exp(-pi^2 m^2/beta^2) < tol */
} or_params;

typedef struct or_sys or_sys;

or_sys *or_create(const or_topo *t, const or_params *p);
void or_destroy(or_sys *s);
int or_set_threads(int n);

/* Fixed point <-> Cartesian (x wrapped into [-L/2, L/2)). */
void or_to_fixed(const double box[3], const double *x, uint64_t *u, int64_t n);
void or_from_fixed(const double box[3], const uint64_t *u, double *x, int64_t n);

/* Forces (accumulated into f, which is zeroed first), energies [OR_NE], virial [9]. */
void or_forces(or_sys *s, const uint64_t *u, double *f, double *e, double *vir, int flags);

/* Exclusion bookkeeping: counts and sorted (i<j) pair lists. */
int64_t or_n_pairs(or_sys *s, int which); /* 0 = all exclusions, 1 = Ewald-corrected (1-2,1-3), 2 = 1-4 */
void or_get_pairs(or_sys *s, int which, int32_t *ij);

/* Non-excluded atom pairs with r < rcut (count only; brute force if n small). */
int64_t or_count_pairs_within(or_sys *s, const uint64_t *u, double rcut);

/* Column/cluster binning and cluster-pair list, restating the GPU definition.
 * nx, ny: column grid; M = 8 atoms per cluster.  Outputs (caller-allocated):
 *   atom_column[N], slot_atom[n_slots_cap] (-1 = padding),
 *   pairs[2 * cap] (I, J with I <= J), masks[cap]; returns number of pairs
 *   (or -needed if cap too small); *n_slots receives the number of slots. */
int64_t or_cluster_pairs(or_sys *s, const uint64_t *u, int32_t nx, int32_t ny,
                         int32_t *atom_column, int32_t *slot_atom, int64_t n_slots_cap,
                         int64_t *n_slots, int32_t *pairs, uint64_t *masks, int64_t cap);

/* Fixed atoms (consfix; src/prmread.cu:699-722,1665-1728; mobility v.w of the
 * half kick, src/mdcuda.cu:70-80): fixed[i] != 0 holds atom i (zero inverse
 * mass in the kicks, SHAKE and RATTLE, zero Maxwell velocity; constraints
 * between fixed atoms dropped; temperature over 3 N_mobile - 6 - N_cons).
 * NULL releases every atom.  Returns 0, or -1 if every atom would be fixed. */
int or_set_fixed(or_sys *s, const uint8_t *fixed);
/* Harmonic COM restraint (reduce_COM src/cell_based_kernels.cu:713-866,
 * restraint_force src/mdcuda.cu:481-516): group[i] in {0, 1, -1}; with flag
 * OR_RESTRAINT, every atom of group 1 receives fc * (dist + c0 - c1), minimum
 * image, c_g the mass-weighted mean of group g's positions in [-L/2, L/2).
 * No energy.  group NULL or fc 0 removes it. */
int or_set_com_restraint(or_sys *s, const int32_t *group, const double dist[3], double fc);

/* Velocity Verlet with SHAKE/RATTLE: nsteps steps; u, v updated in place. */
int or_step(or_sys *s, uint64_t *u, double *v, int nsteps);
/* or_step that also returns the energies [OR_NE] and the virial [9] of the last
 * step (nullable): the force virial at the final positions plus the
 * constraint virial of the final RATTLE projection, (2/dt) sum_c b_c r_c (x) r_c. */
int or_step_ex(or_sys *s, uint64_t *u, double *v, int nsteps, double *e_out, double *vir_out);
/* Andersen barostat (isotropic): piston state and parameters.  L0: the box at
 * V0 = L0[0] L0[1] L0[2]; the box is cbrt(V / V0) L0.  p_ref in kcal/mol/A^3
 * (pref [atm] x 1.4586e-5, src/defs.in.h:105), mass in amu/A^4 (pmass,
 * src/prmread.cu:664-670); V in A^3; Pi = M dV/dt; Wtr the virial trace at the
 * current positions (constraint part of the last RATTLE included). */
typedef struct or_piston {
  double L0[3];
  double p_ref, mass;
  double V, Pi, Wtr;
} or_piston;
/* nsteps NPT steps (see oracle.c); the oracle's box follows V and is left at
 * the final box. */
int or_step_npt(or_sys *s, or_piston *pc, uint64_t *u, double *v, int nsteps, double *e_out, double *vir_out);
double or_virial_trace(or_sys *s, const uint64_t *u);
void or_set_box(or_sys *s, const double box[3]);
void or_get_box(const or_sys *s, double box[3]);
/* SHAKE of positions u against reference positions uold (minimiser moves). */
int or_shake_positions(or_sys *s, const uint64_t *uold, uint64_t *u);
/* Maxwell velocities (Philox4x32-10 + Box-Muller), momentum removal, RATTLE projection. */
void or_maxwell(or_sys *s, const uint64_t *u, double T, uint64_t seed, uint64_t counter, double *v);
double or_kinetic(or_sys *s, const double *v);
double or_temperature(or_sys *s, const double *v);
/* Constraint residual: max_c |r^2 - d^2| / (2 d^2); velocity residual max |v_ab . r_ab| / |r_ab|. */
double or_constraint_residual(or_sys *s, const uint64_t *u, double *vres, const double *v);

/* Unnormalised complex 3-D DFT, data [K2][K1][K0] interleaved (re, im); sign -1 forward, +1 inverse. */
void or_fft3d(const int32_t K[3], const double *in, double *out, int sign);
/* Philox4x32-10 block. */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
