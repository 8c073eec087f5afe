// fz_integrate.hip -- velocity Verlet, SHAKE/RATTLE, Maxwell velocities, kinetic energy.
//
// Replaces HalfKickGPU / UpdateCoords / CoordsUpdate (src/mdcuda.cu:27-123,
// 354-478), solve_bond_constraints / solve_velocity_constraints
// (src/shake.cu:20-615) and InitVelocity (src/init.cu:19-211).  Positions are
// 64-bit fixed-point fractions, so the periodic wrap is the integer overflow of
// the drift; SHAKE and RATTLE run in FP64 in both precision forms, one thread
// per constraint cluster (<= 4 atoms, <= 3 constraints), converged when
// max |r^2 - d^2| / (2 d^2) < tol.  Velocities are drawn on the device with
// Philox4x32-10 (counter based: reproducible for any thread layout).
#include "fz_device.h"
#include "fz_launch.h"

namespace fz {

// Force on atom a (slot sl): the slot-ordered sum of the pair kernel (and
// restraints) plus the atom-ordered bonded sum and PME gather term.  In the
// mixed form the three int64 fixed-point parts add as integers, so the force
// is the same whatever part a term went to.  zero: the pair and bonded sums
// are cleared (the gather term is overwritten by every evaluation).
template <bool MIXED>
__device__ __forceinline__ double take_force(const SysDev& d, int sl, int a, int comp, bool zero) {
  const size_t ia = (size_t)comp * d.n_atoms + a;
  if (MIXED) {
    unsigned long long* F = (unsigned long long*)d.Fs + (size_t)sl * 3 + comp;
    unsigned long long* B = (unsigned long long*)d.Fa + ia;
    const unsigned long long v = *F + *B + ((const unsigned long long*)d.Fg)[ia];
    if (zero) { *F = 0ull; *B = 0ull; }
    return from_fixed_acc(v);
  } else {
    double* F = (double*)d.Fs + (size_t)sl * 3 + comp;
    double* B = (double*)d.Fa + ia;
    const double v = (*F + *B) + ((const double*)d.Fg)[ia];
    if (zero) { *F = 0.0; *B = 0.0; }
    return v;
  }
}

// clears the pair and bonded force sums of atom a (slot sl)
template <bool MIXED>
__device__ __forceinline__ void clear_force(const SysDev& d, int sl, int a) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (MIXED) {
      ((unsigned long long*)d.Fs)[(size_t)sl * 3 + c] = 0ull;
      ((unsigned long long*)d.Fa)[(size_t)c * d.n_atoms + a] = 0ull;
    } else {
      ((double*)d.Fs)[(size_t)sl * 3 + c] = 0.0;
      ((double*)d.Fa)[(size_t)c * d.n_atoms + a] = 0.0;
    }
  }
}

// Kicks are explicit fused multiply-adds, so a kick rounds the same whether it
// runs alone or next to another kick with the same product (a separate
// product could otherwise be shared and rounded once for both).
// first half kick + drift of atom i (an atom of no constraint cluster); clears
// its force accumulators.  FINISH: the previous step's second half kick (same
// forces) comes first, as a separate rounded addition.
template <bool MIXED, bool FINISH = false>
__device__ __forceinline__ void kick_drift_atom(const SysDev& d, int i) {
  const int sl = d.atom_slot[i];
  const double h = 0.5 * d.dt * d.mdyn[i];
  const double fx = take_force<MIXED>(d, sl, i, 0, true), fy = take_force<MIXED>(d, sl, i, 1, true),
               fz = take_force<MIXED>(d, sl, i, 2, true);
  double vx = d.vx[i], vy = d.vy[i], vz = d.vz[i];
  if (FINISH) { vx = fma(h, fx, vx); vy = fma(h, fy, vy); vz = fma(h, fz, vz); }
  vx = fma(h, fx, vx); vy = fma(h, fy, vy); vz = fma(h, fz, vz);
  d.vx[i] = vx; d.vy[i] = vy; d.vz[i] = vz;
  const unsigned long long ux = d.ux[i] + (unsigned long long)__double2ll_rn(d.dt * vx * d.isc[0]);
  const unsigned long long uy = d.uy[i] + (unsigned long long)__double2ll_rn(d.dt * vy * d.isc[1]);
  const unsigned long long uz = d.uz[i] + (unsigned long long)__double2ll_rn(d.dt * vz * d.isc[2]);
  d.ux[i] = ux; d.uy[i] = uy; d.uz[i] = uz;
  store_slot_pos(d, sl, ux, uy, uz, d.q[i]);  // the pair kernel's slot copy
}

// second half kick of atom i
template <bool MIXED>
__device__ __forceinline__ void kick_atom(const SysDev& d, int i) {
  const int sl = d.atom_slot[i];
  const double h = 0.5 * d.dt * d.mdyn[i];
  d.vx[i] = fma(h, take_force<MIXED>(d, sl, i, 0, false), d.vx[i]);
  d.vy[i] = fma(h, take_force<MIXED>(d, sl, i, 1, false), d.vy[i]);
  d.vz[i] = fma(h, take_force<MIXED>(d, sl, i, 2, false), d.vz[i]);
}

// ---- constraint clusters ------------------------------------------------------

struct CCl {
  int na, nc;
  int at[4];
  int c1[3], c2[3];
  double d2[3];
};

__device__ __forceinline__ CCl load_ccl(const SysDev& d, int c) {
  CCl k;
  const int4 a = d.ccl_atoms[c];
  const int4 p = d.ccl_cons[c];
  const double4 l = d.ccl_d2[c];
  k.at[0] = a.x; k.at[1] = a.y; k.at[2] = a.z; k.at[3] = a.w;
  k.na = (a.x >= 0) + (a.y >= 0) + (a.z >= 0) + (a.w >= 0);
  k.nc = p.w;
  const int pp[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int j = 0; j < 3; j++) { k.c1[j] = pp[j] & 0xff; k.c2[j] = (pp[j] >> 8) & 0xff; }
  k.d2[0] = l.x; k.d2[1] = l.y; k.d2[2] = l.z;
  return k;
}

// b <- A^{-1} b for a 3 x 3 system (cofactor expansion; the constraint
// matrices are small and well conditioned); false if singular
__device__ __forceinline__ bool solve3(const double (&A)[3][3], double (&b)[3], int n) {
  (void)n;
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  if (det == 0.0 || !isfinite(det)) return false;
  const double id = 1.0 / det;
  const double x0 = (b[0] * c00 + b[1] * (A[0][2] * A[2][1] - A[0][1] * A[2][2]) +
                     b[2] * (A[0][1] * A[1][2] - A[0][2] * A[1][1])) * id;
  const double x1 = (b[0] * c01 + b[1] * (A[0][0] * A[2][2] - A[0][2] * A[2][0]) +
                     b[2] * (A[0][2] * A[1][0] - A[0][0] * A[1][2])) * id;
  const double x2 = (b[0] * c02 + b[1] * (A[0][1] * A[2][0] - A[0][0] * A[2][1]) +
                     b[2] * (A[0][0] * A[1][1] - A[0][1] * A[1][0])) * id;
  b[0] = x0; b[1] = x1; b[2] = x2;
  return true;
}

// The constraints of a cluster enter through its incidence matrix
// B[j][m] = +1 if atom m is the first atom of constraint j, -1 if it is the
// second, else 0.  Differences such as r[c1[j]] - r[c2[j]] are formed as sums
// over m weighted by B: every index is then a compile-time constant and the
// cluster state stays in registers (indexing by c1/c2 at run time put it in
// scratch memory).  Each sum has at most two non-zero terms, so it rounds
// exactly as the plain difference.
struct V3 { double x, y, z; };
__device__ __forceinline__ V3 v3sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double v3dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ void incidence(const CCl& k, double (&B)[3][4]) {
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int m = 0; m < 4; m++) B[j][m] = (double)((k.c1[j] == m) - (k.c2[j] == m));
}
__device__ __forceinline__ V3 bsum(const double (&Bj)[4], const V3 (&p)[4]) {
  V3 s = {0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < 4; m++) { s.x += Bj[m] * p[m].x; s.y += Bj[m] * p[m].y; s.z += Bj[m] * p[m].z; }
  return s;
}
// coupling weights w[j][e] = sum_m B[j][m] B[e][m] mi[m]
__device__ __forceinline__ void couplings(const double (&B)[3][4], const double (&mi)[4], double (&w)[3][3]) {
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int e = 0; e < 3; e++) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 4; m++) s += (B[j][m] * B[e][m]) * mi[m];
      w[j][e] = s;
    }
}

__device__ __forceinline__ void load_rel(const SysDev& d, const CCl& k, const unsigned long long* ux,
                                         const unsigned long long* uy, const unsigned long long* uz, V3 (&r)[4]) {
  const int a0 = k.at[0];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int a = k.at[m] >= 0 ? k.at[m] : a0;
    r[m] = V3{fx_d(ux[a0], ux[a], d.sc[0]), fx_d(uy[a0], uy[a], d.sc[1]), fx_d(uz[a0], uz[a], d.sc[2])};
  }
}

// SHAKE iterations on one cluster: rn and ro are the new and old positions
// relative to the cluster's first atom, mi the inverse masses (0 for padding);
// dr receives the position corrections.  False if it did not converge.
__device__ __forceinline__ bool shake_core(const SysDev& d, const CCl& k, const V3 (&rn)[4], const V3 (&ro)[4],
                                           const double (&mi)[4], V3 (&dr)[4]) {
  double B[3][4];
  incidence(k, B);
#pragma unroll
  for (int m = 0; m < 4; m++) dr[m] = V3{0.0, 0.0, 0.0};
  V3 g[3], r0[3];
  double lam[3] = {0.0, 0.0, 0.0};
  // coupling weights w[j][e]: d(r_c1j - r_c2j)/d(lambda_e) = w[j][e] * g_e
  double w[3][3];
  couplings(B, mi, w);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    g[j] = bsum(B[j], ro);   // old constraint vectors
    r0[j] = bsum(B[j], rn);  // unconstrained new ones
  }
  bool ok = false;
  for (int it = 0; it < 200; it++) {
    V3 rc[3];
    double sig[3], worst = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const V3 e = bsum(B[j], dr);
      rc[j] = V3{r0[j].x + e.x, r0[j].y + e.y, r0[j].z + e.z};
      sig[j] = j < k.nc ? v3dot(rc[j], rc[j]) - k.d2[j] : 0.0;
      if (j < k.nc) worst = fmax(worst, fabs(sig[j]) / (2.0 * k.d2[j]));
    }
    if (worst < d.shake_tol) { ok = true; break; }
    double J[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int e = 0; e < 3; e++)
        J[j][e] = (j < k.nc && e < k.nc) ? 2.0 * v3dot(rc[j], g[e]) * w[j][e] : (j == e ? 1.0 : 0.0);
#pragma unroll
    for (int j = 0; j < 3; j++) sig[j] = -sig[j];
    if (!solve3(J, sig, 3)) break;
#pragma unroll
    for (int j = 0; j < 3; j++) lam[j] += j < k.nc ? sig[j] : 0.0;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      V3 s = {0.0, 0.0, 0.0};
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const double l = j < k.nc ? lam[j] * B[j][m] * mi[m] : 0.0;
        s.x += l * g[j].x; s.y += l * g[j].y; s.z += l * g[j].z;
      }
      dr[m] = s;
    }
  }
  return ok;
}

__device__ __forceinline__ void shake_ccl(const SysDev& d, const CCl& k) {
  V3 rn[4], ro[4], dr[4];
  double mi[4];
  load_rel(d, k, d.ux, d.uy, d.uz, rn);
  load_rel(d, k, d.ox, d.oy, d.oz, ro);
#pragma unroll
  for (int m = 0; m < 4; m++) mi[m] = k.at[m] >= 0 ? d.mdyn[k.at[m]] : 0.0;
  if (!shake_core(d, k, rn, ro, mi, dr)) d.flags[F_SHAKE_FAIL] = 1;
  const double idt = 1.0 / d.dt;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int a = k.at[m];
    if (a < 0) continue;
    d.ux[a] += (unsigned long long)__double2ll_rn(dr[m].x * d.isc[0]);
    d.uy[a] += (unsigned long long)__double2ll_rn(dr[m].y * d.isc[1]);
    d.uz[a] += (unsigned long long)__double2ll_rn(dr[m].z * d.isc[2]);
    d.vx[a] += dr[m].x * idt;
    d.vy[a] += dr[m].y * idt;
    d.vz[a] += dr[m].z * idt;
  }
}

__global__ __launch_bounds__(128) void k_shake(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int c = (blockIdx.x - block_base(pre, nsys, s_)) * 128 + threadIdx.x;
  if (c < d.n_ccl) shake_ccl(d, load_ccl(d, c));
}

// RATTLE on one cluster's velocities v (in place), r the positions relative to
// the cluster's first atom, mi the inverse masses (0 for padding); v is left as
// it is if the constraint system is singular.  lam receives the multipliers b_j
// of v_m += mi_m sum_j b_j B_jm r_j (zero when singular).
__device__ __forceinline__ void rattle_core(const CCl& k, const V3 (&r)[4], V3 (&v)[4], const double (&mi)[4],
                                            double (&lam)[3]) {
  V3 rc[3];
  double A[3][3], b[3], B[3][4], w[3][3];
  incidence(k, B);
  couplings(B, mi, w);
#pragma unroll
  for (int j = 0; j < 3; j++) rc[j] = bsum(B[j], r);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    b[j] = j < k.nc ? -v3dot(bsum(B[j], v), rc[j]) : 0.0;
#pragma unroll
    for (int e = 0; e < 3; e++)
      A[j][e] = (j < k.nc && e < k.nc) ? v3dot(rc[j], rc[e]) * w[j][e] : (j == e ? 1.0 : 0.0);
  }
  lam[0] = lam[1] = lam[2] = 0.0;
  if (!solve3(A, b, 3)) return;
#pragma unroll
  for (int j = 0; j < 3; j++) lam[j] = j < k.nc ? b[j] : 0.0;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    V3 s = v[m];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double l = j < k.nc ? b[j] * B[j][m] * mi[m] : 0.0;
      s.x += l * rc[j].x; s.y += l * rc[j].y; s.z += l * rc[j].z;
    }
    v[m] = s;
  }
}

// RATTLE of one cluster.  wv (nullable) receives the virial of its constraint
// forces xx yy zz xy xz yz: the projection is the half kick dt/(2 m_a) G_a of
// G_a = (2/dt) sum_j b_j B_ja r_j, so sum_a r_a (x) G_a = (2/dt) sum_j b_j r_j (x) r_j
// (the constraint part of the reference's PCONSTANT virial, src/shake.cu:300-306).
__device__ __forceinline__ void rattle_ccl(const SysDev& d, const CCl& k, double* wv = nullptr) {
  V3 r[4], v[4];
  double mi[4], lam[3];
  load_rel(d, k, d.ux, d.uy, d.uz, r);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int a = k.at[m] >= 0 ? k.at[m] : k.at[0];
    v[m] = V3{d.vx[a], d.vy[a], d.vz[a]};
    mi[m] = k.at[m] >= 0 ? d.mdyn[a] : 0.0;
  }
  rattle_core(k, r, v, mi, lam);
  if (wv) {
    double B[3][4];
    incidence(k, B);
    const double s2 = 2.0 / d.dt;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const V3 c = bsum(B[j], r);
      const double l = s2 * lam[j];
      wv[0] += l * c.x * c.x; wv[1] += l * c.y * c.y; wv[2] += l * c.z * c.z;
      wv[3] += l * c.x * c.y; wv[4] += l * c.x * c.z; wv[5] += l * c.y * c.z;
    }
  }
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int a = k.at[m];
    if (a < 0) continue;
    d.vx[a] = v[m].x; d.vy[a] = v[m].y; d.vz[a] = v[m].z;
  }
}

__global__ __launch_bounds__(128) void k_rattle(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int c = (blockIdx.x - block_base(pre, nsys, s_)) * 128 + threadIdx.x;
  if (c < d.n_ccl) rattle_ccl(d, load_ccl(d, c));
}

// ---- fused step halves ---------------------------------------------------------
// Thread t of a system handles constraint cluster t (t < n_ccl: kick and drift
// of its atoms, then SHAKE) or the atom free_atoms[t - n_ccl] of no cluster.

// true when atom i moved more than half the skin since the last list build
// (the same test as k_check)
__device__ __forceinline__ bool moved_atom(const SysDev& d, int i) {
  const double dx = fx_d(d.rx[i], d.ux[i], d.sc[0]);
  const double dy = fx_d(d.ry[i], d.uy[i], d.sc[1]);
  const double dz = fx_d(d.rz[i], d.uz[i], d.sc[2]);
  return dx * dx + dy * dy + dz * dz > d.skin2;
}

// Kick, drift and SHAKE of one constraint cluster with its state in registers:
// every load is issued before the first store, so a thread waits on memory
// about twice instead of once per atom and pass.  The arithmetic is that of
// kick_drift_atom followed by shake_ccl; FINISH puts the previous step's second
// half kick and RATTLE (same forces, k_step_second's arithmetic) in front.
// Returns the list-rebuild test of the cluster's atoms.
template <bool MIXED, bool FINISH = false>
__device__ __forceinline__ bool step_first_ccl(const SysDev& d, const CCl& k) {
  unsigned long long* U[3] = {d.ux, d.uy, d.uz};
  const unsigned long long* R[3] = {d.rx, d.ry, d.rz};
  double* Vp[3] = {d.vx, d.vy, d.vz};
  int sl[4], at[4];
  double mi[4], v[4][3], f[4][3];
  unsigned long long u[4][3], o[4][3];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int a = k.at[m] >= 0 ? k.at[m] : k.at[0];  // padding repeats the first atom and is not stored
    at[m] = a;
    sl[m] = d.atom_slot[a];
    mi[m] = k.at[m] >= 0 ? d.mdyn[a] : 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) { v[m][c] = Vp[c][a]; u[m][c] = U[c][a]; }
  }
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int c = 0; c < 3; c++) f[m][c] = take_force<MIXED>(d, sl[m], at[m], c, false);
  const double h0 = 0.5 * d.dt;
  if (FINISH) {
    V3 r[4], w[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const double h = h0 * mi[m];
#pragma unroll
      for (int c = 0; c < 3; c++) v[m][c] = fma(h, f[m][c], v[m][c]);
      r[m] = V3{fx_d(u[0][0], u[m][0], d.sc[0]), fx_d(u[0][1], u[m][1], d.sc[1]), fx_d(u[0][2], u[m][2], d.sc[2])};
      w[m] = V3{v[m][0], v[m][1], v[m][2]};
    }
    double lam[3];
    rattle_core(k, r, w, mi, lam);
#pragma unroll
    for (int m = 0; m < 4; m++) { v[m][0] = w[m].x; v[m][1] = w[m].y; v[m][2] = w[m].z; }
  }
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const double h = h0 * mi[m];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      v[m][c] = fma(h, f[m][c], v[m][c]);
      o[m][c] = u[m][c];
      u[m][c] = u[m][c] + (unsigned long long)__double2ll_rn(d.dt * v[m][c] * d.isc[c]);
    }
  }
  V3 rn[4], ro[4], dr[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    rn[m] = V3{fx_d(u[0][0], u[m][0], d.sc[0]), fx_d(u[0][1], u[m][1], d.sc[1]), fx_d(u[0][2], u[m][2], d.sc[2])};
    ro[m] = V3{fx_d(o[0][0], o[m][0], d.sc[0]), fx_d(o[0][1], o[m][1], d.sc[1]), fx_d(o[0][2], o[m][2], d.sc[2])};
  }
  if (!shake_core(d, k, rn, ro, mi, dr)) d.flags[F_SHAKE_FAIL] = 1;
  const double idt = 1.0 / d.dt;
  bool move = false;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int a = k.at[m];
    if (a < 0) continue;
    const double e[3] = {dr[m].x, dr[m].y, dr[m].z};
    double r2 = 0.0;
    unsigned long long un3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const unsigned long long un = u[m][c] + (unsigned long long)__double2ll_rn(e[c] * d.isc[c]);
      un3[c] = un;
      U[c][a] = un;
      Vp[c][a] = v[m][c] + e[c] * idt;
      const double x = fx_d(R[c][a], un, d.sc[c]);
      r2 = c == 0 ? x * x : r2 + x * x;
    }
    move = move || r2 > d.skin2;
    store_slot_pos(d, sl[m], un3[0], un3[1], un3[2], d.q[a]);  // the pair kernel's slot copy
    clear_force<MIXED>(d, sl[m], a);
  }
  return move;
}

// first half kick, drift, SHAKE, and the list-rebuild test; clears the force
// and energy accumulators.  FINISH: the previous step's second half kick and
// RATTLE first (one kernel between two force evaluations).
template <bool MIXED, bool FINISH>
__global__ __launch_bounds__(128) void k_step_first(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lb = blockIdx.x - block_base(pre, nsys, s_);
  if (lb == 0 && threadIdx.x < A_NSLOT) d.acc[threadIdx.x] = 0ull;
  const int t = lb * 128 + threadIdx.x;
  bool move = false;
  if (t < d.n_ccl) {
    move = step_first_ccl<MIXED, FINISH>(d, load_ccl(d, t));
  } else if (t < d.n_ccl + d.n_free) {
    const int i = d.free_atoms[t - d.n_ccl];
    kick_drift_atom<MIXED, FINISH>(d, i);
    move = moved_atom(d, i);
  }
  if (__any(move) && (threadIdx.x & 63) == 0) d.flags[F_REBUILD] = 1;
}

// second half kick and RATTLE; the constraint virial of the RATTLE projection
// is added to the step's virial (the last step of a run, whose energies and
// virial are reported)
template <bool MIXED>
__global__ __launch_bounds__(128) void k_step_second(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int t = (blockIdx.x - block_base(pre, nsys, s_)) * 128 + threadIdx.x;
  double wv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (t < d.n_ccl) {
    const CCl k = load_ccl(d, t);
#pragma unroll
    for (int m = 0; m < 4; m++)
      if (k.at[m] >= 0) kick_atom<MIXED>(d, k.at[m]);
    rattle_ccl(d, k, wv);
  } else if (t < d.n_ccl + d.n_free) {
    kick_atom<MIXED>(d, d.free_atoms[t - d.n_ccl]);
  }
  double w9[9] = {wv[0], wv[3], wv[4], wv[3], wv[1], wv[5], wv[4], wv[5], wv[2]};
  const int slots[9] = {A_VIR + 0, A_VIR + 1, A_VIR + 2, A_VIR + 3, A_VIR + 4, A_VIR + 5, A_VIR + 6, A_VIR + 7, A_VIR + 8};
  block_acc_fixed<9>(w9, d.acc, slots);
}

void launch_step_first(const Launch& L, bool finish) {
  const dim3 g(L.w_step.total), b(128);
  if (L.mixed) {
    if (finish) hipLaunchKernelGGL((k_step_first<true, true>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
    else hipLaunchKernelGGL((k_step_first<true, false>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
  } else {
    if (finish) hipLaunchKernelGGL((k_step_first<false, true>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
    else hipLaunchKernelGGL((k_step_first<false, false>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
  }
}
void launch_step_second(const Launch& L) {
  if (L.mixed) hipLaunchKernelGGL(k_step_second<true>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
  else hipLaunchKernelGGL(k_step_second<false>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
}

// ---- Andersen barostat (NPT) -----------------------------------------------------
// Isotropic piston on the volume, V'' = (P - p_ref) / M (the piston of the
// reference's reduce_virial, src/cell_based_kernels.cu:403-600; its NPT step is
// a stub, src/npt_kernels.cu:26-33).  Scaling is molecular for the constraint
// clusters: a box change from scale l to l' moves each cluster's centre of mass
// and each free atom with the box (scaled coordinates fixed) and keeps a
// cluster's internal geometry (its fixed-point offsets from the centre of mass
// scale by l / l'), so constraints are never strained; centre-of-mass and
// free-atom velocities (peculiar) scale by l / l'.  The piston sees the
// molecular pressure (2 KE_mol + W_mol) / 3V: KE_mol of the cluster centres and
// free atoms, W_mol = tr sum r F minus sum_c sum_{i in c} (r_i - R_c) . F_i
// (constraint forces are internal to a cluster and drop out).  One step, h = dt/2:
//   k_npt_a  Pi += h (W_mol / 3V - p_ref); V1 = V + h Pi / M
//   k_npt_b  kick; box l -> l1; drift at box l1; KE_mol
//   k_npt_c  Pi += dt 2 KE_mol / 3V1; V2 = V1 + h Pi / M; box fields and PME tables at l2
//   k_npt_d  box l1 -> l2; SHAKE against the cluster geometry before the drift;
//            slot copies; list-rebuild test
//   forces; k_npt_intra (W_mol's cluster part); k_npt_g (Pi += h (W_mol / 3V2 - p_ref));
//   k_step_second (kick, RATTLE)
// Without constraint clusters each piece is the exact flow of part of the
// extended Hamiltonian KE + U + p_ref V + Pi^2 / 2M (symplectic, time reversible).

__global__ void k_npt_a(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  if (threadIdx.x != 0 || !d.npt) return;
  NptState& st = *d.nst;
  const double h = 0.5 * d.dt;
  const double Pi = st.Pi + h * (st.Wtr / (3.0 * st.V) - d.npt_pref);
  const double V1 = st.V + h * Pi / d.npt_mass;
  const double ell1 = cbrt(V1 / d.npt_V0);
  st.Pi = Pi;
  st.V1 = V1;
  st.ell0 = st.ell;
  st.ell1 = ell1;
  st.a1 = st.ell / ell1;
  st.ke1 = 0ull;
  st.wintra = 0ull;
}

// A box change by f = l_old / l_new of one cluster held in registers (m: masses,
// u: fixed-point positions, v: velocities): offsets from the centre of mass
// scale by f (their length in A stays), the centre-of-mass velocity by f.
// Returns the cluster's centre-of-mass kinetic energy after the change.
__device__ __forceinline__ double npt_cluster_box(int na, const double (&m)[4], unsigned long long (&u)[4][3],
                                                  double (&v)[4][3], double f) {
  double M = 0.0, off[3] = {0.0, 0.0, 0.0}, P[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double mk = k < na ? m[k] : 0.0;
    M += mk;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      off[c] += mk * (double)(long long)(u[k][c] - u[0][c]);
      P[c] += mk * v[k][c];
    }
  }
  double ke = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    off[c] /= M;
    const double Vc = P[c] / M;
    double dl[4];
#pragma unroll
    for (int k = 0; k < 4; k++) dl[k] = (double)(long long)(u[k][c] - u[0][c]) - off[c];  // all before any moves
#pragma unroll
    for (int k = 0; k < 4; k++) {
      u[k][c] += (unsigned long long)__double2ll_rn((f - 1.0) * dl[k]);
      v[k][c] += (f - 1.0) * Vc;
    }
    ke += 0.5 * M * (f * Vc) * (f * Vc);
  }
  return ke;
}

// kick, box change l0 -> l1 and drift at box l1 of one cluster or free atom;
// positions before the step into o; clears the step's energy accumulators
template <bool MIXED>
__global__ __launch_bounds__(128) void k_npt_b(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lb = blockIdx.x - block_base(pre, nsys, s_);
  if (lb == 0 && threadIdx.x < A_NSLOT) d.acc[threadIdx.x] = 0ull;
  const int t = lb * 128 + threadIdx.x;
  double ke[1] = {0.0};
  unsigned long long* U[3] = {d.ux, d.uy, d.uz};
  unsigned long long* O[3] = {d.ox, d.oy, d.oz};
  double* Vp[3] = {d.vx, d.vy, d.vz};
  const double f1 = d.nst->a1, ell1 = d.nst->ell1, h0 = 0.5 * d.dt;
  double isc1[3];
#pragma unroll
  for (int c = 0; c < 3; c++) isc1[c] = 18446744073709551616.0 / (ell1 * d.npt_L0[c]);
  if (t < d.n_ccl) {
    const CCl k = load_ccl(d, t);
    double m[4], v[4][3];
    unsigned long long u[4][3];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int a = k.at[j] >= 0 ? k.at[j] : k.at[0];
      const int sl = d.atom_slot[a];
      m[j] = 1.0 / d.minv[a];
      const double hm = h0 * d.mdyn[a];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        u[j][c] = U[c][a];
        v[j][c] = fma(hm, take_force<MIXED>(d, sl, a, c, false), Vp[c][a]);
      }
    }
    ke[0] = npt_cluster_box(k.na, m, u, v, f1);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int a = k.at[j];
      if (a < 0) continue;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        O[c][a] = U[c][a];
        U[c][a] = u[j][c] + (unsigned long long)__double2ll_rn(d.dt * v[j][c] * isc1[c]);
        Vp[c][a] = v[j][c];
      }
    }
  } else if (t < d.n_ccl + d.n_free) {
    const int i = d.free_atoms[t - d.n_ccl];
    const int sl = d.atom_slot[i];
    const double hm = h0 * d.mdyn[i];
    double v2 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double v = fma(hm, take_force<MIXED>(d, sl, i, c, false), Vp[c][i]) * f1;
      const unsigned long long u = U[c][i];
      O[c][i] = u;
      U[c][i] = u + (unsigned long long)__double2ll_rn(d.dt * v * isc1[c]);
      Vp[c][i] = v;
      v2 += v * v;
    }
    ke[0] = 0.5 / d.minv[i] * v2;
  }
  const int slot[1] = {0};
  block_acc_fixed<1>(ke, &d.nst->ke1, slot);
}

// PME per-axis tables m_d = j' / L_d and e^{-pi^2 m_d^2 / beta^2} at box scale ell
__device__ __forceinline__ void npt_pme_tables(const SysDev& d, double ell) {
  const int KS = d.K[0] + d.K[1] + d.K[2];
  for (int t = threadIdx.x; t < KS; t += blockDim.x) {
    const int ax = t < d.K[0] ? 0 : (t < d.K[0] + d.K[1] ? 1 : 2);
    const int j = t - (ax == 0 ? 0 : (ax == 1 ? d.K[0] : d.K[0] + d.K[1]));
    const int K = d.K[ax];
    const int jj = j <= K / 2 ? j : j - K;
    const double m = jj / (ell * d.npt_L0[ax]);
    d.pme_m[t] = m;
    d.pme_ex[t] = exp(-PI_D * PI_D * m * m / (d.beta * d.beta));
  }
}

__global__ __launch_bounds__(256) void k_npt_c(SysDev* __restrict__ S, int nsys) {
  SysDev& d = S[blockIdx.x];
  if (!d.npt) return;
  NptState& st = *d.nst;
  const double h = 0.5 * d.dt;
  const double Pi = st.Pi + d.dt * 2.0 * from_fixed_e(st.ke1) / (3.0 * st.V1);
  double V2 = st.V1 + h * Pi / d.npt_mass;
  double ell2 = cbrt(V2 / d.npt_V0);
  // below the smallest scale the search ranges cover: flagged (the call fails
  // and is restored), and the step goes on at the old box
  const bool shrink = !(ell2 >= NPT_MIN_SCALE);
  if (shrink) { V2 = st.V; ell2 = st.ell; }
  const double ell1 = st.ell1, ell_b = st.ell_b;
  npt_pme_tables(d, ell2);
  __syncthreads();  // every thread has read the state before it changes
  if (threadIdx.x == 0) {
    st.Pi = Pi;
    st.V = V2;
    st.a2 = ell1 / ell2;
    st.ell = ell2;
    npt_box_fields(d, ell2, ell_b);
    if (shrink) d.flags[F_OVERFLOW] |= OV_BOX;
  }
}

// descriptor fields and PME tables at the state's box, or at the box of the
// last list build (build_scale: a restore rebuilds that list there)
__global__ __launch_bounds__(256) void k_npt_apply(SysDev* __restrict__ S, int nsys, int build_scale) {
  SysDev& d = S[blockIdx.x];
  if (!d.npt) return;
  const double ell_b = d.nst->ell_b, ell = build_scale ? ell_b : d.nst->ell;
  npt_pme_tables(d, ell);
  if (threadIdx.x == 0) npt_box_fields(d, ell, ell_b);
}

// a list build since the state last looked moves the build scale to the
// current box (stand-alone evaluations)
__global__ void k_npt_mark(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  if (threadIdx.x != 0 || !d.npt) return;
  NptState& st = *d.nst;
  if (d.flags[F_NBUILD] != st.nbuild) {
    st.nbuild = d.flags[F_NBUILD];
    st.ell_b = st.ell;
  }
}

// box change l1 -> l2, SHAKE at the new box against the cluster geometry
// before the drift (positions o at box l0), slot copies, force clearing and the
// list-rebuild test of one constraint cluster (cf. step_first_ccl)
template <bool MIXED>
__device__ __forceinline__ bool npt_d_ccl(const SysDev& d, const CCl& k, double f2, double ell0) {
  unsigned long long* U[3] = {d.ux, d.uy, d.uz};
  const unsigned long long* O[3] = {d.ox, d.oy, d.oz};
  const unsigned long long* R[3] = {d.rx, d.ry, d.rz};
  double* Vp[3] = {d.vx, d.vy, d.vz};
  int sl[4];
  double mi[4], m[4], v[4][3];
  unsigned long long u[4][3], o[4][3];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int a = k.at[j] >= 0 ? k.at[j] : k.at[0];
    sl[j] = d.atom_slot[a];
    mi[j] = k.at[j] >= 0 ? d.mdyn[a] : 0.0;
    m[j] = 1.0 / d.minv[a];
#pragma unroll
    for (int c = 0; c < 3; c++) { v[j][c] = Vp[c][a]; u[j][c] = U[c][a]; o[j][c] = O[c][a]; }
  }
  (void)npt_cluster_box(k.na, m, u, v, f2);
  double sc0[3];
#pragma unroll
  for (int c = 0; c < 3; c++) sc0[c] = ell0 * d.npt_L0[c] / 18446744073709551616.0;
  V3 rn[4], ro[4], dr[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    rn[j] = V3{fx_d(u[0][0], u[j][0], d.sc[0]), fx_d(u[0][1], u[j][1], d.sc[1]), fx_d(u[0][2], u[j][2], d.sc[2])};
    ro[j] = V3{fx_d(o[0][0], o[j][0], sc0[0]), fx_d(o[0][1], o[j][1], sc0[1]), fx_d(o[0][2], o[j][2], sc0[2])};
  }
  if (!shake_core(d, k, rn, ro, mi, dr)) d.flags[F_SHAKE_FAIL] = 1;
  const double idt = 1.0 / d.dt;
  bool move = false;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int a = k.at[j];
    if (a < 0) continue;
    const double e[3] = {dr[j].x, dr[j].y, dr[j].z};
    double r2 = 0.0;
    unsigned long long un3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const unsigned long long un = u[j][c] + (unsigned long long)__double2ll_rn(e[c] * d.isc[c]);
      un3[c] = un;
      U[c][a] = un;
      Vp[c][a] = v[j][c] + e[c] * idt;
      const double x = fx_d(R[c][a], un, d.sc[c]);
      r2 = c == 0 ? x * x : r2 + x * x;
    }
    move = move || r2 > d.skin2;
    store_slot_pos(d, sl[j], un3[0], un3[1], un3[2], d.q[a]);
    clear_force<MIXED>(d, sl[j], a);
  }
  return move;
}

template <bool MIXED>
__global__ __launch_bounds__(128) void k_npt_d(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int t = (blockIdx.x - block_base(pre, nsys, s_)) * 128 + threadIdx.x;
  const double f2 = d.nst->a2, ell0 = d.nst->ell0;
  bool move = false;
  if (t < d.n_ccl) {
    move = npt_d_ccl<MIXED>(d, load_ccl(d, t), f2, ell0);
  } else if (t < d.n_ccl + d.n_free) {
    const int i = d.free_atoms[t - d.n_ccl];
    const int sl = d.atom_slot[i];
    d.vx[i] *= f2; d.vy[i] *= f2; d.vz[i] *= f2;
    store_slot_pos(d, sl, d.ux[i], d.uy[i], d.uz[i], d.q[i]);
    clear_force<MIXED>(d, sl, i);
    move = moved_atom(d, i);
  }
  if (__any(move) && (threadIdx.x & 63) == 0) d.flags[F_REBUILD] = 1;
}

// the clusters' internal virial sum_{i in c} (r_i - R_c) . F_i at the current
// positions and forces, into the state's wintra (zeroed at the step's start)
template <bool MIXED>
__global__ __launch_bounds__(128) void k_npt_intra(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int t = (blockIdx.x - block_base(pre, nsys, s_)) * 128 + threadIdx.x;
  double w[1] = {0.0};
  if (t < d.n_ccl) {
    const CCl k = load_ccl(d, t);
    double m[4], f[4][3], M = 0.0, Rc[3] = {0.0, 0.0, 0.0};
    V3 r[4];
    load_rel(d, k, d.ux, d.uy, d.uz, r);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int a = k.at[j] >= 0 ? k.at[j] : k.at[0];
      const int sl = d.atom_slot[a];
      m[j] = k.at[j] >= 0 ? 1.0 / d.minv[a] : 0.0;
#pragma unroll
      for (int c = 0; c < 3; c++) f[j][c] = take_force<MIXED>(d, sl, a, c, false);
      M += m[j];
      Rc[0] += m[j] * r[j].x; Rc[1] += m[j] * r[j].y; Rc[2] += m[j] * r[j].z;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) Rc[c] /= M;
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (k.at[j] >= 0) w[0] += (r[j].x - Rc[0]) * f[j][0] + (r[j].y - Rc[1]) * f[j][1] + (r[j].z - Rc[2]) * f[j][2];
  }
  const int slot[1] = {0};
  block_acc_fixed<1>(w, &d.nst->wintra, slot);
}

// piston half kick with the molecular virial trace of the step's forces; a list
// build in this step moves the build scale to the current box
__global__ void k_npt_g(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  if (threadIdx.x != 0 || !d.npt) return;
  NptState& st = *d.nst;
  const double W = from_fixed_e(d.acc[A_VIR]) + from_fixed_e(d.acc[A_VIR + 4]) + from_fixed_e(d.acc[A_VIR + 8]) -
                   from_fixed_e(st.wintra);
  st.Wtr = W;
  st.Pi = st.Pi + 0.5 * d.dt * (W / (3.0 * st.V) - d.npt_pref);
  if (d.flags[F_NBUILD] != st.nbuild) {
    st.nbuild = d.flags[F_NBUILD];
    st.ell_b = st.ell;
  }
}

// the molecular virial of a stand-alone evaluation (forces only) as the piston's W
__global__ void k_npt_zero_intra(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  if (threadIdx.x == 0 && d.npt) d.nst->wintra = 0ull;
}
__global__ void k_npt_wtr(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  if (threadIdx.x != 0 || !d.npt) return;
  d.nst->Wtr = from_fixed_e(d.acc[A_VIR]) + from_fixed_e(d.acc[A_VIR + 4]) + from_fixed_e(d.acc[A_VIR + 8]) -
               from_fixed_e(d.nst->wintra);
}

void launch_npt_a(const Launch& L) { hipLaunchKernelGGL(k_npt_a, dim3(L.nsys), dim3(64), 0, L.stream, L.d_sys, L.nsys); }
void launch_npt_b(const Launch& L) {
  if (L.mixed) hipLaunchKernelGGL(k_npt_b<true>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
  else hipLaunchKernelGGL(k_npt_b<false>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
}
void launch_npt_intra(const Launch& L) {
  if (L.mixed) hipLaunchKernelGGL(k_npt_intra<true>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
  else hipLaunchKernelGGL(k_npt_intra<false>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
}
void launch_npt_c(const Launch& L) {
  hipLaunchKernelGGL(k_npt_c, dim3(L.nsys), dim3(256), 0, L.stream, const_cast<SysDev*>(L.d_sys), L.nsys);
}
void launch_npt_apply(const Launch& L, bool build_scale) {
  hipLaunchKernelGGL(k_npt_apply, dim3(L.nsys), dim3(256), 0, L.stream, const_cast<SysDev*>(L.d_sys), L.nsys,
                     build_scale ? 1 : 0);
}
void launch_npt_mark(const Launch& L) { hipLaunchKernelGGL(k_npt_mark, dim3(L.nsys), dim3(64), 0, L.stream, L.d_sys, L.nsys); }
void launch_npt_wtr(const Launch& L) {
  hipLaunchKernelGGL(k_npt_zero_intra, dim3(L.nsys), dim3(64), 0, L.stream, L.d_sys, L.nsys);
  launch_npt_intra(L);
  hipLaunchKernelGGL(k_npt_wtr, dim3(L.nsys), dim3(64), 0, L.stream, L.d_sys, L.nsys);
}
void launch_npt_d(const Launch& L) {
  if (L.mixed) hipLaunchKernelGGL(k_npt_d<true>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
  else hipLaunchKernelGGL(k_npt_d<false>, dim3(L.w_step.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_step.pre);
}
void launch_npt_g(const Launch& L) { hipLaunchKernelGGL(k_npt_g, dim3(L.nsys), dim3(64), 0, L.stream, L.d_sys, L.nsys); }

// ---- kinetic energy, velocities ---------------------------------------------------

__global__ __launch_bounds__(256) void k_kinetic(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  double ke[1] = {0.0};
  if (i < d.n_atoms) ke[0] = 0.5 / d.minv[i] * (d.vx[i] * d.vx[i] + d.vy[i] * d.vy[i] + d.vz[i] * d.vz[i]);
  const int slots[1] = {A_KIN};
  block_acc_fixed<1>(ke, d.acc, slots);
}

__device__ __forceinline__ void philox4x32(unsigned c[4], unsigned k0, unsigned k1) {
  // Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const unsigned hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    const unsigned hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
}

// v_d = sqrt(RC T / m) N(0,1) (src/init.cu:80-92); momentum accumulated for removal
__global__ __launch_bounds__(256) void k_maxwell(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre,
                                                 double T, unsigned long long seed, unsigned long long counter) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  double p[3] = {0.0, 0.0, 0.0};
  if (i < d.n_atoms) {
    unsigned c[4] = {(unsigned)i, 0u, (unsigned)counter, (unsigned)(counter >> 32)};
    philox4x32(c, (unsigned)seed, (unsigned)(seed >> 32));
    const double s = 1.0 / 4294967296.0;
    const double u1 = ((double)c[0] + 0.5) * s, u2 = ((double)c[1] + 0.5) * s;
    const double u3 = ((double)c[2] + 0.5) * s, u4 = ((double)c[3] + 0.5) * s;
    const double m = 1.0 / d.minv[i], vm = sqrt(RCG * T / m);
    const double ra = sqrt(-2.0 * log(u1)), rb = sqrt(-2.0 * log(u3));
    const double mob = d.mdyn[i] != 0.0 ? 1.0 : 0.0;  // fixed atoms: zero velocity
    const double vx = mob * vm * ra * cos(2.0 * PI_D * u2), vy = mob * vm * ra * sin(2.0 * PI_D * u2);
    const double vz = mob * vm * rb * cos(2.0 * PI_D * u4);
    d.vx[i] = vx; d.vy[i] = vy; d.vz[i] = vz;
    p[0] = m * vx; p[1] = m * vy; p[2] = m * vz;
  }
  const int slots[3] = {A_MOM, A_MOM + 1, A_MOM + 2};
  block_acc_fixed<3>(p, d.acc, slots);
}

// v_i -= P / (N m_i) over the N mobile atoms (src/init.cu:94-118; with fixed
// atoms the reference sums over its first nAtomwoseg atoms, src/init.cu:77-127)
__global__ __launch_bounds__(256) void k_momentum(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (i >= d.n_atoms) return;
  const double f = d.n_mobile > 0 ? d.mdyn[i] / d.n_mobile : 0.0;
  d.vx[i] -= from_fixed_e(d.acc[A_MOM]) * f;
  d.vy[i] -= from_fixed_e(d.acc[A_MOM + 1]) * f;
  d.vz[i] -= from_fixed_e(d.acc[A_MOM + 2]) * f;
}

// ---- steepest-descent minimiser (replaces SteepestDescent, src/mdcuda.cu:245) ----

// largest |F|^2 per system; non-negative doubles order like their bit patterns
template <bool MIXED>
__global__ __launch_bounds__(256) void k_fmax(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  double f2 = 0.0;
  if (i < d.n_atoms && d.mdyn[i] != 0.0) {
    const int sl = d.atom_slot[i];
    const double fx = take_force<MIXED>(d, sl, i, 0, false), fy = take_force<MIXED>(d, sl, i, 1, false),
                 fz = take_force<MIXED>(d, sl, i, 2, false);
    f2 = fx * fx + fy * fy + fz * fz;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) f2 = fmax(f2, __shfl_xor(f2, o, 64));
  if ((threadIdx.x & 63) == 0 && f2 > 0.0) atomicMax(d.acc + A_FMAX, (unsigned long long)__double_as_longlong(f2));
}

// u <- u + mstep * F / max|F| (the atom with the largest force moves mstep);
// the previous positions are kept in o for SHAKE and for a rejected move
template <bool MIXED>
__global__ __launch_bounds__(256) void k_min_move(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (i >= d.n_atoms) return;
  const double f2max = __longlong_as_double((long long)d.acc[A_FMAX]);
  const double a = f2max > 0.0 ? *d.mstep / sqrt(f2max) : 0.0;
  const int sl = d.atom_slot[i];
  const unsigned long long ux = d.ux[i], uy = d.uy[i], uz = d.uz[i];
  d.ox[i] = ux; d.oy[i] = uy; d.oz[i] = uz;
  if (d.mdyn[i] == 0.0) return;  // fixed atom
  d.ux[i] = ux + (unsigned long long)__double2ll_rn(a * take_force<MIXED>(d, sl, i, 0, false) * d.isc[0]);
  d.uy[i] = uy + (unsigned long long)__double2ll_rn(a * take_force<MIXED>(d, sl, i, 1, false) * d.isc[1]);
  d.uz[i] = uz + (unsigned long long)__double2ll_rn(a * take_force<MIXED>(d, sl, i, 2, false) * d.isc[2]);
}

// The reference's minimiser moves (ConjugatedGradient / SteepestDescent,
// src/mdcuda.cu:126-330): method 1, per-atom Fletcher-Reeves direction
// H = (F + gamma H_prev) * fac with gamma = |F|^2 / |F_prev|^2 (F_prev taken as F
// when |F_prev|^2 < 1e-8); method 2, H = F * fac; method 0, no move.  u += H.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_min_ref_move(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  const int n = d.n_atoms;
  if (i >= n) return;
  const unsigned long long ux = d.ux[i], uy = d.uy[i], uz = d.uz[i];
  d.ox[i] = ux; d.oy[i] = uy; d.oz[i] = uz;
  const int method = (int)d.mstep[1];
  if (method == 0 || d.mdyn[i] == 0.0) return;  // no move / fixed atom
  const double fac = d.mstep[0];
  const int sl = d.atom_slot[i];
  const double F[3] = {take_force<MIXED>(d, sl, i, 0, false), take_force<MIXED>(d, sl, i, 1, false),
                       take_force<MIXED>(d, sl, i, 2, false)};
  double h[3];
  if (method == 1) {
    double p[3] = {d.mF[i], d.mF[n + i], d.mF[2 * n + i]};
    double p2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    const double f2 = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    if (p2 < 1e-8) p2 = f2;
    const double g = p2 > 0.0 ? f2 / p2 : 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      h[k] = (F[k] + g * d.mH[k * n + i]) * fac;
      d.mH[k * n + i] = h[k];
      d.mF[k * n + i] = F[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) h[k] = F[k] * fac;
  }
  d.ux[i] = ux + (unsigned long long)__double2ll_rn(h[0] * d.isc[0]);
  d.uy[i] = uy + (unsigned long long)__double2ll_rn(h[1] * d.isc[1]);
  d.uz[i] = uz + (unsigned long long)__double2ll_rn(h[2] * d.isc[2]);
}

void launch_min_ref_move(const Launch& L) {
  if (L.mixed) hipLaunchKernelGGL(k_min_ref_move<true>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  else hipLaunchKernelGGL(k_min_ref_move<false>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  launch_shake(L);
}

__global__ __launch_bounds__(256) void k_min_restore(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre,
                                                     const int* __restrict__ reject) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (i >= d.n_atoms || !reject[s_]) return;
  d.ux[i] = d.ox[i]; d.uy[i] = d.oy[i]; d.uz[i] = d.oz[i];
}

__global__ __launch_bounds__(256) void k_zero_vel(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (i >= d.n_atoms) return;
  d.vx[i] = 0.0; d.vy[i] = 0.0; d.vz[i] = 0.0;
}

void launch_min_move(const Launch& L) {
  if (L.mixed) {
    hipLaunchKernelGGL(k_fmax<true>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    hipLaunchKernelGGL(k_min_move<true>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  } else {
    hipLaunchKernelGGL(k_fmax<false>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    hipLaunchKernelGGL(k_min_move<false>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  }
  launch_shake(L);
}
void launch_min_restore(const Launch& L, const int* d_reject) {
  hipLaunchKernelGGL(k_min_restore, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre, d_reject);
}
void launch_zero_vel(const Launch& L) {
  hipLaunchKernelGGL(k_zero_vel, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
}

__global__ void k_zero_acc(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  if (threadIdx.x < A_NSLOT) d.acc[threadIdx.x] = 0ull;
}

// clears the slot-ordered pair and atom-ordered bonded force sums (before a
// stand-alone evaluation)
__global__ __launch_bounds__(256) void k_zero_forces(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int sl = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (sl >= d.n_slots_cap) return;
  unsigned long long* F = (unsigned long long*)d.Fs + (size_t)sl * 3;
  F[0] = 0ull;
  F[1] = 0ull;
  F[2] = 0ull;
  unsigned long long* B = (unsigned long long*)d.Fa;
  for (int a = sl; a < d.n_atoms; a += d.n_slots_cap) {
    B[a] = 0ull;
    B[d.n_atoms + a] = 0ull;
    B[2 * d.n_atoms + a] = 0ull;
  }
}

void launch_shake(const Launch& L) {
  if (L.w_ccl.total) hipLaunchKernelGGL(k_shake, dim3(L.w_ccl.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_ccl.pre);
}
void launch_rattle(const Launch& L) {
  if (L.w_ccl.total) hipLaunchKernelGGL(k_rattle, dim3(L.w_ccl.total), dim3(128), 0, L.stream, L.d_sys, L.nsys, L.w_ccl.pre);
}
void launch_kinetic(const Launch& L) {
  hipLaunchKernelGGL(k_kinetic, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
}
void launch_zero_acc(const Launch& L) {
  hipLaunchKernelGGL(k_zero_acc, dim3(L.nsys), dim3(64), 0, L.stream, L.d_sys, L.nsys);
}

__global__ void k_clear_shake_flag(const SysDev* __restrict__ S, int nsys) {
  if (threadIdx.x == 0) S[blockIdx.x].flags[F_SHAKE_FAIL] = 0;
}
void launch_clear_shake_flag(const Launch& L) {
  hipLaunchKernelGGL(k_clear_shake_flag, dim3(L.nsys), dim3(64), 0, L.stream, L.d_sys, L.nsys);
}
__global__ __launch_bounds__(256) void k_snapshot(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre,
                                                 int clear_shake) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lb = blockIdx.x - block_base(pre, nsys, s_);
  const int n = d.n_atoms;
  unsigned long long* sn = d.snap;
  if (lb == 0 && threadIdx.x < F_NSLOT) {
    int f = d.flags[threadIdx.x];
    if (clear_shake && threadIdx.x == F_SHAKE_FAIL) {
      f = 0;
      d.flags[F_SHAKE_FAIL] = 0;
    }
    ((int*)(sn + (size_t)9 * n))[threadIdx.x] = f;
  }
  if (lb == 0 && d.npt && threadIdx.x == 0) d.nst[1] = d.nst[0];
  const int i = lb * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned long long* src[9] = {d.ux, d.uy, d.uz, (const unsigned long long*)d.vx, (const unsigned long long*)d.vy,
                                      (const unsigned long long*)d.vz, d.rx, d.ry, d.rz};
#pragma unroll
  for (int k = 0; k < 9; k++) sn[(size_t)k * n + i] = src[k][i];
}

void launch_snapshot(const Launch& L, bool clear_shake) {
  hipLaunchKernelGGL(k_snapshot, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre,
                     clear_shake ? 1 : 0);
}

void launch_zero_forces(const Launch& L) {
  hipLaunchKernelGGL(k_zero_forces, dim3(L.w_slots.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_slots.pre);
}
void launch_maxwell(const Launch& L, double T, unsigned long long seed, unsigned long long counter) {
  launch_zero_acc(L);
  hipLaunchKernelGGL(k_maxwell, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre, T, seed, counter);
  hipLaunchKernelGGL(k_momentum, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  launch_rattle(L);
}

}  // namespace fz
