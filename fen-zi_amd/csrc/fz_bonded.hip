// fz_bonded.hip -- bonds, angles + Urey-Bradley, dihedrals, impropers, the
// Ewald exclusion correction (1-2, 1-3) and 1-4 pairs, one thread per term.
//
// Replaces the reference's atom-centric bondedforce/bondedforce_c36
// (src/bonded_forces.cu:597,1384), which evaluates every term once per member
// atom (a dihedral 4x), and ewaldCorrection (src/mdcuda.cu:520).  Terms are
// evaluated once, in FP64 in both precision forms, and their forces are added
// to the atom-ordered bonded sums (fixed point in the mixed form, so the sums
// are order independent), which need no slot layout: the kernel runs beside
// the neighbour build.  Functional forms: src/bonded_forces.cu:22-392 (bond
// K(r-r0)^2; angle K(theta-theta0)^2 with cos clamped at 0.9999; UB K(S-S0)^2;
// dihedral K(1+cos(n phi - delta)); improper K psi^2) and src/mdcuda.cu:534-552
// (exclusion correction -CC qi qj erf(beta r)/r).
#include "fz_device.h"
#include "fz_launch.h"

namespace fz {

__device__ __forceinline__ void relv(const SysDev& d, int a, int b, double* v) {  // r_b - r_a
  v[0] = fx_d(d.ux[a], d.ux[b], d.sc[0]);
  v[1] = fx_d(d.uy[a], d.uy[b], d.sc[1]);
  v[2] = fx_d(d.uz[a], d.uz[b], d.sc[2]);
}
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Forces of a block's terms are first summed in LDS over the block's atom
// window (consecutive terms share atoms: chain neighbours, the members of one
// water), then added to the atom-ordered sums once per atom and component.
// A block whose terms span more than BOND_WIN atoms adds directly.  In the
// mixed form the LDS sums are fixed-point integers, so the result is the same
// bit for bit either way.
constexpr int BOND_WIN = 1024;

template <class P>
struct Sink {
  bool lds;
  int w0;
  typename P::acc_t* lacc;  // [3][BOND_WIN]
};

__device__ __forceinline__ void lds_add(unsigned long long* p, double v) { atomicAdd(p, to_fixed_acc(v)); }
__device__ __forceinline__ void lds_add(double* p, double v) { atomicAdd(p, v); }

template <class P>
__device__ __forceinline__ void addf(const Sink<P>& sk, const SysDev& d, int atom, const double* f) {
  if (sk.lds) {
    const int k = atom - sk.w0;
    lds_add(sk.lacc + k, f[0]);
    lds_add(sk.lacc + BOND_WIN + k, f[1]);
    lds_add(sk.lacc + 2 * BOND_WIN + k, f[2]);
    return;
  }
  typename P::acc_t* F = (typename P::acc_t*)d.Fa;
  const int n = d.n_atoms;
  P::add(F + atom, f[0]);
  P::add(F + n + atom, f[1]);
  P::add(F + 2 * n + atom, f[2]);
}

__device__ __forceinline__ void add_acc(unsigned long long* p, unsigned long long v) { atomicAdd(p, v); }
__device__ __forceinline__ void add_acc(double* p, double v) { unsafeAtomicAdd(p, v); }

// smallest and largest atom of flat term t (hi = -1 if t is past the last term)
__device__ __forceinline__ void term_span(const SysDev& d, int t, int& lo, int& hi) {
  int4 v = make_int4(0, 0, 0, 0);
  int na = 0;
  if (t < d.n_bonds) {
    const int2 b = d.bond_ij[t]; v = make_int4(b.x, b.y, b.y, b.y); na = 2;
  } else if ((t -= d.n_bonds) < d.n_angles) {
    v = d.angle_ijk[t]; v.w = v.z; na = 3;
  } else if ((t -= d.n_angles) < d.n_diheds + d.n_imprs) {
    v = t < d.n_diheds ? d.dihe_ijkl[t] : d.impr_ijkl[t - d.n_diheds]; na = 4;
  } else if ((t -= d.n_diheds + d.n_imprs) < d.n_ew) {
    const int2 b = d.ew_ij[t]; v = make_int4(b.x, b.y, b.y, b.y); na = 2;
  } else if ((t -= d.n_ew) < d.n_p14) {
    const int2 b = d.p14_ij[t]; v = make_int4(b.x, b.y, b.y, b.y); na = 2;
  }
  lo = na ? min(min(v.x, v.y), min(v.z, v.w)) : 0x7fffffff;
  hi = na ? max(max(v.x, v.y), max(v.z, v.w)) : -1;
}

__device__ __forceinline__ void vir_outer(double* vir, const double* r, const double* f) {
#pragma unroll
  for (int x = 0; x < 3; x++)
#pragma unroll
    for (int y = 0; y < 3; y++) vir[3 * x + y] += r[x] * f[y];
}

// dihedral angle (IUPAC sign) and gradient, Blondel & Karplus, J. Comput. Chem. 17, 1132 (1996)
__device__ __forceinline__ double dihedral_grad(const double* F, const double* G, const double* H, double g[4][3]) {
  double A[3], B[3], BxA[3];
  cross3(F, G, A);
  cross3(H, G, B);
  const double A2 = dot3(A, A), B2 = dot3(B, B), Gn = sqrt(dot3(G, G));
  cross3(B, A, BxA);
  const double iab = 1.0 / sqrt(A2 * B2);
  const double sinp = dot3(BxA, G) * iab / Gn;
  const double cosp = dot3(A, B) * iab;
  const double phi = atan2(sinp, cosp);
  const double FG = dot3(F, G), HG = dot3(H, G);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    g[0][k] = -Gn / A2 * A[k];
    g[3][k] = Gn / B2 * B[k];
    g[1][k] = Gn / A2 * A[k] + FG / (A2 * Gn) * A[k] - HG / (B2 * Gn) * B[k];
    g[2][k] = HG / (B2 * Gn) * B[k] - FG / (A2 * Gn) * A[k] - Gn / B2 * B[k];
  }
  return phi;
}

// energy slots in the per-block reduction
enum { BE_BOND = 0, BE_ANGLE, BE_UB, BE_DIHED, BE_IMPR, BE_EWEXCL, BE_VDW, BE_ELEC, BE_N };

template <class P, bool EV>
__global__ __launch_bounds__(256) void k_bonded(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  int t = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  // atom window of the block
  __shared__ typename P::acc_t lacc[3 * BOND_WIN];
  __shared__ int wlo, whi;
  if (threadIdx.x == 0) { wlo = 0x7fffffff; whi = -1; }
  __syncthreads();
  {
    int lo, hi;
    term_span(d, t, lo, hi);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      lo = min(lo, __shfl_xor(lo, o, 64));
      hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && hi >= 0) { atomicMin(&wlo, lo); atomicMax(&whi, hi); }
  }
  __syncthreads();
  Sink<P> sk{whi >= 0 && whi - wlo < BOND_WIN, wlo, lacc};
  const int span = sk.lds ? whi - wlo + 1 : 0;
  for (int k = threadIdx.x; k < span; k += 256) {
    lacc[k] = 0;
    lacc[BOND_WIN + k] = 0;
    lacc[2 * BOND_WIN + k] = 0;
  }
  __syncthreads();
  double e[BE_N + 9];
#pragma unroll
  for (int k = 0; k < BE_N + 9; k++) e[k] = 0.0;
  double* vir = e + BE_N;
  do {
    if (t < d.n_bonds) {
      const int2 ab = d.bond_ij[t];
      const double2 p = d.bond_p[t];
      double v[3];
      relv(d, ab.x, ab.y, v);
      const double r = sqrt(dot3(v, v)), dr = r - p.y;
      e[BE_BOND] += p.x * dr * dr;
      const double fr = -2.0 * p.x * dr / r;
      double fb[3] = {fr * v[0], fr * v[1], fr * v[2]}, fa[3] = {-fb[0], -fb[1], -fb[2]};
      addf<P>(sk, d, ab.y, fb);
      addf<P>(sk, d, ab.x, fa);
      if (EV) vir_outer(vir, v, fb);
      break;
    }
    t -= d.n_bonds;
    if (t < d.n_angles) {
      const int4 at = d.angle_ijk[t];
      const double4 p = d.angle_p[t];
      double v1[3], v2[3];
      relv(d, at.y, at.x, v1);
      relv(d, at.y, at.z, v2);
      const double l1 = dot3(v1, v1), l2 = dot3(v2, v2), q = 1.0 / sqrt(l1 * l2);
      double cs = dot3(v1, v2) * q;
      if (fabs(cs) > 0.9999) cs = copysign(0.9999, cs);  // src/bonded_forces.cu:81
      const double sn = sqrt(1.0 - cs * cs), dth = acos(cs) - p.y;
      e[BE_ANGLE] += p.x * dth * dth;
      const double kf = 2.0 * p.x * dth / sn;
      double fa[3], fc[3], fb[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        fa[k] = kf * (v2[k] * q - cs * v1[k] / l1);
        fc[k] = kf * (v1[k] * q - cs * v2[k] / l2);
        fb[k] = -fa[k] - fc[k];
      }
      addf<P>(sk, d, at.x, fa);
      addf<P>(sk, d, at.z, fc);
      addf<P>(sk, d, at.y, fb);
      if (EV) { vir_outer(vir, v1, fa); vir_outer(vir, v2, fc); }
      if (p.z != 0.0) {
        double w[3];
        relv(d, at.x, at.z, w);
        const double r = sqrt(dot3(w, w)), dr = r - p.w;
        e[BE_UB] += p.z * dr * dr;
        const double fr = -2.0 * p.z * dr / r;
        double f2[3] = {fr * w[0], fr * w[1], fr * w[2]}, f1[3] = {-f2[0], -f2[1], -f2[2]};
        addf<P>(sk, d, at.z, f2);
        addf<P>(sk, d, at.x, f1);
        if (EV) vir_outer(vir, w, f2);
      }
      break;
    }
    t -= d.n_angles;
    if (t < d.n_diheds + d.n_imprs) {
      const bool dih = t < d.n_diheds;
      const int tt = dih ? t : t - d.n_diheds;
      const int4 at = dih ? d.dihe_ijkl[tt] : d.impr_ijkl[tt];
      double F[3], G[3], H[3];
      relv(d, at.y, at.x, F);  // r_a - r_b
      relv(d, at.z, at.y, G);  // r_b - r_c
      relv(d, at.z, at.w, H);  // r_d - r_c
      double g[4][3];
      const double phi = dihedral_grad(F, G, H, g);
      double dEdphi;
      if (dih) {
        const double2 p = d.dihe_p[tt];
        const double n = (double)d.dihe_n[tt];
        e[BE_DIHED] += p.x * (1.0 + cos(n * phi - p.y));
        dEdphi = -p.x * n * sin(n * phi - p.y);
      } else {
        const double K = d.impr_k[tt];
        e[BE_IMPR] += K * phi * phi;
        dEdphi = 2.0 * K * phi;
      }
      double fo[4][3];
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int k = 0; k < 3; k++) fo[m][k] = -dEdphi * g[m][k];
      addf<P>(sk, d, at.x, fo[0]);
      addf<P>(sk, d, at.y, fo[1]);
      addf<P>(sk, d, at.z, fo[2]);
      addf<P>(sk, d, at.w, fo[3]);
      if (EV) {
        double ra[3] = {F[0] + G[0], F[1] + G[1], F[2] + G[2]};
        vir_outer(vir, ra, fo[0]);
        vir_outer(vir, G, fo[1]);
        vir_outer(vir, H, fo[3]);
      }
      break;
    }
    t -= d.n_diheds + d.n_imprs;
    if (t < d.n_ew) {
      const int2 ij = d.ew_ij[t];
      double v[3];
      relv(d, ij.x, ij.y, v);  // r_j - r_i
      const double r2 = dot3(v, v), r = sqrt(r2);
      const double qq = CC * d.q[ij.x] * d.q[ij.y];
      const double br = d.beta * r, erfv = erf(br);
      e[BE_EWEXCL] += -qq * erfv / r;
      const double fr = qq * (br * TWO_OVER_SQRT_PI * exp(-br * br) - erfv) / (r2 * r);
      double fj[3] = {fr * v[0], fr * v[1], fr * v[2]}, fi[3] = {-fj[0], -fj[1], -fj[2]};
      addf<P>(sk, d, ij.y, fj);
      addf<P>(sk, d, ij.x, fi);
      if (EV) vir_outer(vir, v, fj);
      break;
    }
    t -= d.n_ew;
    if (t < d.n_p14) {
      const int2 ij = d.p14_ij[t];
      double v[3];
      relv(d, ij.x, ij.y, v);
      const double r2 = dot3(v, v), r = sqrt(r2), ri2 = 1.0 / r2;
      const double2 p = d.lj14[d.type[ij.x] * d.n_types + d.type[ij.y]];
      const double s2 = p.y * ri2, s6 = s2 * s2 * s2;
      const double a6 = p.x * s6, a12 = a6 * s6;
      double elj = a12 - a6, flj = (12.0 * a12 - 6.0 * a6) * ri2;
      if (r2 > d.ron2) {
        if (r2 >= d.rc2) { elj = 0.0; flj = 0.0; }
        else {
          const double c1 = d.rc2 - r2, c2 = d.rc2 + 2.0 * r2 - 3.0 * d.ron2;
          const double sw = c1 * c1 * c2 * d.sw_inv, dsw = 12.0 * c1 * (r2 - d.ron2) * d.sw_inv;
          flj = flj * sw + elj * dsw;
          elj *= sw;
        }
      }
      const double qq = CC * d.q[ij.x] * d.q[ij.y];
      const double br = d.beta * r;
      const double ec = qq * erfc(br) / r;
      const double fc = (ec + qq * TWO_OVER_SQRT_PI * d.beta * exp(-br * br)) * ri2;
      e[BE_VDW] += elj;
      e[BE_ELEC] += ec;
      const double fr = fc + flj;  // F_i = fr (r_i - r_j)
      double fj[3] = {fr * v[0], fr * v[1], fr * v[2]}, fi[3] = {-fj[0], -fj[1], -fj[2]};
      addf<P>(sk, d, ij.y, fj);
      addf<P>(sk, d, ij.x, fi);
      if (EV) vir_outer(vir, v, fj);
      break;
    }
  } while (false);
  __syncthreads();
  if (span) {
    typename P::acc_t* F = (typename P::acc_t*)d.Fa + sk.w0;
    const int n = d.n_atoms;
    for (int k = threadIdx.x; k < span; k += 256) {
      const typename P::acc_t fx = lacc[k], fy = lacc[BOND_WIN + k], fz = lacc[2 * BOND_WIN + k];
      if (fx != 0) add_acc(F + k, fx);
      if (fy != 0) add_acc(F + n + k, fy);
      if (fz != 0) add_acc(F + 2 * n + k, fz);
    }
  }
  if (EV) {
    const int slots[BE_N + 9] = {A_BOND, A_ANGLE, A_UB, A_DIHED, A_IMPR, A_EWEXCL, A_VDW, A_ELEC,
                                 A_VIR + 0, A_VIR + 1, A_VIR + 2, A_VIR + 3, A_VIR + 4,
                                 A_VIR + 5, A_VIR + 6, A_VIR + 7, A_VIR + 8};
    block_acc_fixed<BE_N + 9>(e, d.acc, slots);
  }
}

void launch_bonded(const Launch& L, bool ev) {
  if (L.w_terms.total == 0) return;
  dim3 g(L.w_terms.total), b(256);
  if (L.mixed) {
    if (ev) hipLaunchKernelGGL((k_bonded<PrecMixed, true>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_terms.pre);
    else hipLaunchKernelGGL((k_bonded<PrecMixed, false>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_terms.pre);
  } else {
    if (ev) hipLaunchKernelGGL((k_bonded<PrecF64, true>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_terms.pre);
    else hipLaunchKernelGGL((k_bonded<PrecF64, false>), g, b, 0, L.stream, L.d_sys, L.nsys, L.w_terms.pre);
  }
}

}  // namespace fz
