// fz_pme.hip -- smooth PME (order 4): spread, 3-D FFT, reciprocal convolution,
// force gather.
//
// Replaces the reference's lattice lists + ChargeSpread_small/_medium
// (src/cell_based_kernels.cu:1557-1765, src/pme_kernels.cu:20-380), the two
// cuFFT C2C transforms (src/mdcuda.cu:748,780), BCMultiply
// (src/cell_based_kernels.cu:1770) and PMEForce_medium/_large
// (src/pme_kernels.cu:381-773).  Differences by design: closed-form cardinal
// B-splines instead of the 1001-sample table (defect B2); atom-centric spread
// with order-independent accumulation (int64 fixed point in the mixed form);
// an own LDS FFT for 2,3,5-smooth sizes with real-to-complex / complex-to-real
// x passes, so only the half spectrum is stored and moved; the convolution,
// reciprocal energy and virial are fused into the z pass, which transforms
// forward, multiplies and transforms back while the line is in LDS.
//
// Conventions: Q(k) = sum_i q_i M4(u_x - k_x) M4(u_y - k_y) M4(u_z - k_z),
// u_d = K_d * frac_d; Qhat = unnormalised forward DFT; G(m) = CC/(pi V)
// exp(-pi^2 m^2 / beta^2) / m^2 * B(m), |b_d|^2 = 9 / (2 + cos(2 pi j / K_d))^2;
// E = 1/2 sum_m G |Qhat|^2; phi = unnormalised inverse DFT of G Qhat;
// F_i = -q_i sum_k grad(theta_i) phi(k).
#include "fz_device.h"
#include "fz_launch.h"

namespace fz {

template <class R> struct C2;
template <> struct C2<float> { using t = float2; };
template <> struct C2<double> { using t = double2; };

// The weights are computed with no contraction into fused multiply-adds, so
// that every spread path (atomic, bricks, planes) forms bitwise the same values
// whatever code the compiler inlines them into.
template <class R>
__device__ __forceinline__ void bspline4(R f, R* th, R* dth) {
#pragma clang fp contract(off)
  const R f2 = f * f, f3 = f2 * f, g = R(1) - f;
  th[0] = g * g * g * R(1.0 / 6.0);
  th[1] = (R(3) * f3 - R(6) * f2 + R(4)) * R(1.0 / 6.0);
  th[2] = (R(-3) * f3 + R(3) * f2 + R(3) * f + R(1)) * R(1.0 / 6.0);
  th[3] = f3 * R(1.0 / 6.0);
  if (dth) {
    dth[0] = R(-0.5) * g * g;
    dth[1] = R(0.5) * (R(3) * f2 - R(4) * f);
    dth[2] = R(0.5) * (R(-3) * f2 + R(2) * f + R(1));
    dth[3] = R(0.5) * f2;
  }
}

// grid index (first of 4) and spline fraction from a 64-bit fixed-point fraction
template <class R>
__device__ __forceinline__ int spline_pos(unsigned long long u, int K, R& f) {
  const unsigned long long hi = __umul64hi(u, (unsigned long long)K);
  const unsigned long long lo = u * (unsigned long long)K;
  if (sizeof(R) == 4) f = (R)(float)(unsigned)(lo >> 40) * R(5.9604644775390625e-08);  // 2^-24
  else f = (R)(double)(lo >> 11) * R(1.1102230246251565e-16);                         // 2^-53
  int g = (int)hi - 3;
  return g < 0 ? g + K : g;
}

// weight k (0..3) of the order-4 cardinal B-spline at fraction f
template <class R>
__device__ __forceinline__ R bspline4_k(R f, int k) {
#pragma clang fp contract(off)
  const R f2 = f * f, f3 = f2 * f, g = R(1) - f;
  const R w0 = g * g * g, w1 = R(3) * f3 - R(6) * f2 + R(4), w2 = R(-3) * f3 + R(3) * f2 + R(3) * f + R(1);
  const R w = k == 0 ? w0 : k == 1 ? w1 : k == 2 ? w2 : f3;
  return w * R(1.0 / 6.0);
}

// ---------------------------------------------------------------------------
// Charge spreading with global atomics, for launches over few atoms, where the
// brick pipeline's five launches cost more than the contention they remove.
// One atom per wave iteration, one of its 4x4x4 grid points per lane: the 64
// atomics of an instruction fall on 16 rows of 4 consecutive x points.  Atoms
// are visited with a multiplier coprime to their count, so that concurrent
// waves work on distant atoms; the kernel reads no cluster slots and can run
// before the list build.  The wave's 64 atoms are read and placed on the grid
// in parallel first (lane k: the k-th atom).
// Qacc must be zero before; the x pass of the FFT clears it after reading.
// The mixed-form grid is the same integer sum as the brick path's.
__device__ __forceinline__ float bcast(float v, int k) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k)); }
__device__ __forceinline__ double bcast(double v, int k) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), k), __builtin_amdgcn_readlane(__double2loint(v), k));
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_spread_atomic(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  using R = typename std::conditional<MIXED, float, double>::type;
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lane = threadIdx.x & 63;
  const int t0 = ((blockIdx.x - block_base(pre, nsys, s_)) * 4 + (threadIdx.x >> 6)) * 64;
  const int na = d.n_atoms;
  const unsigned long long mult = d.spread_mult;
  const int K0 = d.K[0], K1 = d.K[1], K2 = d.K[2];
  const int ca = lane & 3, cb = (lane >> 2) & 3, cc = lane >> 4;
  R q = R(0), fx = R(0), fy = R(0), fz = R(0);
  int gx0 = 0, gy0 = 0, gz0 = 0;
  bool has = false;
  if (t0 + lane < na) {
    const int a = (int)(((unsigned long long)(t0 + lane) * mult) % (unsigned long long)na);
    q = (R)d.q[a];
    has = q != R(0);
    gx0 = spline_pos<R>(d.ux[a], K0, fx);
    gy0 = spline_pos<R>(d.uy[a], K1, fy);
    gz0 = spline_pos<R>(d.uz[a], K2, fz);
  }
  unsigned long long bal = __ballot(has);
  while (bal) {
    const int k = __ffsll((long long)bal) - 1;
    bal &= bal - 1ull;
    int gx = __builtin_amdgcn_readlane(gx0, k) + ca;
    int gy = __builtin_amdgcn_readlane(gy0, k) + cb;
    int gz = __builtin_amdgcn_readlane(gz0, k) + cc;
    gx -= gx >= K0 ? K0 : 0;
    gy -= gy >= K1 ? K1 : 0;
    gz -= gz >= K2 ? K2 : 0;
    const R v = bcast(q, k) * bspline4_k<R>(bcast(fx, k), ca) * bspline4_k<R>(bcast(fy, k), cb) *
                bspline4_k<R>(bcast(fz, k), cc);
    const size_t g = ((size_t)gz * K1 + gy) * K0 + gx;
    if (MIXED) atomicAdd((unsigned long long*)d.Qacc + g, fixed_f32((float)v));
    else unsafeAtomicAdd((double*)d.Qacc + g, (double)v);
  }
}

// ---------------------------------------------------------------------------
// Charge spreading by bricks.  The grid is cut into bricks of 8^3 points; an
// atom belongs to the brick holding the first point of its 4^3 stencil, so its
// stencil lies inside the brick's (8+3)^3 block.  Per step:
//   count    atoms per brick (rank of each atom inside its brick),
//   scan     brick offsets (and the counts cleared for the next step),
//   scatter  per atom a record: stencil weights along x, y, z, charge, and the
//            stencil's first point inside the block,
//   spread   one workgroup per brick sums its atoms' 64 weights each into the
//            block in LDS (int64 fixed point at 2^32 per charge unit in the
//            mixed form, FP64 in the FP64 form) and stores the block,
//   reduce   every grid point adds the blocks that overlap it (at most 2 per
//            axis) into Qacc, which it overwrites.
// No global atomics: the grid is an integer sum in the mixed form, the same
// for any order of the atoms, and bitwise equal to adding the 64 weights of
// every atom into the grid one by one.

constexpr int BRK = 8;          // brick edge (grid points)
constexpr int BRH = BRK + 3;    // block edge: brick + the stencil's reach
constexpr int BRV = BRH * BRH * BRH;

// Atoms, not cluster slots, are visited, so the spread needs nothing from the
// list build.  Consecutive atoms are mostly bonded neighbours, so the lanes of
// a wave fall into few bricks: one atomic per distinct brick of the wave,
// ranks in lane order after it.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_brick_count(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  using R = typename std::conditional<MIXED, float, double>::type;
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int a = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  const bool inside = a < d.n_atoms;
  int h = -1;
  if (inside) {
    if ((R)d.q[a] != R(0)) {
      R f;
      const int gx = spline_pos<R>(d.ux[a], d.K[0], f);
      const int gy = spline_pos<R>(d.uy[a], d.K[1], f);
      const int gz = spline_pos<R>(d.uz[a], d.K[2], f);
      h = (gx >> 3) + d.nbr[0] * ((gy >> 3) + d.nbr[1] * (gz >> 3));
    }
  }
  const int lane = threadIdx.x & 63;
  int rank = 0;
  unsigned long long todo = __ballot(h >= 0);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int hl = __builtin_amdgcn_readlane(h, leader);
    const unsigned long long m = __ballot(h == hl);
    int base = 0;
    if (lane == leader) base = atomicAdd(d.br_cnt + hl, __popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);
    if (h == hl)
      rank = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    todo &= ~m;
  }
  if (inside) d.br_atom[a] = make_int2(h, rank);
}

// one block per system
__global__ __launch_bounds__(1024) void k_brick_scan(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  __shared__ int part[1024];
  const int n = d.n_bricks;
  const int per = (n + 1023) / 1024;
  const int c0 = threadIdx.x * per, c1 = min(n, c0 + per);
  int sum = 0;
  for (int c = c0; c < c1; c++) sum += d.br_cnt[c];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - sum;
  for (int c = c0; c < c1; c++) {
    d.br_off[c] = run;
    run += d.br_cnt[c];
    d.br_cnt[c] = 0;
  }
  if (threadIdx.x == 1023) d.br_off[n] = part[1023];
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_brick_scatter(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  using R = typename std::conditional<MIXED, float, double>::type;
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int a = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (a >= d.n_atoms) return;
  const int2 r = d.br_atom[a];
  if (r.x < 0) return;
  R fx, fy, fz;
  const int gx = spline_pos<R>(d.ux[a], d.K[0], fx);
  const int gy = spline_pos<R>(d.uy[a], d.K[1], fy);
  const int gz = spline_pos<R>(d.uz[a], d.K[2], fz);
  R* rec = (R*)d.br_rec + (size_t)(d.br_off[r.x] + r.y) * 16;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    rec[k] = bspline4_k<R>(fx, k);
    rec[4 + k] = bspline4_k<R>(fy, k);
    rec[8 + k] = bspline4_k<R>(fz, k);
  }
  rec[12] = (R)d.q[a];
  rec[13] = (R)((gx & 7) + BRH * ((gy & 7) + BRH * (gz & 7)));  // exact integer
}

// One workgroup per brick; a wave adds one atom's stencil per iteration, lane
// l holding point (l & 3, (l >> 2) & 3, l >> 4) of it.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_brick_spread(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using A = typename std::conditional<MIXED, unsigned long long, double>::type;
  __shared__ A blk[BRV];
  __shared__ R rec[64 * 16];
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int b = blockIdx.x - block_base(pre, nsys, s_);
  const int a0 = d.br_off[b], a1 = d.br_off[b + 1];
  if (a0 == a1) return;  // no block: the reduce skips empty bricks
  for (int i = threadIdx.x; i < BRV; i += 256) blk[i] = A(0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ca = lane & 3, cb = (lane >> 2) & 3, cc = lane >> 4;
  const int lofs = ca + BRH * (cb + BRH * cc);
  const R* src = (const R*)d.br_rec;
  for (int base = a0; base < a1; base += 64) {
    const int n = min(64, a1 - base);
    __syncthreads();  // block cleared / previous records consumed
    for (int i = threadIdx.x; i < n * 16; i += 256) rec[i] = src[(size_t)base * 16 + i];
    __syncthreads();
    for (int k = wv; k < n; k += 4) {
      const R* r = rec + k * 16;
      const R v = r[12] * r[ca] * r[4 + cb] * r[8 + cc];
      const int loc = (int)r[13] + lofs;
      if (MIXED) atomicAdd((unsigned long long*)&blk[loc], fixed_f32((float)v));
      else atomicAdd((double*)&blk[loc], (double)v);
    }
  }
  __syncthreads();
  A* out = (A*)d.br_part + (size_t)b * BRV;
  for (int i = threadIdx.x; i < BRV; i += 256) out[i] = blk[i];
}

// Block points over grid index x along one axis.  Blocks span the unwrapped
// coordinates [0, 8 nb + 3); every X = x + m K in that range is point X & 7 of
// brick X >> 3 (when that brick exists) and, when X & 7 < 3, point (X & 7) + 8
// of the brick before.  At most 4 entries for any K >= 4.
__device__ __forceinline__ int brick_cover(int x, int K, int nb, int* br, int* loc) {
  int n = 0;
  for (int X = x; X < BRK * nb + 3; X += K) {
    const int h = X >> 3, l = X & 7;
    if (h < nb) { br[n] = h; loc[n] = l; n++; }
    if (l < 3 && h >= 1) { br[n] = h - 1; loc[n] = l + BRK; n++; }
  }
  return n;
}

// One workgroup per brick, two of its 8^3 points per thread (x fastest).
template <bool MIXED>
__global__ __launch_bounds__(256) void k_brick_reduce(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  using A = typename std::conditional<MIXED, unsigned long long, double>::type;
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int K0 = d.K[0], K1 = d.K[1], K2 = d.K[2];
  const int nb0 = d.nbr[0], nb1 = d.nbr[1];
  const int b = blockIdx.x - block_base(pre, nsys, s_);
  const int bx = b % nb0, byz = b / nb0, by = byz % nb1, bz = byz / nb1;
  const A* part = (const A*)d.br_part;
  for (int t = threadIdx.x; t < BRK * BRK * BRK; t += 256) {
    const int x = BRK * bx + (t & 7), y = BRK * by + ((t >> 3) & 7), z = BRK * bz + (t >> 6);
    if (x >= K0 || y >= K1 || z >= K2) continue;
    int cbx[4], clx[4], cby[4], cly[4], cbz[4], clz[4];
    const int nx = brick_cover(x, K0, nb0, cbx, clx);
    const int ny = brick_cover(y, K1, nb1, cby, cly);
    const int nz = brick_cover(z, K2, d.nbr[2], cbz, clz);
    A sum = A(0);
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          const int c = cbx[i] + nb0 * (cby[j] + nb1 * cbz[k]);
          if (d.br_off[c + 1] == d.br_off[c]) continue;
          sum += part[(size_t)c * BRV + clx[i] + BRH * (cly[j] + BRH * clz[k])];
        }
    ((A*)d.Qacc)[((size_t)z * K1 + y) * K0 + x] = sum;
  }
}

void launch_spread(const Launch& L) {
  dim3 b(256);
  if (!L.spread_bricks) {
    if (L.mixed) hipLaunchKernelGGL(k_spread_atomic<true>, dim3(L.w_atoms.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    else hipLaunchKernelGGL(k_spread_atomic<false>, dim3(L.w_atoms.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    return;
  }
  if (L.mixed) {
    hipLaunchKernelGGL(k_brick_count<true>, dim3(L.w_atoms.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    hipLaunchKernelGGL(k_brick_scan, dim3(L.nsys), dim3(1024), 0, L.stream, L.d_sys, L.nsys);
    hipLaunchKernelGGL(k_brick_scatter<true>, dim3(L.w_atoms.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    hipLaunchKernelGGL(k_brick_spread<true>, dim3(L.w_bricks.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_bricks.pre);
    hipLaunchKernelGGL(k_brick_reduce<true>, dim3(L.w_bricks.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_bricks.pre);
  } else {
    hipLaunchKernelGGL(k_brick_count<false>, dim3(L.w_atoms.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    hipLaunchKernelGGL(k_brick_scan, dim3(L.nsys), dim3(1024), 0, L.stream, L.d_sys, L.nsys);
    hipLaunchKernelGGL(k_brick_scatter<false>, dim3(L.w_atoms.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    hipLaunchKernelGGL(k_brick_spread<false>, dim3(L.w_bricks.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_bricks.pre);
    hipLaunchKernelGGL(k_brick_reduce<false>, dim3(L.w_bricks.total), b, 0, L.stream, L.d_sys, L.nsys, L.w_bricks.pre);
  }
}

// ---------------------------------------------------------------------------
// LDS line FFTs (Stockham autosort, radix 4/2/3/5)

template <class C>
__device__ __forceinline__ C cmul(C a, C b) { C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r; }
template <class C>
__device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class C>
__device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <class C, class R>
__device__ __forceinline__ C cscale(C a, R s) { C r; r.x = a.x * s; r.y = a.y * s; return r; }
template <class C, class R>
__device__ __forceinline__ C cmuli(C a, R dir) { C r; r.x = -dir * a.y; r.y = dir * a.x; return r; }  // (dir * i) * a

template <class C, class R>
__device__ __forceinline__ void dft_small(C* v, int Rdx, R dir) {
  if (Rdx == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  } else if (Rdx == 4) {
    C a = v[0], b = v[1], c = v[2], e = v[3];
    C s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, e), s3 = cmuli(csub(b, e), dir);
    v[0] = cadd(s0, s2); v[2] = csub(s0, s2);
    v[1] = cadd(s1, s3); v[3] = csub(s1, s3);
  } else if (Rdx == 3) {
    const R h = R(0.86602540378443864676);
    C a = v[0], t1 = cadd(v[1], v[2]), t2 = cmuli(csub(v[1], v[2]), dir * h);
    C m = csub(a, cscale(t1, R(0.5)));
    v[0] = cadd(a, t1); v[1] = cadd(m, t2); v[2] = csub(m, t2);
  } else {  // 5
    const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
    const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
    C a = v[0];
    C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    C a1 = cadd(a, cadd(cscale(t1, c1), cscale(t2, c2)));
    C a2 = cadd(a, cadd(cscale(t1, c2), cscale(t2, c1)));
    C b1 = cmuli(cadd(cscale(t3, s1), cscale(t4, s2)), dir);
    C b2 = cmuli(csub(cscale(t3, s2), cscale(t4, s1)), dir);
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(a1, b1); v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2); v[3] = csub(a2, b2);
  }
}

// nl lines of length K stored line-major in src; returns the buffer holding the
// (unnormalised) transform.  dir = -1 forward, +1 inverse.  tw[k] = e^{-2 pi i k/K}.
// floor(t / d) for t < 2^16 and 2 <= d <= 4096 as mulhi(t, floor((2^32 - 1) / d) + 1)
__device__ __forceinline__ int fdiv(int t, int d, unsigned m) { return d == 1 ? t : (int)__umulhi((unsigned)t, m); }
__device__ __forceinline__ unsigned fdiv_magic(int d) { return d == 1 ? 0u : 0xFFFFFFFFu / (unsigned)d + 1u; }

// tw: the axis' twiddles, staged in LDS by the caller
template <class C, class R>
__device__ C* fft_lines(C* src, C* dst, int K, int nl, const C* tw, R dir) {
  int Ns = 1, rem = K;
  while (rem > 1) {
    const int Rdx = (rem % 4 == 0) ? 4 : (rem % 2 == 0) ? 2 : (rem % 3 == 0) ? 3 : 5;
    const int nb = K / Rdx;
    const int step = K / (Ns * Rdx);
    const unsigned m_nb = fdiv_magic(nb), m_ns = fdiv_magic(Ns);
    for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
      const int line = fdiv(t, nb, m_nb), j = t - line * nb;
      const C* in = src + line * K;
      C* out = dst + line * K;
      const int jq = fdiv(j, Ns, m_ns);
      const int k = j - jq * Ns;
      C v[5];
#pragma unroll
      for (int r = 0; r < 5; r++)
        if (r < Rdx) v[r] = in[j + r * nb];
#pragma unroll
      for (int r = 1; r < 5; r++)
        if (r < Rdx) {
          C w = tw[k * r * step];
          w.y *= -dir;  // forward: e^{-i..}; inverse: conjugate
          v[r] = cmul(v[r], w);
        }
      dft_small<C, R>(v, Rdx, dir);
      const int base = jq * Ns * Rdx + k;
#pragma unroll
      for (int r = 0; r < 5; r++)
        if (r < Rdx) out[base + r * Ns] = v[r];
    }
    __syncthreads();
    C* t = src; src = dst; dst = t;
    Ns *= Rdx;
    rem /= Rdx;
  }
  return src;
}

// the axis' twiddles (K entries) into LDS after the two line buffers
template <class C>
__device__ __forceinline__ const C* stage_tw(C* lds, const C* __restrict__ tw, int K) {
  for (int k = threadIdx.x; k < K; k += blockDim.x) lds[k] = tw[k];
  return lds;
}

// x pass, forward: real grid -> half spectrum (cleared after reading when asked)
template <bool MIXED>
__global__ __launch_bounds__(256) void k_fft_x_r2c(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre, int NL,
                                                    int clear) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using C = typename C2<R>::t;
  extern __shared__ char smem[];
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lb = blockIdx.x - block_base(pre, nsys, s_);
  const int K0 = d.K[0], nlines = d.K[1] * d.K[2], Kxh = d.Kxh;
  C* b0 = (C*)smem;
  C* b1 = b0 + NL * K0;
  const size_t l0 = (size_t)lb * NL;
  for (int e = threadIdx.x; e < NL * K0; e += 256) {
    const int line = e / K0, x = e - line * K0;
    C v; v.x = R(0); v.y = R(0);
    if (l0 + line < (size_t)nlines) {
      const size_t g = (l0 + line) * K0 + x;
      if (MIXED) {
        unsigned long long* Q = (unsigned long long*)d.Qacc;
        v.x = (R)from_fixed_acc(Q[g]);
        if (clear) Q[g] = 0ull;
      } else {
        double* Q = (double*)d.Qacc;
        v.x = (R)Q[g];
        if (clear) Q[g] = 0.0;
      }
    }
    b0[e] = v;
  }
  const C* twl = stage_tw(b1 + NL * K0, (const C*)d.tw, K0);
  __syncthreads();
  C* res = fft_lines<C, R>(b0, b1, K0, NL, twl, R(-1));
  C* Qc = (C*)d.Qc;
  for (int e = threadIdx.x; e < NL * Kxh; e += 256) {
    const int line = e / Kxh, kx = e - line * Kxh;
    if (l0 + line < (size_t)nlines) Qc[(l0 + line) * Kxh + kx] = res[line * K0 + kx];
  }
}

// y pass (either direction): lines along y for NL consecutive kx at fixed z
template <bool MIXED>
__global__ __launch_bounds__(256) void k_fft_y(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre, int NL, int dirv) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using C = typename C2<R>::t;
  extern __shared__ char smem[];
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lb = blockIdx.x - block_base(pre, nsys, s_);
  const int K1 = d.K[1], Kxh = d.Kxh;
  const int ng = (Kxh + NL - 1) / NL;
  const int z = lb / ng, kx0 = (lb - z * ng) * NL;
  C* b0 = (C*)smem;
  C* b1 = b0 + NL * K1;
  C* Qc = (C*)d.Qc;
  for (int e = threadIdx.x; e < NL * K1; e += 256) {
    const int y = e / NL, l = e - y * NL;
    C v; v.x = R(0); v.y = R(0);
    if (kx0 + l < Kxh) v = Qc[((size_t)z * K1 + y) * Kxh + kx0 + l];
    b0[l * K1 + y] = v;
  }
  const C* twl = stage_tw(b1 + NL * K1, (const C*)d.tw + d.K[0], K1);
  __syncthreads();
  C* res = fft_lines<C, R>(b0, b1, K1, NL, twl, R(dirv));
  for (int e = threadIdx.x; e < NL * K1; e += 256) {
    const int y = e / NL, l = e - y * NL;
    if (kx0 + l < Kxh) Qc[((size_t)z * K1 + y) * Kxh + kx0 + l] = res[l * K1 + y];
  }
}

// G(m) of grid point (kx, y, z) and |m|^2 from the per-axis tables, with no
// contraction into fused multiply-adds: the table below and the z pass form
// bitwise the same value.  (kx, y, z) = 0 is not evaluated here.
__device__ __forceinline__ double pme_G(const double* __restrict__ mt, const double* __restrict__ ex,
                                        const double* __restrict__ bsp, double pre, int K0, int K1, int kx, int y, int z,
                                        double& m2) {
#pragma clang fp contract(off)
  const double mx = mt[kx], my = mt[K0 + y], mz = mt[K0 + K1 + z];
  m2 = mx * mx + my * my + mz * mz;
  return pre * ex[kx] * ex[K0 + y] * ex[K0 + K1 + z] * bsp[kx] * bsp[K0 + y] * bsp[K0 + K1 + z] / m2;
}

// G(m) over the half spectrum ([z][y][kx], G(0) = 0) for a fixed box, once
// per system: the z pass reads it instead of forming it per point
__global__ __launch_bounds__(256) void k_pme_gtable(double* __restrict__ G, const double* __restrict__ mt,
                                                    const double* __restrict__ ex, const double* __restrict__ bsp,
                                                    double pre, int K0, int K1, int K2, int Kxh) {
  const size_t n = (size_t)K2 * K1 * Kxh;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const int kx = (int)(i % Kxh), y = (int)((i / Kxh) % K1), z = (int)(i / ((size_t)Kxh * K1));
    double m2;
    G[i] = (kx == 0 && y == 0 && z == 0) ? 0.0 : pme_G(mt, ex, bsp, pre, K0, K1, kx, y, z, m2);
  }
}

void launch_pme_gtable(const SysDev& d, hipStream_t st) {
  const size_t n = (size_t)d.K[2] * d.K[1] * d.Kxh;
  const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_pme_gtable, dim3(blocks), dim3(256), 0, st, d.pme_g, d.pme_m, d.pme_ex, d.pme_bsp, d.pme_pre,
                     d.K[0], d.K[1], d.K[2], d.Kxh);
}

// z pass: forward FFT, multiply by G(m) (mode 0) or 1 (mode 1), energy and
// virial, inverse FFT.  G comes from the system's table when it has one (fixed
// box) and is formed per point otherwise (barostat: the box changes every step).
template <bool MIXED, bool EV>
__global__ __launch_bounds__(256) void k_fft_z_conv(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre, int NL, int mode) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using C = typename C2<R>::t;
  extern __shared__ char smem[];
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lb = blockIdx.x - block_base(pre, nsys, s_);
  const int K0 = d.K[0], K1 = d.K[1], K2 = d.K[2], Kxh = d.Kxh;
  const int ng = (Kxh + NL - 1) / NL;
  const int y = lb / ng, kx0 = (lb - y * ng) * NL;
  C* b0 = (C*)smem;
  C* b1 = b0 + NL * K2;
  C* Qc = (C*)d.Qc;
  const size_t zs = (size_t)K1 * Kxh;
  for (int e = threadIdx.x; e < NL * K2; e += 256) {
    const int z = e / NL, l = e - z * NL;
    C v; v.x = R(0); v.y = R(0);
    if (kx0 + l < Kxh) v = Qc[(size_t)z * zs + (size_t)y * Kxh + kx0 + l];
    b0[l * K2 + z] = v;
  }
  const C* twz = stage_tw(b1 + NL * K2, (const C*)d.tw + K0 + K1, K2);
  __syncthreads();
  C* res = fft_lines<C, R>(b0, b1, K2, NL, twz, R(-1));
  double ev[10];
#pragma unroll
  for (int k = 0; k < 10; k++) ev[k] = 0.0;
  for (int e = threadIdx.x; e < NL * K2; e += 256) {
    const int l = e / K2, z = e - l * K2;
    const int kx = kx0 + l;
    if (kx >= Kxh) continue;
    C v = res[e];
    double G = 1.0;
    if (mode == 0) {
      if (kx == 0 && y == 0 && z == 0) {
        G = 0.0;
      } else if (!EV && d.pme_g) {
        G = d.pme_g[((size_t)z * K1 + y) * Kxh + kx];
      } else {
        double m2;
        G = pme_G(d.pme_m, d.pme_ex, d.pme_bsp, d.pme_pre, K0, K1, kx, y, z, m2);
        if (EV) {
          const double mx = d.pme_m[kx], my = d.pme_m[K0 + y], mz = d.pme_m[K0 + K1 + z];
          const double wgt = (kx == 0 || 2 * kx == K0) ? 1.0 : 2.0;
          const double Em = 0.5 * wgt * G * ((double)v.x * v.x + (double)v.y * v.y);
          const double fac = 2.0 * (1.0 + d.pi2_beta2 * m2) / m2;
          ev[0] += Em;
          const double mm[3] = {mx, my, mz};
#pragma unroll
          for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) ev[1 + 3 * a + b] += Em * ((a == b ? 1.0 : 0.0) - fac * mm[a] * mm[b]);
        }
      }
    }
    v.x = (R)(v.x * G);
    v.y = (R)(v.y * G);
    res[e] = v;
  }
  __syncthreads();
  C* other = res == b0 ? b1 : b0;
  C* res2 = fft_lines<C, R>(res, other, K2, NL, twz, R(1));
  for (int e = threadIdx.x; e < NL * K2; e += 256) {
    const int z = e / NL, l = e - z * NL;
    if (kx0 + l < Kxh) Qc[(size_t)z * zs + (size_t)y * Kxh + kx0 + l] = res2[l * K2 + z];
  }
  if (EV && mode == 0) {
    const int slots[10] = {A_PME, A_VIR + 0, A_VIR + 1, A_VIR + 2, A_VIR + 3, A_VIR + 4,
                           A_VIR + 5, A_VIR + 6, A_VIR + 7, A_VIR + 8};
    block_acc_fixed<10>(ev, d.acc, slots);
  }
}

// x pass, inverse: half spectrum -> real potential grid
template <bool MIXED>
__global__ __launch_bounds__(256) void k_fft_x_c2r(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre, int NL) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using C = typename C2<R>::t;
  extern __shared__ char smem[];
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int lb = blockIdx.x - block_base(pre, nsys, s_);
  const int K0 = d.K[0], nlines = d.K[1] * d.K[2], Kxh = d.Kxh;
  C* b0 = (C*)smem;
  C* b1 = b0 + NL * K0;
  const C* Qc = (const C*)d.Qc;
  const size_t l0 = (size_t)lb * NL;
  for (int e = threadIdx.x; e < NL * K0; e += 256) {
    const int line = e / K0, k = e - line * K0;
    C v; v.x = R(0); v.y = R(0);
    if (l0 + line < (size_t)nlines) {
      if (k < Kxh) v = Qc[(l0 + line) * Kxh + k];
      else { v = Qc[(l0 + line) * Kxh + (K0 - k)]; v.y = -v.y; }
    }
    b0[e] = v;
  }
  const C* twl = stage_tw(b1 + NL * K0, (const C*)d.tw, K0);
  __syncthreads();
  C* res = fft_lines<C, R>(b0, b1, K0, NL, twl, R(1));
  R* phi = (R*)d.phi;
  for (int e = threadIdx.x; e < NL * K0; e += 256) {
    const int line = e / K0;
    if (l0 + line < (size_t)nlines) phi[l0 * K0 + e] = res[e].x;
  }
}

// ---------------------------------------------------------------------------
// Plane kernels, for systems of up to a few ten thousand atoms whose grid
// plane fits in LDS: one workgroup per z plane.
//
// Forward: the plane's charge sum, then its x (real to half spectrum) and y
// transforms, all in LDS.  Each workgroup reads every atom's z position and
// spreads the atoms whose stencil covers its plane (16 points each) with LDS
// atomics, as int64 fixed point in the mixed form: the same terms and integer
// sums as the global-atomic and brick spreads, so the plane holds the same
// values, and the line transforms are those of the x and y passes, so the half
// spectrum is the same to the bit.  No grid accumulator is read or cleared.
// Inverse: y, then x (half spectrum to real) transforms of the plane into the
// potential grid, again the passes' arithmetic.
// LDS: plane_lds_bytes (the charge sum lives in the second buffer until it is read).
// Workgroups of NT threads: 1,024 where a launch holds few planes (the atom
// scan of one workgroup per plane needs many waves to hide its load latency
// behind), 256 where it holds many (a batch of small systems: more workgroups
// per CU).
constexpr int PLANE_BATCH = 8;

template <bool MIXED, int NT>
__global__ __launch_bounds__(NT) void k_pme_plane_fwd(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using C = typename C2<R>::t;
  using A = typename std::conditional<MIXED, unsigned long long, double>::type;
  extern __shared__ char smem[];
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int z = blockIdx.x - block_base(pre, nsys, s_);
  const int K0 = d.K[0], K1 = d.K[1], K2 = d.K[2], Kxh = d.Kxh;
  const int KP = K0 * K1;
  C* b0 = (C*)smem;
  C* b1 = b0 + KP;
  C* twx = b1 + KP;
  C* twy = twx + K0;
  A* acc = (A*)b1;
  for (int e = threadIdx.x; e < KP; e += NT) acc[e] = A(0);
  for (int k = threadIdx.x; k < K0 + K1; k += NT) twx[k] = ((const C*)d.tw)[k];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int h = lane >> 4, ca = lane & 3, cb = (lane >> 2) & 3;
  const int na = d.n_atoms;
  // PLANE_BATCH atoms per thread at a time: their z positions, then the rest
  // of the atoms on this plane, are loaded together (two memory round trips
  // per batch rather than two per atom)
  for (int base = 0; base < na; base += NT * PLANE_BATCH) {
    unsigned long long uz[PLANE_BATCH];
#pragma unroll
    for (int k = 0; k < PLANE_BATCH; k++) {
      const int a = base + k * NT + threadIdx.x;
      uz[k] = a < na ? d.uz[a] : 0ull;
    }
    int cz[PLANE_BATCH];
    R fz[PLANE_BATCH];
#pragma unroll
    for (int k = 0; k < PLANE_BATCH; k++) {
      const int a = base + k * NT + threadIdx.x;
      const int gz0 = spline_pos<R>(uz[k], K2, fz[k]);
      int cc = z - gz0;
      cc += cc < 0 ? K2 : 0;
      cz[k] = a < na && cc < 4 ? cc : -1;
    }
    double qd[PLANE_BATCH];
    unsigned long long ux[PLANE_BATCH], uy[PLANE_BATCH];
#pragma unroll
    for (int k = 0; k < PLANE_BATCH; k++) {
      const int a = base + k * NT + threadIdx.x;
      qd[k] = 0.0;
      ux[k] = uy[k] = 0ull;
      if (cz[k] >= 0) {
        qd[k] = d.q[a];
        ux[k] = d.ux[a];
        uy[k] = d.uy[a];
      }
    }
#pragma unroll
    for (int k = 0; k < PLANE_BATCH; k++) {
      const R q = (R)qd[k];
      const bool hit = cz[k] >= 0 && q != R(0);
      R fx = R(0), fy = R(0), wz = R(0);
      int gx0 = 0, gy0 = 0;
      if (hit) {
        gx0 = spline_pos<R>(ux[k], K0, fx);
        gy0 = spline_pos<R>(uy[k], K1, fy);
        wz = bspline4_k<R>(fz[k], cz[k]);
      }
      unsigned long long m = __ballot(hit);
      while (m) {
        // four atoms at a time: lanes 16 h + p take the h-th, grid point (p & 3, p >> 2) of its plane rows
        int src = -1;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int b = m ? __ffsll((long long)m) - 1 : -1;
          if (b >= 0) m &= m - 1ull;
          src = h == j ? b : src;
        }
        const int sl = src < 0 ? 0 : src;
        const R qs = __shfl(q, sl, 64), fxs = __shfl(fx, sl, 64), fys = __shfl(fy, sl, 64), wzs = __shfl(wz, sl, 64);
        int gx = __shfl(gx0, sl, 64) + ca, gy = __shfl(gy0, sl, 64) + cb;
        if (src >= 0) {
          gx -= gx >= K0 ? K0 : 0;
          gy -= gy >= K1 ? K1 : 0;
          const R v = qs * bspline4_k<R>(fxs, ca) * bspline4_k<R>(fys, cb) * wzs;
          if (MIXED) atomicAdd((unsigned long long*)&acc[gy * K0 + gx], fixed_f32((float)v));
          else atomicAdd((double*)&acc[gy * K0 + gx], (double)v);
        }
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < KP; e += NT) {
    C v;
    if (MIXED) v.x = (R)from_fixed_acc(((const unsigned long long*)acc)[e]);
    else v.x = (R)((const double*)acc)[e];
    v.y = R(0);
    b0[e] = v;
  }
  __syncthreads();
  // x lines (one per y), then y lines (one per kept kx)
  C* r = fft_lines<C, R>(b0, b1, K0, K1, twx, R(-1));
  C* o = r == b0 ? b1 : b0;
  for (int e = threadIdx.x; e < Kxh * K1; e += NT) {
    const int y = e / Kxh, kx = e - y * Kxh;
    o[kx * K1 + y] = r[y * K0 + kx];
  }
  __syncthreads();
  C* r2 = fft_lines<C, R>(o, r, K1, Kxh, twy, R(-1));
  C* Qc = (C*)d.Qc + (size_t)z * K1 * Kxh;
  for (int e = threadIdx.x; e < Kxh * K1; e += NT) {
    const int y = e / Kxh, kx = e - y * Kxh;
    Qc[e] = r2[kx * K1 + y];
  }
}

template <bool MIXED, int NT>
__global__ __launch_bounds__(NT) void k_pme_plane_inv(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using C = typename C2<R>::t;
  extern __shared__ char smem[];
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int z = blockIdx.x - block_base(pre, nsys, s_);
  const int K0 = d.K[0], K1 = d.K[1], Kxh = d.Kxh;
  const int KP = K0 * K1;
  C* b0 = (C*)smem;
  C* b1 = b0 + KP;
  C* twx = b1 + KP;
  C* twy = twx + K0;
  const C* Qc = (const C*)d.Qc + (size_t)z * K1 * Kxh;
  for (int e = threadIdx.x; e < Kxh * K1; e += NT) {
    const int y = e / Kxh, kx = e - y * Kxh;
    b0[kx * K1 + y] = Qc[e];
  }
  for (int k = threadIdx.x; k < K0 + K1; k += NT) twx[k] = ((const C*)d.tw)[k];
  __syncthreads();
  C* r = fft_lines<C, R>(b0, b1, K1, Kxh, twy, R(1));
  C* o = r == b0 ? b1 : b0;
  for (int e = threadIdx.x; e < KP; e += NT) {
    const int y = e / K0, k = e - y * K0;
    C v;
    if (k < Kxh) v = r[k * K1 + y];
    else { v = r[(K0 - k) * K1 + y]; v.y = -v.y; }
    o[e] = v;
  }
  __syncthreads();
  C* r2 = fft_lines<C, R>(o, r, K0, K1, twx, R(1));
  R* phi = (R*)d.phi + (size_t)z * KP;
  for (int e = threadIdx.x; e < KP; e += NT) phi[e] = r2[e].x;
}

// Force gather into the atom-ordered gather term (a store: each atom has one):
// it needs no slot layout, so it runs beside the neighbour build.  The sum over
// the 4x4x4 stencil is formed per stencil plane c (16 potential values each),
// and the planes' partial sums are added as (p0 + p1) + (p2 + p3), with no
// contraction into fused multiply-adds.  LPA lanes per atom: with 4 (launches
// over few atoms, latency bound) lane c forms plane c's sum and the quad adds
// them by lane exchange; with 1 (many atoms) one lane forms all four.  Both
// give the same bits, so a system's trajectory does not depend on the batch
// around it.
template <bool MIXED, int LPA>
__global__ __launch_bounds__(256) void k_pme_gather(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
#pragma clang fp contract(off)
  using R = typename std::conditional<MIXED, float, double>::type;
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int a = (blockIdx.x - block_base(pre, nsys, s_)) * (256 / LPA) + (int)(threadIdx.x / LPA);
  const int c0 = (int)(threadIdx.x % LPA);
  if (a >= d.n_atoms) return;  // the atom's lanes together
  const double qd = d.q[a];
  const int n = d.n_atoms;
  if (qd == 0.0) {
    if (c0 == 0) {
      if (MIXED) {
        unsigned long long* F = (unsigned long long*)d.Fg;
        F[a] = 0ull; F[n + a] = 0ull; F[2 * n + a] = 0ull;
      } else {
        double* F = (double*)d.Fg;
        F[a] = 0.0; F[n + a] = 0.0; F[2 * n + a] = 0.0;
      }
    }
    return;  // the atom's lanes together
  }
  const int K0 = d.K[0], K1 = d.K[1], K2 = d.K[2];
  R fx, fy, fz, tx[4], ty[4], tz[4], dx[4], dy[4], dz[4];
  const int gx = spline_pos<R>(d.ux[a], K0, fx);
  const int gy = spline_pos<R>(d.uy[a], K1, fy);
  const int gz = spline_pos<R>(d.uz[a], K2, fz);
  bspline4<R>(fx, tx, dx);
  bspline4<R>(fy, ty, dy);
  bspline4<R>(fz, tz, dz);
  const R* phi = (const R*)d.phi;
  int ix[4];
#pragma unroll
  for (int k = 0; k < 4; k++) ix[k] = gx + k >= K0 ? gx + k - K0 : gx + k;
  // per plane partial sums (this lane's planes)
  R p0[4 / LPA], p1[4 / LPA], p2[4 / LPA];
#pragma unroll
  for (int j = 0; j < 4 / LPA; j++) {
    const int c = j * LPA + c0;
    const R tzc = LPA == 1 ? tz[j] : (c == 0 ? tz[0] : c == 1 ? tz[1] : c == 2 ? tz[2] : tz[3]);
    const R dzc = LPA == 1 ? dz[j] : (c == 0 ? dz[0] : c == 1 ? dz[1] : c == 2 ? dz[2] : dz[3]);
    const int z = gz + c >= K2 ? gz + c - K2 : gz + c;
    R g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int y = gy + b >= K1 ? gy + b - K1 : gy + b;
      const R* row = phi + ((size_t)z * K1 + y) * K0;
      R sx = 0, sdx = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const R p = row[ix[k]];
        sx += tx[k] * p;
        sdx += dx[k] * p;
      }
      g0 += sdx * ty[b] * tzc;
      g1 += sx * dy[b] * tzc;
      g2 += sx * ty[b] * dzc;
    }
    p0[j] = g0; p1[j] = g1; p2[j] = g2;
  }
  R g0, g1, g2;
  if (LPA == 4) {
    g0 = p0[0]; g1 = p1[0]; g2 = p2[0];
    g0 += __shfl_xor(g0, 1, 64); g1 += __shfl_xor(g1, 1, 64); g2 += __shfl_xor(g2, 1, 64);
    g0 += __shfl_xor(g0, 2, 64); g1 += __shfl_xor(g1, 2, 64); g2 += __shfl_xor(g2, 2, 64);
    if (c0 != 0) return;
  } else {
    g0 = (p0[0] + p0[1 % (4 / LPA)]) + (p0[2 % (4 / LPA)] + p0[3 % (4 / LPA)]);
    g1 = (p1[0] + p1[1 % (4 / LPA)]) + (p1[2 % (4 / LPA)] + p1[3 % (4 / LPA)]);
    g2 = (p2[0] + p2[1 % (4 / LPA)]) + (p2[2 % (4 / LPA)] + p2[3 % (4 / LPA)]);
  }
  const double fxv = -qd * (double)g0 * K0 / d.L[0];
  const double fyv = -qd * (double)g1 * K1 / d.L[1];
  const double fzv = -qd * (double)g2 * K2 / d.L[2];
  if (MIXED) {
    unsigned long long* F = (unsigned long long*)d.Fg;
    F[a] = to_fixed_acc(fxv);
    F[n + a] = to_fixed_acc(fyv);
    F[2 * n + a] = to_fixed_acc(fzv);
  } else {
    double* F = (double*)d.Fg;
    F[a] = fxv;
    F[n + a] = fyv;
    F[2 * n + a] = fzv;
  }
}

template <bool MIXED>
static void launch_gather(const Launch& L) {
  // four lanes per atom up to GATHER_QUAD_BLOCKS blocks of 256 atoms per launch
  if (L.w_atoms.total <= GATHER_QUAD_BLOCKS)
    hipLaunchKernelGGL((k_pme_gather<MIXED, 4>), dim3(L.w_atoms64.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms64.pre);
  else
    hipLaunchKernelGGL((k_pme_gather<MIXED, 1>), dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
}

template <bool MIXED>
static void launch_fft_passes(const Launch& L, bool ev, int mode) {
  const size_t lds = L.fft_lds;
  const int NL = L.fft_nl;
  // the atomic spread adds into a zero grid: the x pass clears it after reading
  hipLaunchKernelGGL(k_fft_x_r2c<MIXED>, dim3(L.w_fftx.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_fftx.pre, NL,
                     L.spread_bricks ? 0 : 1);
  hipLaunchKernelGGL(k_fft_y<MIXED>, dim3(L.w_ffty.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_ffty.pre, NL, -1);
  if (ev) hipLaunchKernelGGL((k_fft_z_conv<MIXED, true>), dim3(L.w_fftz.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_fftz.pre, NL, mode);
  else hipLaunchKernelGGL((k_fft_z_conv<MIXED, false>), dim3(L.w_fftz.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_fftz.pre, NL, mode);
  hipLaunchKernelGGL(k_fft_y<MIXED>, dim3(L.w_ffty.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_ffty.pre, NL, 1);
  hipLaunchKernelGGL(k_fft_x_c2r<MIXED>, dim3(L.w_fftx.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_fftx.pre, NL);
}

template <bool MIXED, int NT>
static void launch_planes(const Launch& L, int part, bool ev) {
  if (part == 0) {
    hipLaunchKernelGGL((k_pme_plane_fwd<MIXED, NT>), dim3(L.w_planes.total), dim3(NT), L.plane_lds, L.stream, L.d_sys, L.nsys,
                       L.w_planes.pre);
    return;
  }
  const size_t lds = L.fft_lds;
  const int NL = L.fft_nl;
  if (ev) hipLaunchKernelGGL((k_fft_z_conv<MIXED, true>), dim3(L.w_fftz.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_fftz.pre, NL, 0);
  else hipLaunchKernelGGL((k_fft_z_conv<MIXED, false>), dim3(L.w_fftz.total), dim3(256), lds, L.stream, L.d_sys, L.nsys, L.w_fftz.pre, NL, 0);
  hipLaunchKernelGGL((k_pme_plane_inv<MIXED, NT>), dim3(L.w_planes.total), dim3(NT), L.plane_lds, L.stream, L.d_sys, L.nsys,
                     L.w_planes.pre);
}

void launch_pme_part(const Launch& L, int part, bool ev) {
  if (L.w_fftx.total == 0) return;
  if (part < 2 && L.pme_planes) {
    const bool wide = L.w_planes.total < PLANE_WIDE_BELOW;
    if (L.mixed) wide ? launch_planes<true, 1024>(L, part, ev) : launch_planes<true, 256>(L, part, ev);
    else wide ? launch_planes<false, 1024>(L, part, ev) : launch_planes<false, 256>(L, part, ev);
  } else if (part == 0) {
    launch_spread(L);
  } else if (part == 1) {
    if (L.mixed) launch_fft_passes<true>(L, ev, 0);
    else launch_fft_passes<false>(L, ev, 0);
  } else {
    if (L.mixed) launch_gather<true>(L);
    else launch_gather<false>(L);
  }
}

void launch_pme_fft_only(const Launch& L, int conv_mode) {
  if (L.mixed) launch_fft_passes<true>(L, true, conv_mode);
  else launch_fft_passes<false>(L, true, conv_mode);
}

}  // namespace fz
