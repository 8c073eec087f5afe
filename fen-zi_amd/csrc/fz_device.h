// fz_device.h -- device helpers shared by the gfx950 kernels: fixed-point
// arithmetic, wave64 reductions, batch work location, precision policies.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fz_sys.h"

// global address space: pointers so qualified compile to global_* memory
// instructions instead of generic (flat) ones
#define FZ_GLOBAL __attribute__((address_space(1)))

namespace fz {

constexpr double CC = 332.0716;     // reference src/defs.in.h:157
constexpr double RCG = 0.001987;    // src/defs.in.h:156
constexpr double TU = 20.4548;      // src/defs.in.h:158
constexpr double PI_D = 3.14159265358979323846;
constexpr double TWO_OVER_SQRT_PI = 1.12837916709551257390;
// Fixed-point scales: forces and PME grid charges 2^32 per unit (range
// +-2.1e9), energies / virial / momentum 2^28 per unit (range +-3.4e10).
constexpr double ACC_SCALE = 4294967296.0;          // 2^32
constexpr double ACC_ISCALE = 1.0 / 4294967296.0;
constexpr double E_SCALE = 268435456.0;             // 2^28
constexpr double E_ISCALE = 1.0 / 268435456.0;

// ---- fixed point -----------------------------------------------------------

// Signed minimum-image displacement (u_j - u_i) in Angstrom.
__device__ __forceinline__ double fx_d(unsigned long long ui, unsigned long long uj, double sc) {
  return (double)(long long)(uj - ui) * sc;
}

__device__ __forceinline__ unsigned long long to_fixed_acc(double v) {
  return (unsigned long long)__double2ll_rn(v * ACC_SCALE);
}
__device__ __forceinline__ double from_fixed_acc(unsigned long long v) {
  return (double)(long long)v * ACC_ISCALE;
}

// Fixed point (2^32) of an FP32 value: round-to-nearest-even of v * 2^32 as an
// int64, the same value to_fixed_acc((double)v) gives.  For |v| < 2^19 the
// FP64 sum v * 2^32 + 1.5 * 2^52 is exact before its single rounding to an
// integer, which it keeps in the low mantissa bits: its bit pattern minus that
// of 1.5 * 2^52 is the result.  The same sum scaled by 2^-32, v + 1.5 * 2^20,
// has the same bits and needs no multiply (one FP64 add, one subtract).  A wave with
// a larger value (a minimiser's first steps from overlapping atoms) takes the
// general conversion.
__device__ __forceinline__ unsigned long long fixed_f32(float v) {
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(!(__builtin_fabsf(v) < 524288.0f)) == 0ull, 1)) {
    // 1.5 * 2^20 + v lies in [2^20, 2^21), where the FP64 spacing is 2^-32
    const double w = (double)v + 1572864.0;
    return (unsigned long long)__double_as_longlong(w) - 0x4138000000000000ull;
  }
  return to_fixed_acc((double)v);
}

// energy / virial / momentum accumulators
__device__ __forceinline__ void acc_add_fixed(unsigned long long* p, double v) {
  atomicAdd(p, (unsigned long long)__double2ll_rn(v * E_SCALE));
}
__device__ __forceinline__ double from_fixed_e(unsigned long long v) {
  return (double)(long long)v * E_ISCALE;
}

// ---- wave64 helpers ---------------------------------------------------------

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Largest s in [0, n) with pre[s] <= b (pre[0] = 0, non-decreasing).
// System of block b of a launch over a work table: the table holds the n + 1
// block prefixes followed by one system index per block (a single load).
// A single-system launch (n == 1) needs no table load at all.
__device__ __forceinline__ int block_sys(const int* __restrict__ pre, int n, int b) { return n == 1 ? 0 : pre[n + 1 + b]; }
// first block of system s
__device__ __forceinline__ int block_base(const int* __restrict__ pre, int n, int s) { return n == 1 ? 0 : pre[s]; }

// index of the prefix interval holding b (binary search over n + 1 prefixes)
__device__ __forceinline__ int locate(const int* __restrict__ pre, int n, int b) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pre[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// Block-wide reduction of NV doubles into fixed-point accumulators (one atomic
// per value per block).  Requires blockDim.x <= 1024 and a multiple of 64.
template <int NV>
__device__ __forceinline__ void block_acc_fixed(double (&v)[NV], unsigned long long* acc, const int* slot) {
  __shared__ double red[16][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) red[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += red[i][threadIdx.x];
    if (s != 0.0) acc_add_fixed(acc + slot[threadIdx.x], s);
  }
  __syncthreads();
}

// ---- precision policies -------------------------------------------------------
// Force accumulators live in SysDev::Fs as 3 blocks of n_slots_cap 64-bit words.
// FP64 form: native double atomics.  Mixed form: int64 fixed point (2^32 per
// kcal/mol/A), which makes every sum independent of the order of additions.

struct PrecF64 {
  using real = double;
  using acc_t = double;
  static constexpr bool fixed = false;
  __device__ static __forceinline__ void add(acc_t* p, double v) { unsafeAtomicAdd(p, v); }
  __device__ static __forceinline__ double get(const acc_t* p) { return *p; }
};

struct PrecMixed {
  using real = float;
  using acc_t = unsigned long long;
  static constexpr bool fixed = true;
  __device__ static __forceinline__ void add(acc_t* p, double v) { atomicAdd(p, to_fixed_acc(v)); }
  __device__ static __forceinline__ double get(const acc_t* p) { return from_fixed_acc(*p); }
};

// Slot-ordered copy of an atom's position and charge for the pair kernel:
// int4 {x, y, z top 32 bits, FP32 charge} in the mixed form, Slot64 in FP64.
struct Slot64 {
  unsigned long long x, y, z;
  double q;
};
__device__ __forceinline__ void store_slot_pos(const SysDev& d, int sl, unsigned long long x, unsigned long long y,
                                               unsigned long long z, double q) {
  if (d.mixed)
    ((int4*)d.xs)[sl] = make_int4((int)(unsigned)(x >> 32), (int)(unsigned)(y >> 32), (int)(unsigned)(z >> 32),
                                  __float_as_int((float)q));
  else
    ((Slot64*)d.xs)[sl] = Slot64{x, y, z, q};
}
// the whole slot (a < 0: padding): position, charge, LJ parameters, type
__device__ __forceinline__ void store_slot(const SysDev& d, int sl, int a) {
  if (a >= 0) store_slot_pos(d, sl, d.ux[a], d.uy[a], d.uz[a], d.q[a]);
  else store_slot_pos(d, sl, 0ull, 0ull, 0ull, 0.0);
  const int t = a >= 0 ? d.type[a] : 0;
  if (d.mixed) d.lp[sl] = d.lj_sep ? d.lj_ah[t] : make_float2(__int_as_float(t), 0.0f);
  d.ts[sl] = t;
}

}  // namespace fz
