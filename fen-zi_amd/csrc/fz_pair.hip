// fz_pair.hip -- real-space LJ + Ewald pair kernel over the cluster-pair list.
//
// Replaces the reference's nonbondforce (src/nonbonded_forces.cu:25-270): same
// functional forms (erfc Coulomb, CHARMM-switched LJ with the exact gradient
// of the switched energy, reference defect B1 not reproduced), Newton's third
// law instead of a full list.
//
// One wave64 evaluates one 8x8 tile (i cluster x j cluster) per list entry:
// lane l holds i = l & 7, j = l >> 3.  j forces are reduced over the 8 lanes
// that share j (xor 1, 2, 4) and added once per tile; i forces stay in
// registers until the list moves to a new i cluster, then are reduced over
// lanes l, l+8, ... (xor 8, 16, 32).  Work items are fixed-size chunks of the
// list, so the load is balanced independently of cluster density.
//
// Mixed form: int32 fixed-point displacements (the top 32 bits of the 64-bit
// fractions; int32 wrap-around is the minimum image), FP32 arithmetic, own
// FP32 erfc fit, int64 fixed-point force sums (order independent).
// FP64 form: int64 displacements, FP64 arithmetic, FP64 atomics.
#include "fz_device.h"
#include "fz_erfc_fit.h"
#include "fz_launch.h"

namespace fz {

// ---------------------------------------------------------------------------
// Slot-ordered copies of positions, charges and LJ parameters for the pair
// kernel (k_step_first and the list build keep them current in dynamics).
template <bool MIXED>
__global__ __launch_bounds__(256) void k_gather_sorted(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  const int sl = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (sl >= *d.n_clusters * CLUSTER) return;
  store_slot(d, sl, d.slot_atom[sl]);
}

// One block: the pair kernel's work items (per system, the PAIR_CHUNK-entry
// chunks of its list; none after an overflow) as a prefix over the systems.
// Systems are taken 1024 at a time, one per thread, with a block scan and a
// running carry.  MODE 0: every system, and the rebuild decisions are spent;
// MODE 1: only the systems without a pending rebuild (their lists are valid
// before the build runs); MODE 2: only the systems with one (after the build),
// and the rebuild decisions are spent.
template <int MODE>
__global__ __launch_bounds__(1024) void k_chunk_prefix(const SysDev* __restrict__ S, int nsys, int* __restrict__ chunk_pre) {
  __shared__ int part[1024];
  __shared__ int carry;
  const int t = threadIdx.x;
  if (t == 0) { carry = 0; chunk_pre[0] = 0; }
  for (int base = 0; base < nsys; base += 1024) {
    const int s = base + t;
    int c = 0;
    if (s < nsys) {
      const bool pending = S[s].flags[F_REBUILD] != 0;
      const bool take = MODE == 0 || (MODE == 1 ? !pending : pending);
      const long long np = (S[s].flags[F_OVERFLOW] || !take) ? 0 : *S[s].n_pairs;
      c = (int)((np + PAIR_CHUNK - 1) / PAIR_CHUNK);
      if (MODE != 1) S[s].flags[F_REBUILD] = 0;
    }
    part[t] = c;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
      const int v = t >= o ? part[t - o] : 0;
      __syncthreads();
      part[t] += v;
      __syncthreads();
    }
    if (s < nsys) chunk_pre[s + 1] = carry + part[t];
    __syncthreads();
    if (t == 1023) carry += part[1023];
    __syncthreads();
  }
}

void launch_chunk_prefix(const Launch& L) {
  hipLaunchKernelGGL(k_chunk_prefix<0>, dim3(1), dim3(1024), 0, L.stream, L.d_sys, L.nsys, L.chunk_pre);
}
void launch_chunk_split(const Launch& L, bool built) {
  if (built) hipLaunchKernelGGL(k_chunk_prefix<2>, dim3(1), dim3(1024), 0, L.stream, L.d_sys, L.nsys, L.chunk_pre);
  else hipLaunchKernelGGL(k_chunk_prefix<1>, dim3(1), dim3(1024), 0, L.stream, L.d_sys, L.nsys, L.chunk_pre_a);
}

void launch_gather_sorted(const Launch& L) {
  if (L.mixed)
    hipLaunchKernelGGL(k_gather_sorted<true>, dim3(L.w_slots.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_slots.pre);
  else
    hipLaunchKernelGGL(k_gather_sorted<false>, dim3(L.w_slots.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_slots.pre);
  launch_chunk_prefix(L);
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long rfl64(unsigned long long v) {
  unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// DPP lane moves within rows: quad_perm [1,0,3,2] (xor 1), quad_perm [2,3,0,1]
// (xor 2), row_half_mirror (i <-> 7 - i inside each group of 8)
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// sum over the 8 lanes sharing lane >> 3; every lane of the group gets the sum
template <class T>
__device__ __forceinline__ T red8(T v) {
  v += dpp<0xB1>(v);
  v += dpp<0x4E>(v);
  v += dpp<0x141>(v);
  return v;
}

// j forces of one tile: group g = lane >> 3 holds the sums of j slot g; lanes
// 0..23 gather (component lane >> 3, slot lane & 7) and issue one atomic each
template <class Acc, class T>
__device__ __forceinline__ void store_j(Acc* F, int J, unsigned long long mask, int lane, T jx, T jy, T jz) {
  const int c = lane >> 3, jj = lane & 7;
  const int src = jj << 3;
  const T gx = __shfl(jx, src, 64), gy = __shfl(jy, src, 64), gz = __shfl(jz, src, 64);
  const T v = c == 0 ? gx : (c == 1 ? gy : gz);
  if (lane < 24 && ((mask >> (8 * jj)) & 0xffull)) {
    Acc* p = F + (size_t)(J * CLUSTER + jj) * 3 + c;
    if constexpr (sizeof(Acc) == 8 && sizeof(T) == 4) atomicAdd((unsigned long long*)p, to_fixed_acc((double)v));
    else unsafeAtomicAdd((double*)p, (double)v);
  }
}
// Global-address-space views of the buffers: loads and atomics through them
// compile to global_* instructions (SGPR base + VGPR offset) instead of flat_*,
// which also keeps them out of the LDS counter the DPP/permute traffic waits on.
typedef FZ_GLOBAL int GI;
typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef FZ_GLOBAL v4i GI4;
typedef FZ_GLOBAL v2f GF2;
typedef FZ_GLOBAL unsigned long long GU64;

__device__ __forceinline__ void atomic_add_g(GU64* p, unsigned long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Raw buffer views of the slot arrays: a load is (SGPR descriptor, SGPR byte
// offset of the cluster, VGPR byte offset of the lane), with no 64-bit address
// arithmetic on the vector ALU.  gfx9 descriptor word 3 (raw, 32-bit format).
// On gfx9 the scalar offset is not part of the range check, so num_records
// only covers the lane offsets; the caller's indices are in range.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ v4i buf_ld4(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ v2f buf_ld2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(v2f, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

// Lanes 8k + c (c < 3) of a 64-bit lane mask whose byte k is not zero.
__device__ __forceinline__ unsigned long long byte_any3(unsigned long long a) {
  a |= a >> 4;
  a |= a >> 2;
  a |= a >> 1;
  a &= 0x0101010101010101ull;
  return a | (a << 1) | (a << 2);
}

template <class T>
__device__ __forceinline__ T red_stride8(T v) {  // sum over lanes sharing lane & 7
  v += __shfl_xor(v, 8, 64);
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// i forces of cluster I at the end of its run of entries: sums over the lanes
// sharing lane & 7; lane i + 8c (c < 3) adds component c of slot i
__device__ __forceinline__ void flush_i(FZ_GLOBAL char* Fb, int I, int lane, float fix, float fiy, float fiz) {
  const float rx = red_stride8(fix), ry = red_stride8(fiy), rz = red_stride8(fiz);
  const int c = lane >> 3;
  const float v = c == 0 ? rx : (c == 1 ? ry : rz);
  if (lane < 24)
    atomic_add_g((GU64*)(Fb + (size_t)I * (CLUSTER * 24) + (unsigned)(lane & 7) * 24u + (unsigned)c * 8u), fixed_f32(v));
}

// Sums of three values over the 8 lanes sharing lane >> 3, every lane of the
// group getting them: nine v_add_f32 with a DPP operand (xor 1, xor 2, mirror
// within 8).  Written out because the compiler fuses the first add with the
// product that feeds it (a mov_dpp + fma pair per value) and sinks the last
// stage into the store's branch (mov_dpp + add).  s_nop 1 gives the two wait
// states a DPP read needs after the VALU write of its operand; inside the block
// each read comes two instructions after the write.  Called with all lanes on.
__device__ __forceinline__ void red8x3(float& x, float& y, float& z) {
  float a, b, c;
  __asm__(
      "s_nop 1\n\t"
      "v_add_f32_dpp %0, %3, %3 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %1, %4, %4 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %2, %5, %5 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %1, %1, %1 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %2, %2, %2 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %1, %1, %1 row_half_mirror row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_add_f32_dpp %2, %2, %2 row_half_mirror row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "=&v"(a), "=&v"(b), "=&v"(c)
      : "v"(x), "v"(y), "v"(z));
  x = a;
  y = b;
  z = c;
}

__device__ __forceinline__ float erfc_fit(float x, float ex) {
  // erfc(x) = exp(-x^2) * t * P(t), t = 1 / (1 + p x); coefficients fitted for
  // relative error (fz_erfc_fit.h, generated by gen_erfc_fit.py)
  float t = __builtin_amdgcn_rcpf(1.0f + FZ_ERFC_P * x);
  float p = FZ_ERFC_C6;
  p = p * t + FZ_ERFC_C5;
  p = p * t + FZ_ERFC_C4;
  p = p * t + FZ_ERFC_C3;
  p = p * t + FZ_ERFC_C2;
  p = p * t + FZ_ERFC_C1;
  p = p * t + FZ_ERFC_C0;
  return ex * t * p;
}

template <class Acc>
__device__ __forceinline__ void add3(Acc* F, int slot, double fx, double fy, double fz);

template <>
__device__ __forceinline__ void add3<unsigned long long>(unsigned long long* F, int slot, double fx, double fy, double fz) {
  atomicAdd(F + (size_t)slot * 3, to_fixed_acc(fx));
  atomicAdd(F + (size_t)slot * 3 + 1, to_fixed_acc(fy));
  atomicAdd(F + (size_t)slot * 3 + 2, to_fixed_acc(fz));
}
template <>
__device__ __forceinline__ void add3<double>(double* F, int slot, double fx, double fy, double fz) {
  unsafeAtomicAdd(F + (size_t)slot * 3, fx);
  unsafeAtomicAdd(F + (size_t)slot * 3 + 1, fy);
  unsafeAtomicAdd(F + (size_t)slot * 3 + 2, fz);
}

// Energy / virial sums of an evaluation step.  Each lane sums one chunk in FP64
// and converts the chunk's sum to fixed point (2^28) before adding it to its
// running total, so the totals do not depend on which wave ran which chunk:
// results are the same for any persistent-grid size and any batch around the
// system.
struct EvAcc {
  long long e[8];  // vdw, elec, virial xx yy zz xy xz yz
  int npr;
  int ntl;  // tiles evaluated (the same in every lane)
  __device__ void zero() {
#pragma unroll
    for (int k = 0; k < 8; k++) e[k] = 0;
    npr = 0;
    ntl = 0;
  }
  __device__ void add_chunk(const double (&c)[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) e[k] += __double2ll_rn(c[k] * E_SCALE);
  }
  // into this wave's copy of the system's spread sums (folded by k_ev_fold)
  __device__ void flush(const SysDev& d, int lane) {
    long long v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = wave_sum(e[k]);
    const int np = wave_sum(npr);
    if (lane == 0) {
      const unsigned wv = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (EV_SPREAD - 1);
      unsigned long long* c = d.acc_ev + (size_t)wv * EV_W;
#pragma unroll
      for (int k = 0; k < 8; k++) atomicAdd(c + k, (unsigned long long)v[k]);
      atomicAdd(c + 8, (unsigned long long)np);
      atomicAdd(c + 9, (unsigned long long)ntl);
    }
  }
};

// Adds each system's EV_SPREAD copies into its accumulators (the symmetric
// virial's mirrored entries too) and clears the copies.  One block of
// EV_SPREAD threads per system; integer sums, so the order does not matter.
__global__ __launch_bounds__(EV_SPREAD) void k_ev_fold(const SysDev* __restrict__ S) {
  const SysDev& d = S[blockIdx.x];
  unsigned long long* c = d.acc_ev + (size_t)threadIdx.x * EV_W;
  long long v[EV_NV];
#pragma unroll
  for (int k = 0; k < EV_NV; k++) {
    v[k] = wave_sum((long long)c[k]);
    c[k] = 0ull;
  }
  if (threadIdx.x == 0) {
    const int slot[EV_NV] = {A_VDW, A_ELEC, A_VIR + 0, A_VIR + 4, A_VIR + 8, A_VIR + 1, A_VIR + 2, A_VIR + 5, A_NPAIR, A_NTILE};
#pragma unroll
    for (int k = 0; k < EV_NV; k++) atomicAdd(d.acc + slot[k], (unsigned long long)v[k]);
    atomicAdd(d.acc + A_VIR + 3, (unsigned long long)v[5]);
    atomicAdd(d.acc + A_VIR + 6, (unsigned long long)v[6]);
    atomicAdd(d.acc + A_VIR + 7, (unsigned long long)v[7]);
  }
}

// ---------------------------------------------------------------------------
// EV: 0 forces only; 1 also energies and virial summed from the FP32 pair terms
// (the minimisers' accept/reject tests); 2 energies recomputed in FP64 from the
// FP32 distance (reported energies).  SEP: every system of the launch follows
// the combination rule (per-slot LJ parameters, no table lookup).
template <int EV, bool SEP>
__global__ __launch_bounds__(256) void k_pair_mixed(const SysDev* __restrict__ S, int nsys, const int* __restrict__ chunk_pre,
                                                    int spend) {
  // spend: a single system's rebuild decision of this step is spent here
  if (spend && blockIdx.x == 0 && threadIdx.x == 0) S[0].flags[F_REBUILD] = 0;
  const int lane = threadIdx.x & 63, il = lane & 7, jl = lane >> 3;
  const int total = chunk_pre[nsys];
  const int nw = gridDim.x * 4;
  // energy / virial totals stay in registers until the wave moves to another system
  int evs = -1;
  EvAcc ev;
  ev.zero();
  for (int w = rfl(blockIdx.x * 4 + (threadIdx.x >> 6)); w < total; w += nw) {
    const int s = rfl(locate(chunk_pre, nsys, w));
    const SysDev& d = S[s];
    if (EV && s != evs) {
      if (evs >= 0) ev.flush(S[evs], lane);
      ev.zero();
      evs = s;
    }
    double cev[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const long long np = *d.n_pairs;
    const long long e0 = (long long)(w - chunk_pre[s]) * PAIR_CHUNK;
    const int ne = rfl((int)(np - e0 < PAIR_CHUNK ? np - e0 : PAIR_CHUNK));  // uniform: scalar loop control
    if (EV) ev.ntl += ne;
    // slot data of cluster C at (uniform base + C * stride) + per-lane offset:
    // global loads with an SGPR base and a 32-bit VGPR offset
    const unsigned cap = (unsigned)d.n_slots_cap;
    const __amdgpu_buffer_rsrc_t xsr = raw_rsrc(d.xs, cap * 16u);
    const __amdgpu_buffer_rsrc_t lpr = raw_rsrc(d.lp, cap * 8u);
    const GF2* lj = (const GF2*)d.lj32;
    const int sep = SEP || d.lj_sep;
    const int T = d.n_types;
    FZ_GLOBAL char* Fb = (FZ_GLOBAL char*)d.Fs;
    // force atomics: lane 8k + c (c < 3) adds component c of slot k of a cluster
    const unsigned foff = (unsigned)(lane >> 3) * 24u + (unsigned)(lane & 7) * 8u;
    const float sx = d.sc32[0], sy = d.sc32[1], sz = d.sc32[2];
    const float rc2 = (float)d.rc2, ron2 = (float)d.ron2;
    // CHARMM switch S(r2) = (rc2 - r2)^2 (rc2 + 2 r2 - 3 ron2) / (rc2 - ron2)^3 in the
    // form 6 S = c1^2 (sw_a r2 + sw_b), c1 = rc2 - r2, and -2 dS/dr2 = sw_d c1 (r2 - ron2);
    // r2 is clamped to ron2 below the switching range, where 6 S = 6 and dS = 0
    const float sw_a = (float)(12.0 * d.sw_inv), sw_b = (float)(6.0 * d.sw_inv * (d.rc2 - 3.0 * d.ron2));
    const float sw_d = (float)(12.0 * d.sw_inv);
    const float cc = (float)CC;
    const float pb = (float)(FZ_ERFC_P * d.beta);
    const float ccb = (float)(TWO_OVER_SQRT_PI * d.beta);
    const float nb2l2e = -(float)(d.beta * d.beta * 1.4426950408889634);
    // the chunk's entries, one per lane, read once and broadcast per entry
    int vI = 0, vJ = 0;
    unsigned vMlo = 0u, vMhi = 0u;
    if (lane < ne) {
      vI = ((const GI*)d.pl_i)[e0 + lane];
      vJ = ((const GI*)d.pl_j)[e0 + lane];
      const unsigned long long m = ((const GU64*)d.pl_mask)[e0 + lane];
      vMlo = (unsigned)m;
      vMhi = (unsigned)(m >> 32);
    }
    // j cluster data run one entry ahead of the arithmetic
    const unsigned xoff = (unsigned)jl * 16u, poff = (unsigned)jl * 8u;
    int Jn = __builtin_amdgcn_readlane(vJ, 0);
    v4i xjn = buf_ld4(xsr, xoff, (unsigned)Jn * (CLUSTER * 16));
    v2f pjn = buf_ld2(lpr, poff, (unsigned)Jn * (CLUSTER * 8));
    int curI = -1;
    int xix = 0, xiy = 0, xiz = 0, ti = 0;
    float qi = 0.f, qi_raw = 0.f;
    v2f pi = {0.f, 0.f};
    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    // The j-force atomic of a tile is issued one tile late, after the next
    // prefetch: the wait for prefetched data (vmcnt counts loads and atomics in
    // issue order) then covers only an atomic issued a whole tile earlier.
    unsigned long long pend_m = 0ull, pend_v = 0ull;
    int pend_J = 0;
    for (int k = 0; k < ne; k++) {
      const int I = __builtin_amdgcn_readlane(vI, k);
      const int J = Jn;
      const unsigned long long mask = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)vMhi, k) << 32) |
                                      (unsigned)__builtin_amdgcn_readlane((int)vMlo, k);
      const v4i xj = xjn;
      const v2f pj = pjn;
      if (k + 1 < ne) {
        Jn = __builtin_amdgcn_readlane(vJ, k + 1);
        xjn = buf_ld4(xsr, xoff, (unsigned)Jn * (CLUSTER * 16));
        pjn = buf_ld2(lpr, poff, (unsigned)Jn * (CLUSTER * 8));
      }
      if (__builtin_amdgcn_inverse_ballot_w64(pend_m))
        atomic_add_g((GU64*)(Fb + (size_t)pend_J * (CLUSTER * 24) + foff), pend_v);
      if (I != curI) {
        // new i cluster's data first, then the old one's sums, so that the wait
        // for the loads below does not include the flush's atomic
        const v4i v = buf_ld4(xsr, (unsigned)il * 16u, (unsigned)I * (CLUSTER * 16));
        const v2f pv = buf_ld2(lpr, (unsigned)il * 8u, (unsigned)I * (CLUSTER * 8));
        if (curI >= 0) flush_i(Fb, curI, lane, fix, fiy, fiz);
        curI = I;
        xix = v.x; xiy = v.y; xiz = v.z;
        qi_raw = __int_as_float(v.w);
        qi = qi_raw * cc;
        pi = pv;
        ti = __float_as_int(pi.x) * T;
        // consume the i loads here: a load still in flight where this branch
        // rejoins the loop makes the compiler wait for every memory operation
        // (the last tile's atomics and the next tile's prefetch) on each tile
        __asm__ volatile("" ::"v"(qi), "v"(xix), "v"(xiy), "v"(xiz), "v"(pi.x), "v"(pi.y));
        fix = fiy = fiz = 0.f;
      }
      const float dx = (float)(int)((unsigned)xj.x - (unsigned)xix) * sx;
      const float dy = (float)(int)((unsigned)xj.y - (unsigned)xiy) * sy;
      const float dz = (float)(int)((unsigned)xj.z - (unsigned)xiz) * sz;
      float r2 = dx * dx + dy * dy + dz * dz;
      // pairs inside the cutoff as a lane mask: lane l's bit of the entry mask is
      // tested in scalar registers, not shifted out per lane
      const unsigned long long actm = mask & __builtin_amdgcn_ballot_w64(r2 < rc2);
      const bool act = __builtin_amdgcn_inverse_ballot_w64(actm);
      r2 = act ? r2 : 1.0f;
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float rinv2 = rinv * rinv;
      const float r = r2 * rinv;
      // Lennard-Jones with CHARMM switch (energy switched, exact gradient):
      // e6 = E_LJ, f6 = r2 F_LJ / (6 r) unswitched
      v2f p;
      if (sep) {
        const float h = pi.y + pj.y;
        p = v2f{pi.x * pj.x, h * h};
      } else {
        p = lj[ti + __float_as_int(pj.x)];
      }
      const float s2 = p.y * rinv2;
      const float s6 = s2 * s2 * s2;
      const float a6 = p.x * s6, a12 = a6 * s6;
      const float e6 = a12 - a6;
      const float f6 = 2.0f * a12 - a6;
      // max(r2, ron2) of two non-negative floats as an integer max of their bits
      const float r2c = __int_as_float(max(__float_as_int(r2), __float_as_int(ron2)));
      const float c1 = rc2 - r2c;
      const float sw6 = c1 * c1 * (sw_a * r2c + sw_b);  // 6 S
      const float dsw = c1 * (r2c - ron2) * sw_d;       // -2 dS/dr2
      // real-space Ewald: erfc(beta r) = exp(-beta^2 r2) t P(t), t = 1 / (1 + p beta r)
      const float t = __builtin_amdgcn_rcpf(pb * r + 1.0f);
      const float ex = __builtin_amdgcn_exp2f(nb2l2e * r2);
      float tp = FZ_ERFC_C6;
      tp = tp * t + FZ_ERFC_C5;
      tp = tp * t + FZ_ERFC_C4;
      tp = tp * t + FZ_ERFC_C3;
      tp = tp * t + FZ_ERFC_C2;
      tp = tp * t + FZ_ERFC_C1;
      tp = tp * t + FZ_ERFC_C0;
      tp *= t;
      const float qex = qi * __int_as_float(xj.w) * ex;
      const float fc2 = (tp * rinv + ccb) * qex;  // r2 F_C / r
      // F / r = (r2 F_C / r + 6 S f6) / r2 + E_LJ (-2 dS/dr2)
      const float fs = act ? (fc2 + f6 * sw6) * rinv2 + e6 * dsw : 0.0f;
      const float fx = fs * dx, fy = fs * dy, fz = fs * dz;
      fix -= fx; fiy -= fy; fiz -= fz;
      // j sums: every lane of group j = lane >> 3 holds slot j's; lane 8j + c adds
      // component c when a pair of slot j is inside the cutoff
      float jx = fx, jy = fy, jz = fz;
      red8x3(jx, jy, jz);
      const int c = lane & 7;
      pend_v = fixed_f32(c == 0 ? jx : (c == 1 ? jy : jz));
      pend_m = byte_any3(actm);
      pend_J = J;
      if (EV == 1) {
        if (act) {
          const float e_lj = e6 * sw6 * (1.0f / 6.0f);
          const float e_c = qex * tp * rinv;
          cev[0] += (double)e_lj;
          cev[1] += (double)e_c;
          ev.npr++;
          cev[2] += (double)(fx * dx); cev[3] += (double)(fy * dy); cev[4] += (double)(fz * dz);
          cev[5] += (double)(fx * dy); cev[6] += (double)(fx * dz); cev[7] += (double)(fy * dz);
        }
      } else if (EV == 2) {
        if (act) {
          // energies of an evaluation step in FP64 from the same FP32 distance, so
          // that no FP32 bias of the erfc/exp approximations sums over millions of pairs
          const double r2d = (double)r2, rdd = sqrt(r2d), ri2d = 1.0 / r2d;
          const double s2d = (double)p.y * ri2d, s6d = s2d * s2d * s2d;
          const double a6d = (double)p.x * s6d;
          double eljd = a6d * s6d - a6d;
          if (r2d > (double)ron2) {
            const double c1 = (double)rc2 - r2d, c2 = (double)rc2 + 2.0 * r2d - 3.0 * (double)ron2;
            eljd *= c1 * c1 * c2 * d.sw_inv;
          }
          cev[0] += eljd;
          cev[1] += CC * (double)qi_raw * (double)__int_as_float(xj.w) * erfc(d.beta * rdd) / rdd;
          ev.npr++;
          cev[2] += (double)(fx * dx); cev[3] += (double)(fy * dy); cev[4] += (double)(fz * dz);
          cev[5] += (double)(fx * dy); cev[6] += (double)(fx * dz); cev[7] += (double)(fy * dz);
        }
      }
    }
    if (__builtin_amdgcn_inverse_ballot_w64(pend_m))
      atomic_add_g((GU64*)(Fb + (size_t)pend_J * (CLUSTER * 24) + foff), pend_v);
    if (curI >= 0) flush_i(Fb, curI, lane, fix, fiy, fiz);
    if (EV) ev.add_chunk(cev);
  }
  if (EV && evs >= 0) ev.flush(S[evs], lane);
}

template <bool EV>
__global__ __launch_bounds__(256) void k_pair_f64(const SysDev* __restrict__ S, int nsys, const int* __restrict__ chunk_pre,
                                                  int spend) {
  if (spend && blockIdx.x == 0 && threadIdx.x == 0) S[0].flags[F_REBUILD] = 0;
  const int lane = threadIdx.x & 63, il = lane & 7, jl = lane >> 3;
  const int total = chunk_pre[nsys];
  const int nw = gridDim.x * 4;
  int evs = -1;
  EvAcc ev;
  ev.zero();
  for (int w = rfl(blockIdx.x * 4 + (threadIdx.x >> 6)); w < total; w += nw) {
    const int s = rfl(locate(chunk_pre, nsys, w));
    const SysDev& d = S[s];
    if (EV && s != evs) {
      if (evs >= 0) ev.flush(S[evs], lane);
      ev.zero();
      evs = s;
    }
    double cev[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const long long np = *d.n_pairs;
    const long long e0 = (long long)(w - chunk_pre[s]) * PAIR_CHUNK;
    const long long e1 = e0 + PAIR_CHUNK < np ? e0 + PAIR_CHUNK : np;
    if (EV) ev.ntl += (int)(e1 - e0);
    const Slot64* __restrict__ xs = (const Slot64*)d.xs;
    const int* __restrict__ ts = d.ts;
    const double2* __restrict__ lj = d.lj64;
    const int T = d.n_types;
    double* F = (double*)d.Fs;
    const double sx = d.sc[0], sy = d.sc[1], sz = d.sc[2];
    const double rc2 = d.rc2, ron2 = d.ron2, beta = d.beta, swi = d.sw_inv;
    const double ccb = TWO_OVER_SQRT_PI * beta;
    int curI = -1;
    unsigned long long xix = 0, xiy = 0, xiz = 0;
    int ti = 0;
    double qi = 0.0;
    double fix = 0.0, fiy = 0.0, fiz = 0.0;
    for (long long e = e0; e < e1; e++) {
      const int I = rfl(d.pl_i[e]);
      const int J = rfl(d.pl_j[e]);
      const unsigned long long mask = rfl64(d.pl_mask[e]);
      if (I != curI) {
        if (curI >= 0) {
          double rx = red_stride8(fix), ry = red_stride8(fiy), rz = red_stride8(fiz);
          if (jl == 0) add3(F, curI * CLUSTER + il, rx, ry, rz);
        }
        curI = I;
        const Slot64 v = xs[I * CLUSTER + il];
        xix = v.x; xiy = v.y; xiz = v.z;
        qi = v.q * CC;
        ti = ts[I * CLUSTER + il] * T;
        fix = fiy = fiz = 0.0;
      }
      const Slot64 xj = xs[J * CLUSTER + jl];
      const int tj = ts[J * CLUSTER + jl];
      const double dx = (double)(long long)(xj.x - xix) * sx;
      const double dy = (double)(long long)(xj.y - xiy) * sy;
      const double dz = (double)(long long)(xj.z - xiz) * sz;
      double r2 = dx * dx + dy * dy + dz * dz;
      const bool act = ((mask >> lane) & 1ull) && r2 < rc2;
      r2 = act ? r2 : 1.0;
      const double r = sqrt(r2);
      const double rinv = 1.0 / r;
      const double rinv2 = rinv * rinv;
      const double2 p = lj[ti + tj];
      const double s2 = p.y * rinv2;
      const double s6 = s2 * s2 * s2;
      const double a6 = p.x * s6, a12 = a6 * s6;
      double e_lj = a12 - a6;
      double f_lj = (12.0 * a12 - 6.0 * a6) * rinv2;
      if (r2 > ron2) {
        const double c1 = rc2 - r2, c2 = rc2 + 2.0 * r2 - 3.0 * ron2;
        const double sw = c1 * c1 * c2 * swi;
        const double dsw = 12.0 * c1 * (r2 - ron2) * swi;
        f_lj = f_lj * sw + e_lj * dsw;
        e_lj *= sw;
      }
      const double qq = qi * xj.q;
      const double x = beta * r;
      const double ex = exp(-x * x);
      const double e_c = qq * erfc(x) * rinv;
      const double f_c = (e_c + qq * ccb * ex) * rinv2;
      const double fs = act ? f_c + f_lj : 0.0;
      const double fx = fs * dx, fy = fs * dy, fz = fs * dz;
      fix -= fx; fiy -= fy; fiz -= fz;
      const double jx = red8(fx), jy = red8(fy), jz = red8(fz);
      store_j(F, J, mask, lane, jx, jy, jz);
      if (EV) {
        if (act) {
          cev[0] += e_lj; cev[1] += e_c; ev.npr++;
          cev[2] += fx * dx; cev[3] += fy * dy; cev[4] += fz * dz;
          cev[5] += fx * dy; cev[6] += fx * dz; cev[7] += fy * dz;
        }
      }
    }
    if (curI >= 0) {
      double rx = red_stride8(fix), ry = red_stride8(fiy), rz = red_stride8(fiz);
      if (jl == 0) add3(F, curI * CLUSTER + il, rx, ry, rz);
    }
    if (EV) ev.add_chunk(cev);
  }
  if (EV && evs >= 0) ev.flush(S[evs], lane);
}

void launch_pair(const Launch& L, int ev, const int* cp, bool fold, bool spend) {
  const int sp = spend ? 1 : 0;
  dim3 g(L.pair_grid), b(256);
  if (!cp) cp = L.chunk_pre;
  if (L.mixed) {
    if (L.lj_sep) {
      if (ev == 2) hipLaunchKernelGGL((k_pair_mixed<2, true>), g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
      else if (ev == 1) hipLaunchKernelGGL((k_pair_mixed<1, true>), g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
      else hipLaunchKernelGGL((k_pair_mixed<0, true>), g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
    } else {
      if (ev == 2) hipLaunchKernelGGL((k_pair_mixed<2, false>), g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
      else if (ev == 1) hipLaunchKernelGGL((k_pair_mixed<1, false>), g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
      else hipLaunchKernelGGL((k_pair_mixed<0, false>), g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
    }
  } else {
    if (ev) hipLaunchKernelGGL(k_pair_f64<true>, g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
    else hipLaunchKernelGGL(k_pair_f64<false>, g, b, 0, L.stream, L.d_sys, L.nsys, cp, sp);
  }
  if (ev && fold) launch_ev_fold(L);
}

void launch_ev_fold(const Launch& L) {
  hipLaunchKernelGGL(k_ev_fold, dim3(L.nsys), dim3(EV_SPREAD), 0, L.stream, L.d_sys);
}

}  // namespace fz
