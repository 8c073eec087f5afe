// fz_nblist.hip -- cell (column) binning and cluster-pair neighbour list.
//
// Replaces the reference's CellBuild/CellUpdate/CellClean and
// nbbuild_exclbitvec/nbbuild_excllist (src/cell_based_kernels.cu:198-399,
// 1048-1493) and the host-side rebuild branch of SingleStep
// (src/dynamics.cu:727-737).  Design:
//   * positions are 64-bit fixed-point fractions of the box, so binning is an
//     exact integer mulhi and every test below is bit-reproducible;
//   * atoms are binned into x-y columns, sorted by (z, atom id) inside each
//     column, and cut into clusters of 8 consecutive atoms;
//   * a cluster pair (I <= J) enters the list iff the periodic gap between the
//     two clusters' bounding boxes is below the pair-list distance; the gap is
//     formed in integers and squared in FP64 without contraction;
//   * each list entry carries a 64-bit mask, one bit per lane of the 8x8 wave64
//     pair tile, clearing padding, the lower triangle of self pairs and
//     excluded (1-2, 1-3, 1-4) pairs;
//   * the rebuild decision is taken on the device (displacement since the
//     last build > half the skin), so a step never waits on the host.
#include "fz_device.h"
#include "fz_launch.h"

namespace fz {

#define SYS_FROM(pre)                                  \
  const int s_ = block_sys(pre, nsys, blockIdx.x);        \
  const SysDev& d = S[s_];                             \
  const int lb = blockIdx.x - block_base(pre, nsys, s_);                 \
  (void)lb;

// ---------------------------------------------------------------------------
// rebuild trigger: any atom moved more than half the skin since the last build
// (reference CoordsUpdate, src/mdcuda.cu:453-458)
__global__ __launch_bounds__(256) void k_check(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  SYS_FROM(pre);
  const int i = lb * 256 + threadIdx.x;
  bool move = false;
  if (i < d.n_atoms) {
    double dx = fx_d(d.rx[i], d.ux[i], d.sc[0]);
    double dy = fx_d(d.ry[i], d.uy[i], d.sc[1]);
    double dz = fx_d(d.rz[i], d.uz[i], d.sc[2]);
    move = dx * dx + dy * dy + dz * dz > d.skin2;
  }
  if (__any(move) && (threadIdx.x & 63) == 0) d.flags[F_REBUILD] = 1;
}

__device__ __forceinline__ bool skip_build(const SysDev& d) {
  return d.flags[F_REBUILD] == 0 || d.flags[F_OVERFLOW] != 0;
}

// Last block done: each of the system's nblocks blocks takes a ticket after a
// fence once its writes are out; the block that draws the last ticket sees
// every other block's writes and runs the system's single-block step (no block
// ever waits for another).  The counter is left at zero for the next launch.
__device__ __forceinline__ bool last_block(unsigned* ctr, int nblocks) {
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    last = atomicAdd(ctr, 1u) == (unsigned)(nblocks - 1);
    if (last) atomicExch(ctr, 0u);
  }
  __syncthreads();
  if (last) __threadfence();
  return last != 0;
}

__device__ __forceinline__ int column_of(const SysDev& d, int i) {
  int cx = (int)__umul64hi(d.ux[i], (unsigned long long)d.ncx);
  int cy = (int)__umul64hi(d.uy[i], (unsigned long long)d.ncy);
  return cy * d.ncx + cx;
}

// Column histograms are formed in LDS per block of 256 atoms (consecutive atoms
// are mostly neighbours, so a block touches few columns) and added to the
// global counts once per touched column: the global atomics no longer queue
// up by the hundred on each column's counter.  Systems with more columns than
// the LDS table holds use global atomics directly.
constexpr int BIN_LDS = 4096;

// clusters per column, exclusive scan, capacity checks (one block of 256 threads)
__device__ void col_scan_block(const SysDev& d) {
  constexpr int NT = 256;
  __shared__ int part[NT];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const int n = d.ncol;
  const int per = (n + NT - 1) / NT;
  const int c0 = threadIdx.x * per, c1 = min(n, c0 + per);
  int sum = 0;
  for (int c = c0; c < c1; c++) {
    int cnt = d.col_count[c];
    if (cnt > MAX_COLUMN_ATOMS) bad = 1;
    sum += (cnt + CLUSTER - 1) / CLUSTER;
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 1; o < NT; o <<= 1) {  // Hillis-Steele inclusive scan
    int v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - sum;
  for (int c = c0; c < c1; c++) {
    d.col_clu0[c] = run;
    d.col_fill[c] = 0;
    run += (d.col_count[c] + CLUSTER - 1) / CLUSTER;
  }
  if (threadIdx.x == NT - 1) {
    int total = part[NT - 1];
    d.col_clu0[n] = total;
    *d.n_clusters = total;
    int ov = 0;
    if (bad) ov |= OV_COLUMN;
    if (total > d.n_clusters_cap || total * CLUSTER > d.n_slots_cap) ov |= OV_SLOTS;
    if (ov) atomicOr(d.flags + F_OVERFLOW, ov);
  }
}

// Binning count.  FUSED: the block that finishes last (ticket) also runs the
// column scan.  Every block then needs a device-scope release fence, which
// writes back its L2 and costs of the order of 0.5 us per block and XCD, so
// launches of many blocks use the separate k_col_scan instead.
template <bool FUSED>
__global__ __launch_bounds__(256) void k_bin_count(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  SYS_FROM(pre);
  if (skip_build(d)) return;
  __shared__ int hist[BIN_LDS];
  const bool lds = d.ncol <= BIN_LDS;  // uniform over the block
  const int i = lb * 256 + threadIdx.x;
  if (lds) {
    for (int k = threadIdx.x; k < d.ncol; k += 256) hist[k] = 0;
    __syncthreads();
  }
  if (i < d.n_atoms) {
    const int c = column_of(d, i);
    d.atom_col[i] = c;
    if (lds) atomicAdd(hist + c, 1);
    else atomicAdd(d.col_count + c, 1);
  }
  // clear the per-tile entry counts of the coming pair search
  const int ntile = (d.n_clusters_cap + 63) >> 6, nb = pre[s_ + 1] - pre[s_];
  for (int k = lb * 256 + threadIdx.x; k < ntile; k += nb * 256) d.tile_cnt[k] = 0;
  if (lds) {
    __syncthreads();
    for (int k = threadIdx.x; k < d.ncol; k += 256) {
      const int v = hist[k];
      if (v) atomicAdd(d.col_count + k, v);
    }
  }
  if (FUSED && last_block(d.tick + 0, pre[s_ + 1] - pre[s_])) col_scan_block(d);
}

// one block per system
__global__ __launch_bounds__(256) void k_col_scan(const SysDev* __restrict__ S, int nsys) {
  const SysDev& d = S[blockIdx.x];
  if (skip_build(d)) return;
  col_scan_block(d);
}

// places each atom in its column's range of bin_tmp (order inside a column is
// arbitrary here; k_col_sort fixes it)
__global__ __launch_bounds__(256) void k_bin_scatter(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  SYS_FROM(pre);
  if (skip_build(d)) return;
  __shared__ int hist[BIN_LDS];
  const bool lds = d.ncol <= BIN_LDS;
  const int i = lb * 256 + threadIdx.x;
  if (lds) {
    for (int k = threadIdx.x; k < d.ncol; k += 256) hist[k] = 0;
    __syncthreads();
  }
  int c = -1, r = 0;
  if (i < d.n_atoms) {
    c = d.atom_col[i];
    r = lds ? atomicAdd(hist + c, 1) : atomicAdd(d.col_fill + c, 1);
  }
  if (lds) {
    __syncthreads();
    for (int k = threadIdx.x; k < d.ncol; k += 256) {
      const int v = hist[k];
      if (v) hist[k] = atomicAdd(d.col_fill + k, v);  // this block's base in column k
    }
    __syncthreads();
    if (c >= 0) r += hist[c];
  }
  if (c >= 0) d.bin_tmp[d.col_clu0[c] * CLUSTER + r] = i;
}

// one block per column: sort by (u_z, id), lay out slots, snapshot positions,
// cluster bounding boxes
__global__ __launch_bounds__(256) void k_col_sort(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  SYS_FROM(pre);
  if (skip_build(d)) return;
  const int c = lb;
  __shared__ unsigned long long key[MAX_COLUMN_ATOMS];
  __shared__ int ids[MAX_COLUMN_ATOMS];
  const int n = d.col_count[c];
  const int clu0 = d.col_clu0[c];
  const int nclu = d.col_clu0[c + 1] - clu0;
  const int base = clu0 * CLUSTER;
  int P = 1;
  while (P < n) P <<= 1;
  for (int k = threadIdx.x; k < P; k += 256) {
    if (k < n) {
      int a = d.bin_tmp[base + k];
      ids[k] = a;
      key[k] = d.uz[a];
    } else {
      ids[k] = 0x7fffffff;
      key[k] = ~0ull;
    }
  }
  __syncthreads();
  // bitonic sort of (key, id) ascending
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < P / 2; t += 256) {
        int lo = 2 * t - (t & (stride - 1));
        int hi = lo + stride;
        bool up = (lo & size) == 0;
        unsigned long long ka = key[lo], kb = key[hi];
        int ia = ids[lo], ib = ids[hi];
        bool gt = ka > kb || (ka == kb && ia > ib);
        if (gt == up) {
          key[lo] = kb; key[hi] = ka;
          ids[lo] = ib; ids[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
  const int nsc = d.n_slots_cap;
  for (int k = threadIdx.x; k < nclu * CLUSTER; k += 256) {
    int a = k < n ? ids[k] : -1;
    d.slot_atom[base + k] = a;
    unsigned long long x = 0, y = 0, z = 0;
    if (a >= 0) {
      d.atom_slot[a] = base + k;
      x = d.ux[a];
      y = d.uy[a];
      z = key[k];
      d.rx[a] = x;
      d.ry[a] = y;
      d.rz[a] = z;
    }
    d.su[base + k] = x;
    d.su[nsc + base + k] = y;
    d.su[2 * nsc + base + k] = z;
    store_slot(d, base + k, a);  // the pair kernel's copy (kept current by k_step_first)
  }
  for (int m = threadIdx.x; m < nclu; m += 256) {
    unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0, 0, 0};
    for (int k = m * CLUSTER; k < min(n, (m + 1) * CLUSTER); k++) {
      int a = ids[k];
      unsigned long long v[3] = {d.ux[a], d.uy[a], key[k]};
#pragma unroll
      for (int e = 0; e < 3; e++) {
        lo[e] = v[e] < lo[e] ? v[e] : lo[e];
        hi[e] = v[e] > hi[e] ? v[e] : hi[e];
      }
    }
    unsigned long long* b = d.bb + 6 * (size_t)(clu0 + m);
#pragma unroll
    for (int e = 0; e < 3; e++) { b[2 * e] = lo[e]; b[2 * e + 1] = hi[e]; }
    d.clu_col[clu0 + m] = c;
  }
  // z-window tables: the clusters' z bounds are non-decreasing along the column
  // (atoms are sorted by z), so each table entry is a binary search
  const int nzb = d.nzb;
  for (int bz = threadIdx.x; bz <= nzb; bz += 256) {
    int lo = 0, hi = nclu;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const unsigned long long zu = key[min(mid * CLUSTER + CLUSTER - 1, n - 1)];
      if ((int)__umul64hi(zu, (unsigned long long)nzb) >= bz) hi = mid; else lo = mid + 1;
    }
    d.zt_hi[(size_t)c * (nzb + 1) + bz] = lo;
    lo = 0; hi = nclu;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const unsigned long long zl = key[mid * CLUSTER];
      if ((int)__umul64hi(zl, (unsigned long long)nzb) >= bz) hi = mid; else lo = mid + 1;
    }
    d.zt_lo[(size_t)c * (nzb + 1) + bz] = lo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    d.col_count[c] = 0;
    d.col_fill[c] = 0;
  }
}

// ---------------------------------------------------------------------------
// periodic gap between [alo, ahi] and [blo, bhi] on the 2^64 circle
__device__ __forceinline__ unsigned long long axis_gap(unsigned long long alo, unsigned long long ahi,
                                                       unsigned long long blo, unsigned long long bhi) {
  if (alo <= bhi && blo <= ahi) return 0ull;
  unsigned long long d1 = blo - ahi, d2 = alo - bhi;
  return d1 < d2 ? d1 : d2;
}

__device__ __forceinline__ bool bbox_within(const unsigned long long* a, const unsigned long long* b, const SysDev& d) {
#pragma clang fp contract(off)
  double gx = (double)axis_gap(a[0], a[1], b[0], b[1]) * d.sc[0];
  double gy = (double)axis_gap(a[2], a[3], b[2], b[3]) * d.sc[1];
  double gz = (double)axis_gap(a[4], a[5], b[4], b[5]) * d.sc[2];
  double d2 = gx * gx + gy * gy;
  d2 = d2 + gz * gz;
  return d2 < d.rl2;
}

// conservative (never larger than exact) periodic gap in Angstrom between an
// interval [a0, a1] and column index c of width w on a circle of length L
__device__ __forceinline__ double col_gap(double a0, double a1, int c, double w, double L) {
  double b0 = c * w - 1e-9, b1 = (c + 1) * w + 1e-9;
  if (a0 <= b1 && b0 <= a1) return 0.0;
  double d1 = b0 - a1, d2 = a0 - b1;
  d1 -= floor(d1 / L) * L;
  d2 -= floor(d2 / L) * L;
  return fmax(0.0, fmin(d1, d2) - 1e-9);
}

__device__ __forceinline__ bool is_excluded(const SysDev& d, int a, int b) {
  if (!d.use_excl_list) {
    int off = b - a;
    if (off < -32 || off > 31) return false;
    return (d.excl_bits[a] >> (off + 32)) & 1ull;
  }
  int lo = d.excl_off[a], hi = d.excl_off[a + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    int v = d.excl_list[mid];
    if (v < b) lo = mid + 1; else hi = mid;
  }
  return lo < d.excl_off[a + 1] && d.excl_list[lo] == b;
}

// squared distance of two atoms from their fixed-point positions, rounded
// exactly as the oracle does (no contraction)
__device__ __forceinline__ double r2_exact(const SysDev& d, unsigned long long ax, unsigned long long ay,
                                           unsigned long long az, unsigned long long bx, unsigned long long by,
                                           unsigned long long bz) {
#pragma clang fp contract(off)
  const double dx = (double)(long long)(bx - ax) * d.sc[0];
  const double dy = (double)(long long)(by - ay) * d.sc[1];
  const double dz = (double)(long long)(bz - az) * d.sc[2];
  double r2 = dx * dx + dy * dy;
  r2 = r2 + dz * dz;
  return r2;
}

// exclusion test with atom a's 64-bit window word already loaded
__device__ __forceinline__ bool excluded_w(const SysDev& d, unsigned long long wa, int a, int b) {
  if (!d.use_excl_list) {
    const int off = b - a;
    return off >= -32 && off <= 31 && ((wa >> (off + 32)) & 1ull);
  }
  return is_excluded(d, a, b);
}

// FP32 atom-pair distance of the list test: 32-bit top words of the fixed-point
// positions, products and sums rounded one by one (no contraction), so the
// oracle restates it bit for bit.  The threshold carries a 1e-6 relative margin
// over rl^2, so every pair with r < rl is kept whatever the rounding.
__device__ __forceinline__ float r2_list32(const SysDev& d, unsigned ax, unsigned ay, unsigned az, unsigned bx,
                                           unsigned by, unsigned bz) {
#pragma clang fp contract(off)
  const float dx = (float)(int)(bx - ax) * d.sc32[0];
  const float dy = (float)(int)(by - ay) * d.sc32[1];
  const float dz = (float)(int)(bz - az) * d.sc32[2];
  float r2 = dx * dx;
  r2 = r2 + dy * dy;
  r2 = r2 + dz * dz;
  return r2;
}

// One wave per I cluster.
//  1. Candidate columns (column-gap prefilter), one per lane; per column, the
//     z-window tables give the (at most two) ranges of clusters whose z bounds
//     can come within the pair-list distance of I's z bounds.
//  2. Those ranges (J >= I) are packed 64 to a block; each lane tests its J's
//     bounding box exactly (periodic gap in integers, FP64 squares).
//  3. Each hit is tested atom by atom: lane l (i = l & 7, j = l >> 3) keeps its
//     pair when both slots hold atoms, the pair is not excluded, i < j inside a
//     self pair, and the FP32 list distance is below the threshold.  An entry
//     whose 64-bit lane mask is empty is dropped.  The next hit's atoms are
//     loaded while the current one is tested.
//  With split > 1, split consecutive waves share one I cluster, each taking a
//  contiguous share of its candidate columns and writing its own buffer part;
//  the fill concatenates the parts in order, which is the order one wave
//  would have produced.
// (five waves per SIMD: the search waits on memory for every group of hits)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_pair_search(
    const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre, int split_cap) {
  SYS_FROM(pre);
  if (skip_build(d)) return;
  // a private copy of the descriptor: the entry stores below cannot alias it,
  // so its fields stay in registers instead of being reloaded per hit
  const SysDev dv = d;
  const int lane = threadIdx.x & 63;
  const int il = lane & 7, jl = lane >> 3;
  const int split = search_split(dv.search_parts, split_cap);
  const int gw = lb * 4 + (threadIdx.x >> 6);
  const int I = gw / split, part = gw - I * split;
  if (I >= *dv.n_clusters) return;
  const int pcap = dv.search_parts * SEARCH_TMP_CAP / split;  // entries per part
  const int nsc = dv.n_slots_cap;
  // global-address-space views: global_* loads and stores, which wait on the
  // vector memory counter only (generic ones also wait on the LDS counter)
  const FZ_GLOBAL unsigned* su32 = (const FZ_GLOBAL unsigned*)dv.su;  // high words at odd indices
  const FZ_GLOBAL int* slot_atom = (const FZ_GLOBAL int*)dv.slot_atom;
  const FZ_GLOBAL unsigned long long* bbg = (const FZ_GLOBAL unsigned long long*)dv.bb;
  // slot j of cluster J: atom at byte 4 (8 J + j) of slot_atom, high words of
  // x, y, z at bytes 8 (8 J + j) + 4 of the x, y, z blocks of su
  const __amdgpu_buffer_rsrc_t sa_r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<int*>(dv.slot_atom), 0, (int)(nsc * 4u), 0x00020000);
  const __amdgpu_buffer_rsrc_t su_r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<unsigned long long*>(dv.su), 0, (int)(3u * nsc * 8u), 0x00020000);
  const unsigned vo_a = 4u * jl, vo_x = 8u * jl + 4u, vo_y = vo_x + 8u * nsc, vo_z = vo_x + 16u * nsc;
  const int ai = slot_atom[I * CLUSTER + il];
  const unsigned aix = su32[2 * (I * CLUSTER + il) + 1], aiy = su32[2 * (nsc + I * CLUSTER + il) + 1],
                 aiz = su32[2 * (2 * nsc + I * CLUSTER + il) + 1];
  const unsigned long long wa = (ai >= 0 && !dv.use_excl_list) ? dv.excl_bits[ai] : 0ull;
  const size_t obase = (size_t)I * dv.search_parts * SEARCH_TMP_CAP + (size_t)part * pcap;
  FZ_GLOBAL unsigned long long* mout = (FZ_GLOBAL unsigned long long*)(dv.search_mtmp + obase);
  FZ_GLOBAL int* out = (FZ_GLOBAL int*)(dv.search_tmp + obase);
  unsigned long long bi[6];
#pragma unroll
  for (int e = 0; e < 6; e++) bi[e] = bbg[6 * (size_t)I + e];
  const int cI = dv.clu_col[I];
  const int cx = cI % dv.ncx, cy = cI / dv.ncx;
  const double ax0 = (double)bi[0] * dv.sc[0], ax1 = (double)bi[1] * dv.sc[0];
  const double ay0 = (double)bi[2] * dv.sc[1], ay1 = (double)bi[3] * dv.sc[1];
  const double wx = dv.L[0] / dv.ncx, wy = dv.L[1] / dv.ncy;
  const bool allx = 2 * dv.kx_range + 1 >= dv.ncx, ally = 2 * dv.ky_range + 1 >= dv.ncy;
  const int nyv = ally ? dv.ncy : 2 * dv.ky_range + 1;
  const int nxv = allx ? dv.ncx : 2 * dv.kx_range + 1;
  // z window [w0, w1] on the 2^64 circle
  const int nzb = dv.nzb;
  const unsigned long long R = dv.zwin;
  const unsigned long long binw = ~0ull / (unsigned long long)nzb;
  const bool zfull = R >= (1ull << 61) || bi[5] - bi[4] >= ~0ull - 2 * R - binw;
  const unsigned long long w0 = bi[4] - R, w1 = bi[5] + R;
  const bool zwrap = !zfull && w0 > w1;
  const int b0 = (int)__umul64hi(w0, (unsigned long long)nzb);
  const int b1 = (int)__umul64hi(w1, (unsigned long long)nzb);

  int count = 0;
  bool over = false;
  int cand = 0, fill = 0;
  // kept entries are buffered one per lane (lane k: the k-th of the current
  // 64) and stored 64 at a time with coalesced stores
  int bufJ = 0, bufMl = 0, bufMh = 0;

  auto test_hits = [&]() {
    bool hit = false;
    if (lane < fill) {
      unsigned long long bj[6];
#pragma unroll
      for (int e = 0; e < 6; e++) bj[e] = bbg[6 * (size_t)cand + e];
      hit = bbox_within(bi, bj, dv);
    }
    unsigned long long bal = __ballot(hit);
    // hits in groups of four: the atoms of all four are loaded before the
    // first is tested, so a wave waits on memory once per group
    while (bal) {
      int Jg[4];
#pragma unroll
      for (int h = 0; h < 4; h++) {
        Jg[h] = -1;
        if (bal) {
          Jg[h] = __builtin_amdgcn_readlane(cand, __ffsll((long long)bal) - 1);
          bal &= bal - 1ull;
        }
      }
      int bg[4];
      unsigned xg[4], yg[4], zg[4];
#pragma unroll
      for (int h = 0; h < 4; h++) {
        bg[h] = -1; xg[h] = yg[h] = zg[h] = 0u;
        if (Jg[h] >= 0) {
          // buffer loads: the cluster part of the offset is scalar, the lane
          // part a per-lane constant, so no vector address arithmetic per hit
          const unsigned so_a = (unsigned)Jg[h] * (CLUSTER * 4u), so_u = (unsigned)Jg[h] * (CLUSTER * 8u);
          bg[h] = (int)__builtin_amdgcn_raw_buffer_load_b32(sa_r, vo_a, so_a, 0);
          xg[h] = __builtin_amdgcn_raw_buffer_load_b32(su_r, vo_x, so_u, 0);
          yg[h] = __builtin_amdgcn_raw_buffer_load_b32(su_r, vo_y, so_u, 0);
          zg[h] = __builtin_amdgcn_raw_buffer_load_b32(su_r, vo_z, so_u, 0);
        }
      }
#pragma unroll
      for (int h = 0; h < 4; h++) {
        if (Jg[h] < 0) break;
        const int Jk = Jg[h], b = bg[h];
        bool on;
        if (!dv.use_excl_list) {
          // branch free: the window bit is read for every lane (the shift is
          // masked) and the distance computed for every lane
          const unsigned sh = (unsigned)(b - ai + 32);
          const bool ex = (sh < 64u) & (bool)((wa >> (sh & 63u)) & 1ull);
          const bool near = r2_list32(dv, aix, aiy, aiz, xg[h], yg[h], zg[h]) < dv.rl2f;
          on = (ai >= 0) & (b >= 0) & ((I != Jk) | (il < jl)) & !ex & near;
        } else {
          on = ai >= 0 && b >= 0 && (I != Jk || il < jl);
          if (on) on = !excluded_w(dv, wa, ai, b);
          if (on) on = r2_list32(dv, aix, aiy, aiz, xg[h], yg[h], zg[h]) < dv.rl2f;
        }
        const unsigned long long m = __ballot(on);
        if (m) {
          if (count < pcap) {
            const int k = count & 63;
            const bool mine = lane == k;
            bufJ = mine ? Jk : bufJ;
            bufMl = mine ? (int)(unsigned)m : bufMl;
            bufMh = mine ? (int)(unsigned)(m >> 32) : bufMh;
            if (k == 63) {
              out[count - 63 + lane] = bufJ;
              mout[count - 63 + lane] = ((unsigned long long)(unsigned)bufMh << 32) | (unsigned)bufMl;
            }
          } else {
            over = true;
          }
          count++;
        }
      }
    }
  };
  const int nq = nyv * nxv;
  const int q0 = part * nq / split, q1 = (part + 1) * nq / split;
  for (int g = q0; g < q1; g += 64) {
    // per lane: one candidate column and its (up to two) cluster ranges
    int r0a = 0, r0e = 0, r1a = 0, r1e = 0;
    const int q = g + lane;
    if (q < q1) {
      const int iy = q / nxv, ix = q - iy * nxv;
      const int yy = ally ? iy : (cy + iy - dv.ky_range + 2 * dv.ncy) % dv.ncy;
      const int xx = allx ? ix : (cx + ix - dv.kx_range + 2 * dv.ncx) % dv.ncx;
      const double gy = col_gap(ay0, ay1, yy, wy, dv.L[1]);
      const double gx = col_gap(ax0, ax1, xx, wx, dv.L[0]);
      if (gx * gx + gy * gy < dv.rl2) {
        const int C = yy * dv.ncx + xx;
        const FZ_GLOBAL int* clu0 = (const FZ_GLOBAL int*)dv.col_clu0;
        const int j0 = clu0[C], nc = clu0[C + 1] - j0;
        const FZ_GLOBAL int* th = (const FZ_GLOBAL int*)dv.zt_hi + (size_t)C * (nzb + 1);
        const FZ_GLOBAL int* tl = (const FZ_GLOBAL int*)dv.zt_lo + (size_t)C * (nzb + 1);
        if (zfull) {
          r0a = 0; r0e = nc;
        } else if (!zwrap) {
          r0a = th[b0]; r0e = tl[b1 + 1];
        } else {
          const int e1 = tl[b1 + 1], s0 = th[b0];
          if (s0 <= e1) { r0a = 0; r0e = nc; }
          else { r0a = 0; r0e = e1; r1a = s0; r1e = nc; }
        }
        r0a += j0; r0e += j0; r1a += j0; r1e += j0;
        r0a = max(r0a, I); r1a = max(r1a, I);
      }
    }
    // The group's candidates in order (column by column, each column's first
    // range then its second) are numbered by an inclusive scan of the per-lane
    // counts; each free slot of the 64-candidate block finds its column by a
    // binary search over the scan (lane permutes), so the block fills without
    // a scalar loop over columns.
    const int l0 = max(r0e - r0a, 0), l1 = max(r1e - r1a, 0);
    int incl = l0 + l1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    const int total = __builtin_amdgcn_readlane(incl, 63);
    for (int base = 0; base < total;) {
      const int m = min(total - base, 64 - fill);
      // every lane takes part in the permutes; lanes outside [fill, fill + m) discard
      const int p = base + lane - fill;
      int lo = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {  // first k with incl[k] > p
        const int v = __shfl(incl, lo + step - 1, 64);
        if (v <= p) lo += step;
      }
      const int k = min(lo, 63);
      const int off = p - (__shfl(incl, k, 64) - __shfl(l0 + l1, k, 64));
      const int c0 = __shfl(r0a, k, 64), n0 = __shfl(l0, k, 64), c1 = __shfl(r1a, k, 64);
      if (lane >= fill && lane < fill + m) cand = off < n0 ? c0 + off : c1 + (off - n0);
      fill += m;
      base += m;
      if (fill == 64) { test_hits(); fill = 0; }
    }
  }
  if (fill) test_hits();
  {  // the buffered tail (pcap is a multiple of 64, so a full buffer was stored)
    const int kept = count < pcap ? count : pcap;
    const int tail = kept & 63, base = kept - tail;
    if (lane < tail) {
      out[base + lane] = bufJ;
      mout[base + lane] = ((unsigned long long)(unsigned)bufMh << 32) | (unsigned)bufMl;
    }
  }
  if (lane == 0) {
    const int kept = count < pcap ? count : pcap;
    dv.search_cnt[I * split + part] = kept;
    if (kept) atomicAdd(dv.tile_cnt + (I >> 6), kept);
    if (over) atomicOr(dv.flags + F_OVERFLOW, OV_SEARCH);
  }
}

// One wave per I cluster: copy its entries into place.  The list offset of I
// is the sum of the entry counts of the 64-cluster tiles before I's tile plus
// the counts of the clusters before I inside it; the list total is the sum of
// every tile.  Each wave forms both from the tile counts, so no separate scan
// pass is needed.  A list over capacity is flagged and left unwritten.
__global__ __launch_bounds__(256) void k_pair_fill(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre,
                                                   int split_cap, int* __restrict__ chunk_pre) {
  SYS_FROM(pre);
  if (skip_build(d)) return;
  const int lane = threadIdx.x & 63;
  const int I = lb * 4 + (threadIdx.x >> 6);
  const int n = *d.n_clusters;
  if (I >= n) return;
  const int split = search_split(d.search_parts, split_cap);
  const int pcap = d.search_parts * SEARCH_TMP_CAP / split;
  const int t = I >> 6, nt = (n + 63) >> 6;
  long long before = 0, total = 0;
  for (int k = lane; k < nt; k += 64) {
    const long long v = d.tile_cnt[k];
    total += v;
    if (k < t) before += v;
  }
  for (int k = (t << 6) * split + lane; k < I * split; k += 64) before += d.search_cnt[k];
  before = wave_sum(before);
  total = wave_sum(total);
  // a single-system launch also lays out the pair kernel's work items (its
  // list's PAIR_CHUNK-entry chunks; batches do it in k_gather_sorted)
  if (total > d.n_pairs_cap) {
    if (I == 0 && lane == 0) {
      atomicOr(d.flags + F_OVERFLOW, OV_PAIRS);
      *d.n_pairs = 0;
      if (nsys == 1) { chunk_pre[0] = 0; chunk_pre[1] = 0; }
    }
    return;
  }
  if (I == 0 && lane == 0) {
    *d.n_pairs = total;
    d.flags[F_NBUILD] += 1;
    if (nsys == 1) { chunk_pre[0] = 0; chunk_pre[1] = (int)((total + PAIR_CHUNK - 1) / PAIR_CHUNK); }
  }
  long long off = before;
  for (int p = 0; p < split; p++) {
    const int cnt = d.search_cnt[I * split + p];
    const size_t base = (size_t)I * d.search_parts * SEARCH_TMP_CAP + (size_t)p * pcap;
    const int* src = d.search_tmp + base;
    const unsigned long long* msrc = d.search_mtmp + base;
    for (int k = lane; k < cnt; k += 64) {
      d.pl_i[off + k] = I;
      d.pl_j[off + k] = src[k];
      d.pl_mask[off + k] = msrc[k];
    }
    off += cnt;
  }
}

// ---------------------------------------------------------------------------
void launch_nblist(const Launch& L, bool check) {
  if (check) hipLaunchKernelGGL(k_check, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  if (L.w_atoms.total <= 512) {
    hipLaunchKernelGGL(k_bin_count<true>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  } else {
    hipLaunchKernelGGL(k_bin_count<false>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
    hipLaunchKernelGGL(k_col_scan, dim3(L.nsys), dim3(256), 0, L.stream, L.d_sys, L.nsys);
  }
  hipLaunchKernelGGL(k_bin_scatter, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  hipLaunchKernelGGL(k_col_sort, dim3(L.w_cols.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_cols.pre);
  hipLaunchKernelGGL(k_pair_search, dim3(L.w_search.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_search.pre,
                     L.search_cap);
  hipLaunchKernelGGL(k_pair_fill, dim3(L.w_clus.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_clus.pre,
                     L.search_cap, L.chunk_pre);
}

}  // namespace fz
