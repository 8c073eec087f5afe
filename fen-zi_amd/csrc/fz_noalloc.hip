// fz_noalloc.hip -- the PME convolution entry point for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops):
// the same FFT and convolution kernels as fz_pme_convolve, with every buffer
// beyond the caller's grids (system descriptor, accumulators, work tables,
// per-axis tables, accumulation grid, half spectrum, potential grid) carved out
// of one file-scope device array.
//
// The array is reached from host code only through the address that
// hipGetSymbolAddress returns for it on each device (taken once per device by
// fz_ctx_create); its host-side name is not a device address.  Every call on a
// device shares that device's array and leaves its contents behind: each call
// rewrites everything it reads before reading it, and calls that overlap on one
// device (different streams or threads) race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "fz_device.h"
#include "fz_engine.h"
#include "fz_launch.h"

namespace fz {

namespace {

constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Largest grid: K0 * K1 * K2 <= 2^18 points (64^3), each axis <= MAX_FFT.
constexpr size_t NA_K3 = size_t(1) << 18;
constexpr int NA_KS = 3 * MAX_FFT;        // per-axis table entries
constexpr int NA_PRE = 2;                 // one-system work tables: {0, total}
constexpr size_t OFF_DESC = 0;
constexpr size_t OFF_ACC = align256(OFF_DESC + sizeof(SysDev));
constexpr size_t OFF_FLAGS = align256(OFF_ACC + sizeof(unsigned long long) * A_NSLOT);
constexpr size_t OFF_PRE = align256(OFF_FLAGS + sizeof(int) * F_NSLOT);
constexpr size_t OFF_BSP = align256(OFF_PRE + sizeof(int) * 3 * NA_PRE);
constexpr size_t OFF_EX = align256(OFF_BSP + sizeof(double) * NA_KS);
constexpr size_t OFF_M = align256(OFF_EX + sizeof(double) * NA_KS);
constexpr size_t OFF_TW = align256(OFF_M + sizeof(double) * NA_KS);
constexpr size_t OFF_QACC = align256(OFF_TW + 2 * sizeof(double) * NA_KS);
constexpr size_t OFF_QC = align256(OFF_QACC + sizeof(double) * NA_K3);  // int64 or double per point
// half spectrum: (K0/2 + 1) K1 K2 <= K3 complex values (K0 >= 2), FP64 complex at most
constexpr size_t OFF_PHI = align256(OFF_QC + 2 * sizeof(double) * NA_K3);
constexpr size_t NA_BYTES = align256(OFF_PHI + sizeof(double) * NA_K3);

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

std::mutex g_mu;
void* g_ptr[64] = {nullptr};  // per device: device address of that device's scratch

bool smooth235_na(int n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Descriptor, zeroed accumulators and flags, one-system work tables and the
// per-axis tables of the grid (the formulas of the system set-up):
// |b(m)|^2 = 9 / (2 + cos(2 pi j / K))^2, e^{-pi^2 m^2 / beta^2}, m = j' / L
// (j' = j - K above K/2) and twiddles e^{-2 pi i j / K}.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_na_setup(SysDev d, SysDev* dst, double b0, double b1, double b2, double beta) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *dst = d;
  if (t < A_NSLOT) d.acc[t] = 0ull;
  if (t < F_NSLOT) d.flags[t] = 0;
  const int KS = d.K[0] + d.K[1] + d.K[2];
  if (t < KS) {
    const int ax = t < d.K[0] ? 0 : (t < d.K[0] + d.K[1] ? 1 : 2);
    const int j = t - (ax == 0 ? 0 : (ax == 1 ? d.K[0] : d.K[0] + d.K[1]));
    const int K = d.K[ax];
    const double L = ax == 0 ? b0 : (ax == 1 ? b1 : b2);
    const int jj = j <= K / 2 ? j : j - K;
    const double m = jj / L, c = cos(2.0 * PI_D * j / K);
    d.pme_bsp[t] = 9.0 / ((2.0 + c) * (2.0 + c));
    d.pme_ex[t] = exp(-PI_D * PI_D * m * m / (beta * beta));
    d.pme_m[t] = m;
    const double a = -2.0 * PI_D * j / K;
    if (MIXED) ((float2*)d.tw)[t] = make_float2((float)cos(a), (float)sin(a));
    else ((double2*)d.tw)[t] = make_double2(cos(a), sin(a));
  }
}

// work tables {0, total} for the x, y and z passes (one-system launches read
// no per-block entries)
__global__ void k_na_pre(int* pre, int tx, int ty, int tz) {
  if (threadIdx.x == 0) {
    pre[0] = 0; pre[1] = tx;
    pre[2] = 0; pre[3] = ty;
    pre[4] = 0; pre[5] = tz;
  }
}

// caller's real grid -> accumulation grid (fixed point 2^32 in the mixed form)
template <bool MIXED>
__global__ __launch_bounds__(256) void k_na_load(const double* __restrict__ q, void* Qacc, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (MIXED) ((unsigned long long*)Qacc)[i] = (unsigned long long)__double2ll_rn(q[i] * ACC_SCALE);
  else ((double*)Qacc)[i] = q[i];
}

// potential grid -> caller's grid (FP64); reciprocal energy -> caller
template <bool MIXED>
__global__ __launch_bounds__(256) void k_na_store(const void* __restrict__ phi, double* __restrict__ out, int n,
                                                  const unsigned long long* __restrict__ acc, double* energy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = MIXED ? (double)((const float*)phi)[i] : ((const double*)phi)[i];
  if (i == 0 && energy) *energy = from_fixed_e(acc[A_PME]);
}

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= 64) return set_error(FZ_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return FZ_OK;
  void* p = nullptr;
  FZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return set_error(FZ_ERR_HIP, "no device address for the scratch array");
  g_ptr[device] = p;
  return FZ_OK;
}

size_t noalloc_bytes() { return NA_BYTES; }

int pme_convolve_noalloc(int device, const int32_t K[3], const double box[3], double beta, int mixed, int mode,
                         const double* q, double* phi, double* energy, hipStream_t stream) {
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device >= 0 && device < 64) base = g_ptr[device];
  }
  if (!base) return set_error(FZ_ERR_STATE, "device %d has no scratch address (create its context first)", device);
  for (int k = 0; k < 3; k++)
    if (!smooth235_na(K[k]) || K[k] < 2 || K[k] > MAX_FFT)
      return set_error(FZ_ERR_ARG, "grid must be 2,3,5-smooth, 2..%d", MAX_FFT);
  const size_t K3 = (size_t)K[0] * K[1] * K[2];
  if (K3 > NA_K3)
    return set_error(FZ_ERR_ARG, "grid of %zu points exceeds the %zu points of the no-allocation scratch", K3, NA_K3);
  if (!(box[0] > 0 && box[1] > 0 && box[2] > 0) || !(beta > 0))
    return set_error(FZ_ERR_ARG, "box and beta must be positive");
  if (mode != 0 && mode != 1) return set_error(FZ_ERR_ARG, "mode must be 0 or 1");
  FZ_HIP(hipSetDevice(device));
  char* b = (char*)base;
  SysDev d;
  memset(&d, 0, sizeof d);
  for (int k = 0; k < 3; k++) { d.K[k] = K[k]; d.L[k] = box[k]; }
  d.Kxh = K[0] / 2 + 1;
  d.beta = beta;
  d.pme_pre = CC / (PI_D * (box[0] * box[1] * box[2]));
  d.pi2_beta2 = PI_D * PI_D / (beta * beta);
  d.acc = (unsigned long long*)(b + OFF_ACC);
  d.flags = (int*)(b + OFF_FLAGS);
  d.pme_bsp = (double*)(b + OFF_BSP);
  d.pme_ex = (double*)(b + OFF_EX);
  d.pme_m = (double*)(b + OFF_M);
  d.tw = b + OFF_TW;
  d.Qacc = b + OFF_QACC;
  d.Qc = b + OFF_QC;
  d.phi = b + OFF_PHI;
  SysDev* dd = (SysDev*)(b + OFF_DESC);
  int* pre = (int*)(b + OFF_PRE);
  // launch shapes as the engine picks them for one system (16 lines per FFT
  // workgroup, fewer when the two line buffers and the twiddles exceed 48 KiB)
  const int kmax = std::max({4, K[0], K[1], K[2]});
  const int csz = mixed ? 8 : 16;
  int NL = 16;
  while (NL > 1 && (size_t)(NL * 2 + 1) * kmax * csz > 49152) NL >>= 1;
  const int ng = (d.Kxh + NL - 1) / NL;
  Launch L;
  L.stream = stream;
  L.d_sys = dd;
  L.nsys = 1;
  L.mixed = mixed;
  L.fft_nl = NL;
  L.fft_lds = (NL * 2 + 1) * kmax * csz;
  L.spread_bricks = 0;  // the x pass clears the accumulation grid after reading it
  L.w_fftx = Work{pre + 0, (K[1] * K[2] + NL - 1) / NL};
  L.w_ffty = Work{pre + 2, K[2] * ng};
  L.w_fftz = Work{pre + 4, K[1] * ng};
  const int KS = K[0] + K[1] + K[2];
  const int nset = std::max(KS, (int)A_NSLOT);
  const int nK3 = (int)K3;
  if (mixed)
    hipLaunchKernelGGL(k_na_setup<true>, dim3((nset + 255) / 256), dim3(256), 0, stream, d, dd, box[0], box[1],
                       box[2], beta);
  else
    hipLaunchKernelGGL(k_na_setup<false>, dim3((nset + 255) / 256), dim3(256), 0, stream, d, dd, box[0], box[1],
                       box[2], beta);
  hipLaunchKernelGGL(k_na_pre, dim3(1), dim3(64), 0, stream, pre, L.w_fftx.total, L.w_ffty.total, L.w_fftz.total);
  if (mixed) hipLaunchKernelGGL(k_na_load<true>, dim3((nK3 + 255) / 256), dim3(256), 0, stream, q, d.Qacc, nK3);
  else hipLaunchKernelGGL(k_na_load<false>, dim3((nK3 + 255) / 256), dim3(256), 0, stream, q, d.Qacc, nK3);
  launch_pme_fft_only(L, mode);
  if (mixed)
    hipLaunchKernelGGL(k_na_store<true>, dim3((nK3 + 255) / 256), dim3(256), 0, stream, d.phi, phi, nK3,
                       (const unsigned long long*)d.acc, energy);
  else
    hipLaunchKernelGGL(k_na_store<false>, dim3((nK3 + 255) / 256), dim3(256), 0, stream, d.phi, phi, nK3,
                       (const unsigned long long*)d.acc, energy);
  FZ_HIP(hipGetLastError());
  return FZ_OK;
}

}  // namespace fz
