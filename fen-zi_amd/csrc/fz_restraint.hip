// fz_restraint.hip -- harmonic centre-of-mass restraint between two atom groups.
//
// Replaces reduce_COM (src/cell_based_kernels.cu:713-866) and restraint_force
// (src/mdcuda.cu:481-516), enabled by the consharm / consharmdist / consharmfc
// keywords (src/prmread.cu:674-690):
//   c_g = sum_{i in g} m_i x_i / M_g over positions taken in [-L/2, L/2),
//   dr  = d + c_0 - c_1, minimum image,
//   F_i += fc * dr for every atom i of group 1 (group 0 receives no force;
//   the reference defines no energy for the term).
// The group sums go through fixed-point accumulators (per-block FP64 sums,
// added as int64), so the result does not depend on block or batch layout.
#include "fz_device.h"
#include "fz_launch.h"

namespace fz {

__global__ __launch_bounds__(256) void k_com_sum(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  if (!d.com_on) return;  // uniform over the block
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < d.n_atoms) {
    const int g = d.com_grp[i];
    if (g >= 0) {
      // (int64) u is the position in [-L/2, L/2) in units of L / 2^64
      const double m = 1.0 / d.minv[i];
      const double x = (double)(long long)d.ux[i] * d.sc[0];
      const double y = (double)(long long)d.uy[i] * d.sc[1];
      const double z = (double)(long long)d.uz[i] * d.sc[2];
      v[3 * g] = m * x;
      v[3 * g + 1] = m * y;
      v[3 * g + 2] = m * z;
    }
  }
  const int slots[6] = {A_COM, A_COM + 1, A_COM + 2, A_COM + 3, A_COM + 4, A_COM + 5};
  block_acc_fixed<6>(v, d.acc, slots);
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_com_force(const SysDev* __restrict__ S, int nsys, const int* __restrict__ pre) {
  const int s_ = block_sys(pre, nsys, blockIdx.x);
  const SysDev& d = S[s_];
  if (!d.com_on) return;
  const int i = (blockIdx.x - block_base(pre, nsys, s_)) * 256 + threadIdx.x;
  if (i >= d.n_atoms || d.com_grp[i] != 1) return;
  double dr[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double c0 = from_fixed_e(d.acc[A_COM + k]) * d.com_im[0];
    const double c1 = from_fixed_e(d.acc[A_COM + 3 + k]) * d.com_im[1];
    double x = d.com_d[k] + (c0 - c1);
    x -= d.L[k] * rint(x / d.L[k]);
    dr[k] = d.com_fc * x;
  }
  const int sl = d.atom_slot[i];
  const size_t cap = (size_t)d.n_slots_cap;
  if (MIXED) {
    unsigned long long* F = (unsigned long long*)d.Fs;
#pragma unroll
    for (int k = 0; k < 3; k++) atomicAdd(F + (size_t)sl * 3 + k, to_fixed_acc(dr[k]));
  } else {
    double* F = (double*)d.Fs;
#pragma unroll
    for (int k = 0; k < 3; k++) unsafeAtomicAdd(F + (size_t)sl * 3 + k, dr[k]);
  }
}

void launch_restraint(const Launch& L) {
  hipLaunchKernelGGL(k_com_sum, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  if (L.mixed)
    hipLaunchKernelGGL(k_com_force<true>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
  else
    hipLaunchKernelGGL(k_com_force<false>, dim3(L.w_atoms.total), dim3(256), 0, L.stream, L.d_sys, L.nsys, L.w_atoms.pre);
}

}  // namespace fz
