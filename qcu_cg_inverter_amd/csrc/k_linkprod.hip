// k_linkprod.hip -- measurement of the multi-right-hand-side link product on VALU and
// on MFMA (qcuBenchLinkProduct).
//
// In one leg of a 12-right-hand-side FP64 hop every site multiplies its link U (3x3
// complex, as a 6x6 real matrix) into the 24 projected colour vectors of its
// right-hand sides (2 per right-hand side; as a 6x24 real matrix H): Y = U H, 864
// FMA per site and leg.  Both kernels below compute exactly that for S site-legs, with
// operands made from the site-leg index in registers so that no memory traffic enters
// the time:
//   U(s)[r][c] = b(s) + 0.01 (6 r + c),  H(s)[c][j] = b(s) / 2 + 0.001 (j + 1) + 0.1 c,
//   b(s) = 0.001 (s mod 1024),
// and both return per-thread sums of every Y entry, so the host can check that the two
// computed the same products.
//  - VALU: one site-leg per lane, U in registers, the 24 columns in turn (36 FMA each).
//  - MFMA: one site-leg per wave on v_mfma_f64_16x16x4f64 (A 16x4, B 4x16, D 16x16),
//    U padded to 16 x 8 and H to 8 x 32: 2 K steps x 2 N tiles = 4 MFMAs, 21 % of
//    whose multiply slots hold a nonzero product.
#include "qcu_kernels.h"

namespace qcu {

typedef double v4d_t __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) linkprod_valu_kernel(long iters, double* out) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (long it = 0; it < iters; ++it) {
    const long s = gid * iters + it;
    const double b = 1e-3 * (double)(s & 1023);
    double u[36], hv[6];
#pragma unroll
    for (int k = 0; k < 36; ++k) u[k] = b + 0.01 * k;
#pragma unroll
    for (int c = 0; c < 6; ++c) hv[c] = 0.5 * b + 0.1 * c;
#pragma unroll 4
    for (int j = 0; j < 24; ++j) {
      const double dj = 0.001 * (j + 1);
      double h[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) h[c] = hv[c] + dj;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[r] = __builtin_fma(u[6 * r + c], h[c], acc[r]);
    }
  }
  out[gid] = acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5];
}

__global__ void __launch_bounds__(256) linkprod_mfma_kernel(long iters, double* out) {
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  // the lane's operand positions (v_mfma_f64_16x16x4f64: A[i][k] with i = lane % 16,
  // k = lane / 16; B[k][j] with k = lane / 16, j = lane % 16)
  const int i = lane & 15, kq = lane >> 4;
  v4d_t c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
  for (long it = 0; it < iters; ++it) {
    const long s = wave * iters + it;
    const double b = 1e-3 * (double)(s & 1023);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int k = kq + 4 * q;  // K index of this step
      const double a = (i < 6 && k < 6) ? b + 0.01 * (6 * i + k) : 0.0;
      const double b0 = (k < 6) ? 0.5 * b + 0.001 * (i + 1) + 0.1 * k : 0.0;         // columns 0..15
      const double b1 = (k < 6 && i < 8) ? 0.5 * b + 0.001 * (i + 17) + 0.1 * k : 0.0;  // columns 16..23
      c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, c1, 0, 0, 0);
    }
  }
  out[wave * 64 + lane] = c0[0] + c0[1] + c0[2] + c0[3] + c1[0] + c1[1] + c1[2] + c1[3];
}

void launch_linkprod(int mfma, long site_legs, double* out, hipStream_t st) {
  // 1024 blocks of 256 threads: S / 262144 site-legs per lane (VALU) or S / 4096 per
  // wave (MFMA); out holds 262144 doubles either way
  const long lanes = 1024L * 256;
  if (mfma) hipLaunchKernelGGL(linkprod_mfma_kernel, dim3(1024), dim3(256), 0, st, site_legs / (lanes / 64), out);
  else hipLaunchKernelGGL(linkprod_valu_kernel, dim3(1024), dim3(256), 0, st, site_legs / lanes, out);
}

}  // namespace qcu
