// k_gen.hip -- seeded synthetic inputs on the device (the reference has none; its
// caller supplies the fields).  Counter-based: value k of stream `seed` depends only
// on (seed, k), so any subset can be regenerated anywhere.
//   u64(seed, k) = mix(mix(seed) ^ k), mix = splitmix64 finaliser
//   uniform     = ((u64 >> 11) + 0.5) / 2^53
//   gauss(k)    = sqrt(-2 ln u(2k)) cos(2 pi u(2k+1))
//   link l      : H = sum_a h_a lambda_a / 2 with h_a = gauss(8 l + a),
//                 U = SU3-project(exp(i eps H)), exp by 6 squarings of a 12-term Taylor
//                 series of X / 64, projection = Gram-Schmidt of rows 0, 1 and
//                 row 2 = conj(row0 x row1).
#include "qcu_fields.h"
#include "qcu_kernels.h"

namespace qcu {

typedef cx<double> zd;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uni(uint64_t seed, uint64_t k) {
  return ((double)(mix64(mix64(seed) ^ k) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double gauss(uint64_t seed, uint64_t k) {
  double u1 = uni(seed, 2 * k), u2 = uni(seed, 2 * k + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ void m3mul_g(zd r[9], const zd a[9], const zd b[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      zd t = {0.0, 0.0};
      for (int k = 0; k < 3; ++k) t = t + a[i * 3 + k] * b[k * 3 + j];
      r[i * 3 + j] = t;
    }
}

__global__ void gauge_random_kernel(double2* g, long nlinks, uint64_t seed, double eps) {
  long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  double h[8];
  for (int a = 0; a < 8; ++a) h[a] = gauss(seed, (uint64_t)l * 8 + a);
  const double r3 = 0.57735026918962576451;
  zd H[9];
  H[0] = {0.5 * (h[2] + h[7] * r3), 0.0};
  H[4] = {0.5 * (-h[2] + h[7] * r3), 0.0};
  H[8] = {-h[7] * r3, 0.0};
  H[1] = {0.5 * h[0], -0.5 * h[1]};
  H[3] = conj(H[1]);
  H[2] = {0.5 * h[3], -0.5 * h[4]};
  H[6] = conj(H[2]);
  H[5] = {0.5 * h[5], -0.5 * h[6]};
  H[7] = conj(H[5]);
  zd xs[9], e[9], term[9], tmp[9];
  for (int k = 0; k < 9; ++k) {
    xs[k] = {-eps * H[k].im * (1.0 / 64.0), eps * H[k].re * (1.0 / 64.0)};
    e[k] = {k % 4 == 0 ? 1.0 : 0.0, 0.0};
    term[k] = e[k];
  }
  for (int n = 1; n <= 12; ++n) {
    m3mul_g(tmp, term, xs);
    for (int k = 0; k < 9; ++k) {
      term[k] = (1.0 / n) * tmp[k];
      e[k] = e[k] + term[k];
    }
  }
  for (int s = 0; s < 6; ++s) {
    m3mul_g(tmp, e, e);
    for (int k = 0; k < 9; ++k) e[k] = tmp[k];
  }
  double n0 = 0.0;
  for (int j = 0; j < 3; ++j) n0 += norm2(e[j]);
  n0 = 1.0 / sqrt(n0);
  for (int j = 0; j < 3; ++j) e[j] = n0 * e[j];
  zd d = {0.0, 0.0};
  for (int j = 0; j < 3; ++j) d = d + conj(e[j]) * e[3 + j];
  for (int j = 0; j < 3; ++j) e[3 + j] = e[3 + j] - d * e[j];
  double n1 = 0.0;
  for (int j = 0; j < 3; ++j) n1 += norm2(e[3 + j]);
  n1 = 1.0 / sqrt(n1);
  for (int j = 0; j < 3; ++j) e[3 + j] = n1 * e[3 + j];
  recon_row2(e);
  double2* o = g + l * 9;
  for (int k = 0; k < 9; ++k) o[k] = double2{e[k].re, e[k].im};
}

__global__ void source_gauss_kernel(double* v, long n_real, uint64_t seed) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_real) v[k] = gauss(seed, (uint64_t)k);
}

void launch_gauge_random(void* aos, long vh, uint64_t seed, double eps, hipStream_t st) {
  long n = 8 * vh;
  hipLaunchKernelGGL(gauge_random_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st,
                     static_cast<double2*>(aos), n, seed, eps);
}
void launch_source_gauss(void* aos, long n_complex, uint64_t seed, hipStream_t st) {
  long n = 2 * n_complex;
  hipLaunchKernelGGL(source_gauss_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     static_cast<double*>(aos), n, seed);
}

}  // namespace qcu
