// k_clover.hip -- clover term A = 1 + ck T and its inverse, per chiral 6x6 block.
//
// T(x) = -(1/8) sum_{mu<nu} (g_mu g_nu) (x) (Q_munu - Q_munu^dag) with Q the sum of the
// four plaquette leaves at x (reference: src/qcu_clover_dslash.cu:329-458 leaves and
// sum, :210-267 the g_mu g_nu table with factor -1/16 * 2, :537-540 A = 1 + T).
// The reference reads the off-origin links from 32 shifted copies of the gauge
// (src/qcu_communicator.cu:1270-1297); here every link is read from the one packed
// gauge at its neighbour coordinate.  The reference inverts with unpivoted
// Gauss-Jordan (src/qcu_clover_dslash.cu:42-100); here Cholesky in FP64, with a
// partially pivoted Gauss-Jordan fallback for blocks that are not positive definite
// (they occur at the reference normalisation ck = 1 on rough gauges).
#include "qcu_fields.h"
#include "qcu_kernels.h"

namespace qcu {

typedef cx<double> zd;

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ void load_link_at(const double2* g, const Geom& geo, int x, int y, int z, int t, int mu,
                                             zd u[9]) {
  x = x < 0 ? x + geo.L[0] : (x >= geo.L[0] ? x - geo.L[0] : x);
  y = y < 0 ? y + geo.L[1] : (y >= geo.L[1] ? y - geo.L[1] : y);
  z = z < 0 ? z + geo.L[2] : (z >= geo.L[2] ? z - geo.L[2] : z);
  t = t < 0 ? t + geo.L[3] : (t >= geo.L[3] ? t - geo.L[3] : t);
  int par = (x + y + z + t) & 1;
  long idx = t * geo.st + z * geo.sz + y * geo.sy + (x >> 1);
  const double2* b = link_word<F64, 18>(g, geo.vh, mu * 2 + par, idx);
  const long ws = link_stride<F64, 18>(geo.vh);
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    double2 q = b[e * ws];
    u[e] = {q.x, q.y};
  }
}

__device__ __forceinline__ void m3mul(zd r[9], const zd a[9], const zd b[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}
__device__ __forceinline__ void m3dag(zd r[9], const zd a[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = conj(a[j * 3 + i]);
}

// thread per (site, plane): G = Q - Q^dag written to G[(pl * 9 + e) * vh + site]
__global__ void clover_leaves_kernel(const double2* gauge, double2* G, Geom geo, int parity) {
  long site = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int pl = blockIdx.y;
  if (site >= geo.vh) return;
  const int pmu[6] = {0, 0, 0, 1, 1, 2}, pnu[6] = {1, 2, 3, 2, 3, 3};
  int mu = pmu[pl], nu = pnu[pl];
  long r = site;
  int xh = (int)(r % geo.lxh); r /= geo.lxh;
  int y = (int)(r % geo.L[1]); r /= geo.L[1];
  int z = (int)(r % geo.L[2]); r /= geo.L[2];
  int t = (int)r;
  int x = 2 * xh + ((parity + y + z + t) & 1);
  int c[4] = {x, y, z, t};
  auto at = [&](int dm, int dn, int dir, zd u[9]) {
    int d[4] = {c[0], c[1], c[2], c[3]};
    d[mu] += dm;
    d[nu] += dn;
    load_link_at(gauge, geo, d[0], d[1], d[2], d[3], dir, u);
  };
  zd a[9], b[9], t1[9], t2[9], q[9];
  // F1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
  at(0, 0, mu, a); at(1, 0, nu, b); m3mul(t1, a, b);
  at(0, 1, mu, a); m3dag(b, a); m3mul(t2, t1, b);
  at(0, 0, nu, a); m3dag(b, a); m3mul(q, t2, b);
  // F2 = U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)
  at(0, 0, nu, a); at(-1, 1, mu, t2); m3dag(b, t2); m3mul(t1, a, b);
  at(-1, 0, nu, a); m3dag(b, a); m3mul(t2, t1, b);
  at(-1, 0, mu, a); m3mul(t1, t2, a);
#pragma unroll
  for (int k = 0; k < 9; ++k) q[k] = q[k] + t1[k];
  // F3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
  at(-1, 0, mu, b); m3dag(a, b); at(-1, -1, nu, t2); m3dag(b, t2); m3mul(t1, a, b);
  at(-1, -1, mu, a); m3mul(t2, t1, a);
  at(0, -1, nu, a); m3mul(t1, t2, a);
#pragma unroll
  for (int k = 0; k < 9; ++k) q[k] = q[k] + t1[k];
  // F4 = U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu) U_mu(x)^dag
  at(0, -1, nu, b); m3dag(a, b); at(0, -1, mu, b); m3mul(t1, a, b);
  at(1, -1, nu, a); m3mul(t2, t1, a);
  at(0, 0, mu, a); m3dag(b, a); m3mul(t1, t2, b);
#pragma unroll
  for (int k = 0; k < 9; ++k) q[k] = q[k] + t1[k];
  m3dag(b, q);
  double2* o = G + (long)pl * 9 * geo.vh + site;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    zd v = q[k] - b[k];
    o[k * geo.vh] = double2{v.re, v.im};
  }
}

// g_mu g_nu entries per plane: (row spin, col spin, phase i^k), the reference's table
__device__ constexpr int CL_COL[6][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {1, 0, 3, 2}, {1, 0, 3, 2}, {1, 0, 3, 2}, {0, 1, 2, 3}};
__device__ constexpr int CL_PH[6][4] = {{3, 1, 3, 1}, {2, 0, 2, 0}, {1, 1, 3, 3}, {3, 3, 3, 3}, {2, 0, 0, 2}, {1, 3, 3, 1}};

// Cholesky inverse of a Hermitian 6x6 (full storage m[36]); false if not
// positive definite
__device__ bool chol_inv6(const zd m[36], zd inv[36]) {
  zd l[36];
  double dinv[6];
#pragma unroll
  for (int k = 0; k < 36; ++k) l[k] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double d = m[j * 6 + j].re;
#pragma unroll
    for (int k = 0; k < j; ++k) d -= norm2(l[j * 6 + k]);
    if (!(d > 0.0)) return false;
    double lj = sqrt(d);
    l[j * 6 + j] = {lj, 0.0};
    dinv[j] = 1.0 / lj;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      zd s = m[i * 6 + j];
#pragma unroll
      for (int k = 0; k < j; ++k) s = s - mulc(l[i * 6 + k], l[j * 6 + k]);
      l[i * 6 + j] = dinv[j] * s;
    }
  }
  zd w[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) w[k] = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    w[j * 6 + j] = {dinv[j], 0.0};
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      zd s = {0.0, 0.0};
#pragma unroll
      for (int k = j; k < i; ++k) s = s + l[i * 6 + k] * w[k * 6 + j];
      w[i * 6 + j] = (-dinv[i]) * s;
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      zd s = {0.0, 0.0};
#pragma unroll
      for (int k = i; k < 6; ++k) s = s + conj(w[k * 6 + i]) * w[k * 6 + j];
      inv[i * 6 + j] = s;
      inv[j * 6 + i] = conj(s);
    }
  return true;
}

// Gauss-Jordan with partial pivoting (any nonsingular 6x6)
__device__ void gj_inv6(const zd m[36], zd inv[36]) {
  zd a[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) {
    a[k] = m[k];
    inv[k] = {(k % 7 == 0) ? 1.0 : 0.0, 0.0};
  }
  for (int c = 0; c < 6; ++c) {
    int piv = c;
    double best = norm2(a[c * 6 + c]);
    for (int r = c + 1; r < 6; ++r) {
      double v = norm2(a[r * 6 + c]);
      if (v > best) { best = v; piv = r; }
    }
    if (piv != c) {
      for (int k = 0; k < 6; ++k) {
        zd t = a[c * 6 + k]; a[c * 6 + k] = a[piv * 6 + k]; a[piv * 6 + k] = t;
        t = inv[c * 6 + k]; inv[c * 6 + k] = inv[piv * 6 + k]; inv[piv * 6 + k] = t;
      }
    }
    zd p = a[c * 6 + c];
    double d = 1.0 / norm2(p);
    zd pi = {p.re * d, -p.im * d};
    for (int k = 0; k < 6; ++k) {
      a[c * 6 + k] = a[c * 6 + k] * pi;
      inv[c * 6 + k] = inv[c * 6 + k] * pi;
    }
    for (int r = 0; r < 6; ++r) {
      if (r == c) continue;
      zd f = a[r * 6 + c];
      for (int k = 0; k < 6; ++k) {
        a[r * 6 + k] = a[r * 6 + k] - f * a[c * 6 + k];
        inv[r * 6 + k] = inv[r * 6 + k] - f * inv[c * 6 + k];
      }
    }
  }
}

__device__ __forceinline__ void pack_block(const zd m[36], double2* out, long vh, long site, int b) {
  double r[36];
  for (int i = 0; i < 6; ++i) r[i] = m[i * 7].re;
  int k = 6;
  for (int i = 1; i < 6; ++i)
    for (int j = 0; j < i; ++j) {
      r[k] = m[i * 6 + j].re;
      r[k + 1] = m[i * 6 + j].im;
      k += 2;
    }
  double2* o = site_block<36>(out, site) + b * 18 * kSiteBlock;
#pragma unroll
  for (int q = 0; q < 18; ++q) o[q * kSiteBlock] = double2{r[2 * q], r[2 * q + 1]};
}

// thread per (site, chiral block b)
__global__ void clover_assemble_kernel(const double2* G, double2* A, double2* Ainv, long vh, double ck, int* nbad) {
  long site = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (site >= vh) return;
  zd m[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) m[k] = {0.0, 0.0};
  for (int pl = 0; pl < 6; ++pl) {
    zd g[9];
    const double2* gp = G + (long)pl * 9 * vh + site;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      double2 q = gp[e * vh];
      g[e] = {q.x, q.y};
    }
#pragma unroll
    for (int e = 2 * b; e < 2 * b + 2; ++e) {  // entries whose row spin lies in block b
      int rs = e, cs = CL_COL[pl][e];
      zd ph = iph(zd{-0.125, 0.0}, CL_PH[pl][e]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          int row = (rs & 1) * 3 + i, col = (cs & 1) * 3 + j;
          m[row * 6 + col] = m[row * 6 + col] + ph * g[i * 3 + j];
        }
    }
  }
#pragma unroll
  for (int k = 0; k < 36; ++k) m[k] = ck * m[k];
#pragma unroll
  for (int i = 0; i < 6; ++i) m[i * 7].re += 1.0;
  if (A) pack_block(m, A, vh, site, b);
  if (Ainv) {  // the even-odd solve needs the inverse on one parity only
    zd inv[36];
    if (!chol_inv6(m, inv)) {
      gj_inv6(m, inv);
      atomicAdd(nbad, 1);
    }
    pack_block(inv, Ainv, vh, site, b);
  }
}

// packed F64 -> F32 (18 float4 words per site) / F16 (9 uint4 words); one parity
template <int P>
__global__ void clover_convert_kernel(const double2* src, void* dst, long vh) {
  long site = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= vh) return;
  float r[72];
  const double2* sp = site_block<36>(src, site);
#pragma unroll
  for (int k = 0; k < 36; ++k) {
    double2 q = sp[k * kSiteBlock];
    r[2 * k] = (float)q.x;
    r[2 * k + 1] = (float)q.y;
  }
  if (P == F32) {
    float4* o = site_block<18>(static_cast<float4*>(dst), site);
#pragma unroll
    for (int k = 0; k < 18; ++k)
      o[k * kSiteBlock] = float4{r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]};
  } else {
    uint4* o = site_block<9>(static_cast<uint4*>(dst), site);
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k * kSiteBlock] = pack8(r + 8 * k);
  }
}

// packed F64 -> complex128 [vh][2][6][6] (reference layout, include/qcu_point.cuh:91-93)
__global__ void clover_unpack_kernel(const double2* src, double2* out, long vh) {
  long site = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= vh) return;
  double a[72];
  const double2* sp = site_block<36>(src, site);
#pragma unroll
  for (int k = 0; k < 36; ++k) {
    double2 q = sp[k * kSiteBlock];
    a[2 * k] = q.x;
    a[2 * k + 1] = q.y;
  }
  double2* o = out + site * 72;
  for (int b = 0; b < 2; ++b) {
    const double* m = a + 36 * b;
    for (int i = 0; i < 6; ++i) o[b * 36 + i * 7] = double2{m[i], 0.0};
    int k = 6;
    for (int i = 1; i < 6; ++i)
      for (int j = 0; j < i; ++j) {
        o[b * 36 + i * 6 + j] = double2{m[k], m[k + 1]};
        o[b * 36 + j * 6 + i] = double2{m[k], -m[k + 1]};
        k += 2;
      }
  }
}

void launch_clover_leaves(const void* gauge, void* G, const Geom& g, int parity, hipStream_t st) {
  hipLaunchKernelGGL(clover_leaves_kernel, dim3(nblk(g.vh, 128), 6), dim3(128), 0, st,
                     static_cast<const double2*>(gauge), static_cast<double2*>(G), g, parity);
}
void launch_clover_assemble(const void* G, void* A, void* Ainv, long vh, double ck, int* nbad, hipStream_t st) {
  hipLaunchKernelGGL(clover_assemble_kernel, dim3(nblk(vh, 64), 2), dim3(64), 0, st, static_cast<const double2*>(G),
                     static_cast<double2*>(A), static_cast<double2*>(Ainv), vh, ck, nbad);
}
void launch_clover_convert(int prec, const void* src, void* dst, long vh, hipStream_t st) {
  if (prec == F32)
    hipLaunchKernelGGL(clover_convert_kernel<F32>, dim3(nblk(vh, 256)), dim3(256), 0, st,
                       static_cast<const double2*>(src), dst, vh);
  else
    hipLaunchKernelGGL(clover_convert_kernel<F16>, dim3(nblk(vh, 256)), dim3(256), 0, st,
                       static_cast<const double2*>(src), dst, vh);
}
void launch_clover_unpack(const void* packed, void* out, long vh, hipStream_t st) {
  hipLaunchKernelGGL(clover_unpack_kernel, dim3(nblk(vh, 256)), dim3(256), 0, st,
                     static_cast<const double2*>(packed), static_cast<double2*>(out), vh);
}

}  // namespace qcu
