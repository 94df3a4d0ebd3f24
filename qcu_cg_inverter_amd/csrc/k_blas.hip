// k_blas.hip -- level-1 kernels and the CGLS scalar steps.
//
// The reference does each BLAS-1 op as its own kernel plus a host round trip per
// reduction (src/qcu_complex_computation.cu:10-134, src/cg_inverter.cu:181-256).
// Here reductions are per-16-site partials (fixed butterfly order within 16 lanes)
// summed by a one-block kernel in a fixed order, so every result is bitwise reproducible and
// independent of how right-hand sides are batched, and all scalars stay on the
// device: a CG iteration needs no host synchronisation.
#include "qcu_fields.h"
#include "qcu_kernels.h"

namespace qcu {

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

// the same 16-lane butterfly as the hopping kernel's partials
__device__ __forceinline__ double group16_sum(double v) {
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 16);
  return v;
}

// write each 16-lane group's partial sum (its first lane) at slot site/16
__device__ __forceinline__ void put_partial(double* part, long vh, int r, long i, double v) {
  v = group16_sum(v);
  if ((threadIdx.x & 15) == 0 && i < vh) part[(long)r * n_partials(vh) + i / 16] = v;
}

template <int P>
__global__ void zero_kernel(Vec v, long vh) {
  using T = typename Real<P>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cx<T> z[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) z[c] = {T(0), T(0)};
  SpinorIO<P>::store(static_cast<char*>(v.p) + blockIdx.y * v.stride, vh, i, z);
}

template <int P>
__global__ void norm2_kernel(Vec v, long vh, double* part) {
  using T = typename Real<P>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  double s = 0.0;
  if (i < vh) {
    cx<T> a[12];
    SpinorIO<P>::load(static_cast<const char*>(v.p) + r * v.stride, vh, i, a);
    T t = T(0);
#pragma unroll
    for (int c = 0; c < 12; ++c) t += norm2(a[c]);
    s = (double)t;
  }
  put_partial(part, vh, r, i, s);
}

// One block of kSumThreads threads per right-hand side.  Every thread adds its strided
// partials in 16 independent chains (16 loads in flight; the tail as predicated loads
// rather than a dependent loop), the chains are combined in a fixed tree, each wave
// folds its 64 values with lane shuffles, and one thread adds the 16 wave sums.  The
// order depends on nw only, so the sum of a right-hand side never depends on the batch.
constexpr int kSumThreads = 1024;
constexpr int kSumChains = 16;
__device__ __forceinline__ double block_sum_partials(const double* part, long nw) {
  __shared__ double red[kSumThreads / 64];
  double c[kSumChains];
#pragma unroll
  for (int k = 0; k < kSumChains; ++k) c[k] = 0.0;
  long w = threadIdx.x;
  for (; w + (kSumChains - 1) * kSumThreads < nw; w += kSumChains * kSumThreads) {
#pragma unroll
    for (int k = 0; k < kSumChains; ++k) c[k] += part[w + k * kSumThreads];
  }
#pragma unroll
  for (int k = 0; k < kSumChains - 1; ++k) {
    const long j = w + (long)k * kSumThreads;
    if (j < nw) c[k] += part[j];
  }
#pragma unroll
  for (int h = kSumChains / 2; h >= 1; h >>= 1)
#pragma unroll
    for (int k = 0; k < h; ++k) c[k] += c[k + h];
  double v = c[0];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kSumThreads / 64; ++k) s += red[k];
  }
  return s;  // valid in thread 0
}

__global__ void __launch_bounds__(kSumThreads) sum_partials_kernel(const double* part, long nw, double* out) {
  double s = block_sum_partials(part + (long)blockIdx.x * nw, nw);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

template <int P>
__global__ void axpby_kernel(double a, Vec x, double b, Vec y, long vh, int nts) {
  using T = typename Real<P>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (i >= vh) return;
  cx<T> xv[12], yv[12];
  SpinorIO<P>::load(static_cast<const char*>(x.p) + r * x.stride, vh, i, xv);
  char* yp = static_cast<char*>(y.p) + r * y.stride;
  SpinorIO<P>::load(yp, vh, i, yv);
#pragma unroll
  for (int c = 0; c < 12; ++c) yv[c] = T(a) * xv[c] + T(b) * yv[c];
  if (nts) SpinorIO<P>::template store<1>(yp, vh, i, yv);
  else SpinorIO<P>::template store<0>(yp, vh, i, yv);
}

template <int P>
__global__ void clover_apply_kernel(const void* clov, Vec in, Vec out, long vh) {
  using T = typename Real<P>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (i >= vh) return;
  T cl[72];
  CloverIO<P>::load(clov, vh, i, cl);
  cx<T> v[12], o[12];
  SpinorIO<P>::load(static_cast<const char*>(in.p) + r * in.stride, vh, i, v);
  clover_mul(cl, v, o);
  SpinorIO<P>::store(static_cast<char*>(out.p) + r * out.stride, vh, i, o);
}

// ---------------------------------------------------------------- CGLS ------
__global__ void __launch_bounds__(kSumThreads) cg_beta_kernel(const double* part, long nw, CgScal* sc) {
  CgScal& s = sc[blockIdx.x];
  if (!s.active) return;
  double g = block_sum_partials(part + (long)blockIdx.x * nw, nw);
  if (threadIdx.x == 0) {
    if (s.first || s.converged || !(s.gamma > 0.0)) {
      s.beta = 0.0;
    } else {
      s.beta = g / s.gamma;
    }
    s.first = 0;
    s.gamma = g;
  }
}

__global__ void __launch_bounds__(kSumThreads) cg_alpha_kernel(const double* part, long nw, CgScal* sc) {
  CgScal& s = sc[blockIdx.x];
  if (!s.active) return;
  double tt = block_sum_partials(part + (long)blockIdx.x * nw, nw);
  if (threadIdx.x == 0) {
    s.tt = tt;
    if (s.converged) {
      s.alpha = 0.0;
    } else if (!(tt > 0.0) || !(s.gamma == s.gamma) || tt != tt || tt > 1.0e300) {
      s.alpha = 0.0;  // breakdown (p^dag Mhat^dag Mhat p <= 0 or not finite): freeze
      s.breakdown = 1;
    } else {
      s.alpha = s.gamma / tt;
    }
  }
}

__global__ void __launch_bounds__(kSumThreads) cg_ss_kernel(const double* part, long nw, CgScal* sc) {
  CgScal& s = sc[blockIdx.x];
  if (!s.active) return;
  double ss = block_sum_partials(part + (long)blockIdx.x * nw, nw);
  if (threadIdx.x == 0) {
    if (!s.converged) {
      s.ss = ss;
      s.iter += 1;
      if (ss <= s.target2) s.converged = 1;
    }
  }
}

// x += alpha_k p_k of the previous iteration (still pending: alpha is that iteration's,
// p not yet overwritten), then p = r + beta p.  Folding the iterate update in here
// saves one read of p per iteration against doing it in the s update; the host
// applies a pending update itself (cg_xflush) before it reads x.
template <int P, int PX>
__global__ void cg_xpay_kernel(Vec r, Vec p, Vec x, const CgScal* sc, long vh, int nts) {
  using T = typename Real<P>::T;
  using TX = typename Real<PX>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (i >= vh || !sc[k].active) return;
  const double alpha = sc[k].alpha;
  const T beta = T(sc[k].beta);
  char* pp = static_cast<char*>(p.p) + k * p.stride;
  char* xp = static_cast<char*>(x.p) + k * x.stride;
  const char* rp = static_cast<const char*>(r.p) + k * r.stride;
  if (P == F64 && PX == F64) {
    // FP64 in halves of 6 components (p, r and x at once would spill)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      cx<double> pv[6], rv[6];
      if (alpha != 0.0 || beta != T(0)) SpinorIO<F64>::load_half<1>(pp, vh, i, h, pv);
      if (alpha != 0.0) {
        cx<double> xv[6];
        SpinorIO<F64>::load_half<1>(xp, vh, i, h, xv);
#pragma unroll
        for (int c = 0; c < 6; ++c) xv[c] = {xv[c].re + alpha * pv[c].re, xv[c].im + alpha * pv[c].im};
        if (nts) SpinorIO<F64>::store_half<1>(xp, vh, i, h, xv);
        else SpinorIO<F64>::store_half<0>(xp, vh, i, h, xv);
      }
      SpinorIO<F64>::load_half<1>(rp, vh, i, h, rv);
      if (beta != T(0)) {
#pragma unroll
        for (int c = 0; c < 6; ++c) rv[c] = rv[c] + (double)beta * pv[c];
      }
      if (nts) SpinorIO<F64>::store_half<1>(pp, vh, i, h, rv);
      else SpinorIO<F64>::store_half<0>(pp, vh, i, h, rv);
    }
    return;
  }
  cx<T> pv[12];
  if (alpha != 0.0 || beta != T(0)) SpinorIO<P>::load(pp, vh, i, pv);
  if (alpha != 0.0) {
    const TX ax = TX(alpha);
    cx<TX> xv[12];
    SpinorIO<PX>::load(xp, vh, i, xv);
#pragma unroll
    for (int c = 0; c < 12; ++c) xv[c] = {xv[c].re + ax * TX(pv[c].re), xv[c].im + ax * TX(pv[c].im)};
    if (nts) SpinorIO<PX>::template store<1>(xp, vh, i, xv);
    else SpinorIO<PX>::template store<0>(xp, vh, i, xv);
  }
  cx<T> rv[12];
  SpinorIO<P>::load(rp, vh, i, rv);
  if (beta == T(0)) {
    // restart (first iteration, after a restart, or converged): p = r
#pragma unroll
    for (int c = 0; c < 12; ++c) pv[c] = rv[c];
  } else {
#pragma unroll
    for (int c = 0; c < 12; ++c) pv[c] = rv[c] + beta * pv[c];
  }
  if (nts) SpinorIO<P>::template store<1>(pp, vh, i, pv);
  else SpinorIO<P>::template store<0>(pp, vh, i, pv);
}

// s -= alpha t, with the 16-site partials of |s|^2
template <int P>
__global__ void cg_update_kernel(Vec t, Vec s, const CgScal* sc, double* part, long vh, int nts) {
  using T = typename Real<P>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (!sc[k].active) return;
  double acc = 0.0;
  if (i < vh) {
    const T a = T(sc[k].alpha);
    char* sp = static_cast<char*>(s.p) + k * s.stride;
    const char* tp = static_cast<const char*>(t.p) + k * t.stride;
    T n = T(0);
    if (P == F64) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        cx<double> tv[6], sv[6];
        SpinorIO<F64>::load_half<1>(tp, vh, i, h, tv);
        SpinorIO<F64>::load_half<1>(sp, vh, i, h, sv);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          sv[c] = sv[c] - (double)a * tv[c];
          n += T(norm2(sv[c]));
        }
        if (nts) SpinorIO<F64>::store_half<1>(sp, vh, i, h, sv);
        else SpinorIO<F64>::store_half<0>(sp, vh, i, h, sv);
      }
    } else {
      cx<T> tv[12], sv[12];
      SpinorIO<P>::load(tp, vh, i, tv);
      SpinorIO<P>::load(sp, vh, i, sv);
#pragma unroll
      for (int c = 0; c < 12; ++c) {
        sv[c] = sv[c] - a * tv[c];
        n += norm2(sv[c]);
      }
      if (nts) SpinorIO<P>::template store<1>(sp, vh, i, sv);
      else SpinorIO<P>::template store<0>(sp, vh, i, sv);
    }
    acc = (double)n;
  }
  put_partial(part, vh, k, i, acc);
}

// x += alpha p for one slot: a pending iterate update applied by the host
template <int P, int PX>
__global__ void cg_xflush_kernel(Vec x, Vec p, double alpha, long vh) {
  using T = typename Real<P>::T;
  using TX = typename Real<PX>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cx<T> pv[12];
  cx<TX> xv[12];
  SpinorIO<P>::load(p.p, vh, i, pv);
  SpinorIO<PX>::load(x.p, vh, i, xv);
  const TX ax = TX(alpha);
#pragma unroll
  for (int c = 0; c < 12; ++c) xv[c] = {xv[c].re + ax * TX(pv[c].re), xv[c].im + ax * TX(pv[c].im)};
  SpinorIO<PX>::store(x.p, vh, i, xv);
}

// plain streaming copy, the bandwidth and PMC-counter reference kernel: one 16-byte
// element per lane over a full grid (on MI355X this moves about 6.2 TB/s at 1 GiB,
// where a grid-stride loop capped at 32 blocks per CU moves 4.5 TB/s)
__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  long n = (long)(bytes / 16);
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n);
}

// ------------------------------------------------------------- launchers ----
#define QCU_PSW(PV, K, ...)                                                      \
  do {                                                                           \
    if ((PV) == F64) hipLaunchKernelGGL(K<F64>, __VA_ARGS__);                    \
    else if ((PV) == F32) hipLaunchKernelGGL(K<F32>, __VA_ARGS__);               \
    else hipLaunchKernelGGL(K<F16>, __VA_ARGS__);                                \
  } while (0)

void launch_zero(int prec, Vec v, long vh, int nrhs, hipStream_t st) {
  QCU_PSW(prec, zero_kernel, dim3(nblk(vh, 256), nrhs), dim3(256), 0, st, v, vh);
}
void launch_norm2(int prec, Vec v, long vh, int nrhs, double* part, hipStream_t st) {
  QCU_PSW(prec, norm2_kernel, dim3(nblk(vh, 256), nrhs), dim3(256), 0, st, v, vh, part);
}
void launch_sum_partials(const double* part, long nw, int nrhs, double* out, hipStream_t st) {
  hipLaunchKernelGGL(sum_partials_kernel, dim3(nrhs), dim3(kSumThreads), 0, st, part, nw, out);
}
void launch_axpby(int prec, double a, Vec x, double b, Vec y, long vh, int nrhs, hipStream_t st) {
  QCU_PSW(prec, axpby_kernel, dim3(nblk(vh, 256), nrhs), dim3(256), 0, st, a, x, b, y, vh, nt_store(prec, vh, nrhs));
}
void launch_clover_apply(int prec, const void* clov, Vec in, Vec out, long vh, int nrhs, hipStream_t st) {
  QCU_PSW(prec, clover_apply_kernel, dim3(nblk(vh, 256), nrhs), dim3(256), 0, st, clov, in, out, vh);
}
void launch_cg_beta(const double* part, long nw, CgScal* sc, int nrhs, hipStream_t st) {
  hipLaunchKernelGGL(cg_beta_kernel, dim3(nrhs), dim3(kSumThreads), 0, st, part, nw, sc);
}
void launch_cg_alpha(const double* part, long nw, CgScal* sc, int nrhs, hipStream_t st) {
  hipLaunchKernelGGL(cg_alpha_kernel, dim3(nrhs), dim3(kSumThreads), 0, st, part, nw, sc);
}
void launch_cg_ss(const double* part, long nw, CgScal* sc, int nrhs, hipStream_t st) {
  hipLaunchKernelGGL(cg_ss_kernel, dim3(nrhs), dim3(kSumThreads), 0, st, part, nw, sc);
}
void launch_cg_xpay(int prec, int xprec, Vec r, Vec p, Vec x, const CgScal* sc, long vh, int nrhs, hipStream_t st) {
  dim3 g(nblk(vh, 256), nrhs), b(256);
  const int n = nt_store(prec, vh, nrhs);  // p decides for x as well
  if (prec == F64) hipLaunchKernelGGL((cg_xpay_kernel<F64, F64>), g, b, 0, st, r, p, x, sc, vh, n);
  else if (prec == F32 && xprec == F64) hipLaunchKernelGGL((cg_xpay_kernel<F32, F64>), g, b, 0, st, r, p, x, sc, vh, n);
  else if (prec == F32) hipLaunchKernelGGL((cg_xpay_kernel<F32, F32>), g, b, 0, st, r, p, x, sc, vh, n);
  else if (xprec == F64) hipLaunchKernelGGL((cg_xpay_kernel<F16, F64>), g, b, 0, st, r, p, x, sc, vh, n);
  else hipLaunchKernelGGL((cg_xpay_kernel<F16, F32>), g, b, 0, st, r, p, x, sc, vh, n);
}
void launch_cg_xflush(int prec, int xprec, Vec x, Vec p, double alpha, long vh, hipStream_t st) {
  dim3 g(nblk(vh, 256)), b(256);
  if (prec == F64) hipLaunchKernelGGL((cg_xflush_kernel<F64, F64>), g, b, 0, st, x, p, alpha, vh);
  else if (prec == F32 && xprec == F64) hipLaunchKernelGGL((cg_xflush_kernel<F32, F64>), g, b, 0, st, x, p, alpha, vh);
  else if (prec == F32) hipLaunchKernelGGL((cg_xflush_kernel<F32, F32>), g, b, 0, st, x, p, alpha, vh);
  else if (xprec == F64) hipLaunchKernelGGL((cg_xflush_kernel<F16, F64>), g, b, 0, st, x, p, alpha, vh);
  else hipLaunchKernelGGL((cg_xflush_kernel<F16, F32>), g, b, 0, st, x, p, alpha, vh);
}
void launch_cg_update(int prec, Vec t, Vec s, const CgScal* sc, double* part, long vh, int nrhs, hipStream_t st) {
  QCU_PSW(prec, cg_update_kernel, dim3(nblk(vh, 256), nrhs), dim3(256), 0, st, t, s, sc, part, vh,
          nt_store(prec, vh, nrhs));
}

}  // namespace qcu
