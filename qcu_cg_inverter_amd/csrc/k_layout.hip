// k_layout.hip -- conversions between the interface layouts (complex128 AoS, as the
// reference's entry points take them: include/qcu_point.cuh:65-72) and the internal
// SoA layouts of each precision, plus precision conversions between internal fields.
// Role of src/qcu_shift_storage.cu:74-194 (AoS <-> SoA, gauge packing); written for
// wave64 with one thread per site.
#include "qcu_fields.h"
#include "qcu_kernels.h"

namespace qcu {

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

template <int P>
__global__ void spinor_from_aos_kernel(const double2* aos, long aos_stride, Vec out, long vh) {
  using T = typename Real<P>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (i >= vh) return;
  const double2* s = reinterpret_cast<const double2*>(reinterpret_cast<const char*>(aos) + r * aos_stride) + i * 12;
  cx<T> v[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    double2 q = s[c];
    v[c] = {T(q.x), T(q.y)};
  }
  SpinorIO<P>::store(static_cast<char*>(out.p) + r * out.stride, vh, i, v);
}

template <int P>
__global__ void spinor_to_aos_kernel(Vec in, double2* aos, long aos_stride, long vh) {
  using T = typename Real<P>::T;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (i >= vh) return;
  cx<T> v[12];
  SpinorIO<P>::load_cached(static_cast<const char*>(in.p) + r * in.stride, vh, i, v);
  double2* d = reinterpret_cast<double2*>(reinterpret_cast<char*>(aos) + r * aos_stride) + i * 12;
#pragma unroll
  for (int c = 0; c < 12; ++c) d[c] = double2{(double)v[c].re, (double)v[c].im};
}

void launch_spinor_from_aos(int prec, const void* aos, long aos_stride, Vec out, long vh, int nrhs, hipStream_t st) {
  dim3 grid(nblk(vh, 256), nrhs);
  const double2* a = static_cast<const double2*>(aos);
  if (prec == F64) hipLaunchKernelGGL(spinor_from_aos_kernel<F64>, grid, dim3(256), 0, st, a, aos_stride, out, vh);
  else if (prec == F32) hipLaunchKernelGGL(spinor_from_aos_kernel<F32>, grid, dim3(256), 0, st, a, aos_stride, out, vh);
  else hipLaunchKernelGGL(spinor_from_aos_kernel<F16>, grid, dim3(256), 0, st, a, aos_stride, out, vh);
}

void launch_spinor_to_aos(int prec, Vec in, void* aos, long aos_stride, long vh, int nrhs, hipStream_t st) {
  dim3 grid(nblk(vh, 256), nrhs);
  double2* a = static_cast<double2*>(aos);
  if (prec == F64) hipLaunchKernelGGL(spinor_to_aos_kernel<F64>, grid, dim3(256), 0, st, in, a, aos_stride, vh);
  else if (prec == F32) hipLaunchKernelGGL(spinor_to_aos_kernel<F32>, grid, dim3(256), 0, st, in, a, aos_stride, vh);
  else hipLaunchKernelGGL(spinor_to_aos_kernel<F16>, grid, dim3(256), 0, st, in, a, aos_stride, vh);
}

// gauge: thread per (link, mu, parity); the interface link is 9 complex row-major; the
// packed field is written in the layout LinkIO reads (LinkLayout)
template <int P, int R>
__global__ void gauge_pack_kernel(const double2* aos, void* out, long vh) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int mp = blockIdx.y;  // mu * 2 + parity
  if (i >= vh) return;
  const double2* s = aos + ((long)mp * vh + i) * 9;
  constexpr int NE = R == 12 ? 6 : 9;
  static_assert(LinkLayout<F32, 12>::blocked && LinkLayout<F32, 18>::blocked && LinkLayout<F16, 12>::blocked &&
                    LinkLayout<F16, 18>::blocked,
                "the FP32 and FP16 stores below step by kSiteBlock");
  if (P == F64) {
    double2* d = link_word<F64, R>(static_cast<double2*>(out), vh, mp, i);
    const long ws = link_stride<F64, R>(vh);
#pragma unroll
    for (int e = 0; e < NE; ++e) d[e * ws] = s[e];
  } else if (P == F32) {
    if (R == 12) {
      float4* d = link_word<F32, 12>(static_cast<float4*>(out), vh, mp, i);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double2 a = s[2 * k], b = s[2 * k + 1];
        d[k * kSiteBlock] = float4{(float)a.x, (float)a.y, (float)b.x, (float)b.y};
      }
    } else {
      float2* d = link_word<F32, 18>(static_cast<float2*>(out), vh, mp, i);
#pragma unroll
      for (int e = 0; e < 9; ++e) d[e * kSiteBlock] = float2{(float)s[e].x, (float)s[e].y};
    }
  } else {
    if (R == 12) {
      uint2* d = link_word<F16, 12>(static_cast<uint2*>(out), vh, mp, i);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double2 a = s[2 * k], b = s[2 * k + 1];
        unsigned lo = (unsigned)f2h((float)a.x) | ((unsigned)f2h((float)a.y) << 16);
        unsigned hi = (unsigned)f2h((float)b.x) | ((unsigned)f2h((float)b.y) << 16);
        d[k * kSiteBlock] = uint2{lo, hi};
      }
    } else {
      unsigned* d = link_word<F16, 18>(static_cast<unsigned*>(out), vh, mp, i);
#pragma unroll
      for (int e = 0; e < 9; ++e)
        d[e * kSiteBlock] = (unsigned)f2h((float)s[e].x) | ((unsigned)f2h((float)s[e].y) << 16);
    }
  }
}

void launch_gauge_pack(int prec, int recon, const void* aos, void* out, long vh, hipStream_t st) {
  dim3 grid(nblk(vh, 256), 8);
  const double2* a = static_cast<const double2*>(aos);
#define QCU_GP(P, R) \
  if (prec == P && recon == R) { hipLaunchKernelGGL((gauge_pack_kernel<P, R>), grid, dim3(256), 0, st, a, out, vh); return; }
  QCU_GP(F64, 18) QCU_GP(F64, 12) QCU_GP(F32, 18) QCU_GP(F32, 12) QCU_GP(F16, 18) QCU_GP(F16, 12)
#undef QCU_GP
}

// row 2 of every interface-layout link rebuilt from rows 0 and 1, the reference's
// reading of a gauge (src/qcu_wilson_dslash_neo.cu:73-86)
__global__ void gauge_reconstruct_kernel(double2* aos, long nlinks) {
  long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  double2* s = aos + l * 9;
  cx<double> u[9];
#pragma unroll
  for (int e = 0; e < 6; ++e) u[e] = {s[e].x, s[e].y};
  recon_row2(u);
#pragma unroll
  for (int e = 6; e < 9; ++e) s[e] = double2{u[e].re, u[e].im};
}

void launch_gauge_reconstruct(void* aos, long vh, hipStream_t st) {
  long n = 8 * vh;
  hipLaunchKernelGGL(gauge_reconstruct_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, static_cast<double2*>(aos), n);
}

// ----------------------------------------------- reference storage shifts ---
// AoS <-> the reference's coalesced SoA (src/qcu_shift_storage.cu:74-161): an AoS
// record of `wa` doubles per site becomes `w` planes of n doubles (plane e holds
// element e of every site).  A block moves 64 sites through an LDS tile, so both the
// record reads and the plane writes are contiguous runs; blockIdx.y picks the field
// (the 8 (mu, parity) link fields of a gauge, or 1 for a half spinor).
constexpr int kTile = 64, kTilePad = kTile + 1, kTileMaxW = 24;

__global__ void __launch_bounds__(256) aos_to_soa_kernel(const double* src, double* dst, long n, int wa, int w,
                                                         long src_field, long dst_field) {
  __shared__ double tile[kTileMaxW * kTilePad];
  long s0 = (long)blockIdx.x * kTile;
  int ns = (int)min((long)kTile, n - s0);
  src += blockIdx.y * src_field + s0 * wa;
  dst += blockIdx.y * dst_field + s0;
  for (int j = threadIdx.x; j < ns * wa; j += blockDim.x) {
    int e = j % wa, s = j / wa;
    if (e < w) tile[e * kTilePad + s] = src[j];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kTile * w; j += blockDim.x) {
    int e = j / kTile, s = j % kTile;
    if (s < ns) dst[e * n + s] = tile[e * kTilePad + s];
  }
}

// the way back; RECON rebuilds link row 2 = conj(row0 x row1) from the 12 stored reals
template <bool RECON>
__global__ void __launch_bounds__(256) soa_to_aos_kernel(const double* src, double* dst, long n, int w, int wa,
                                                         long src_field, long dst_field) {
  __shared__ double tile[kTileMaxW * kTilePad];
  long s0 = (long)blockIdx.x * kTile;
  int ns = (int)min((long)kTile, n - s0);
  src += blockIdx.y * src_field + s0;
  dst += blockIdx.y * dst_field + s0 * wa;
  for (int j = threadIdx.x; j < kTile * w; j += blockDim.x) {
    int e = j / kTile, s = j % kTile;
    if (s < ns) tile[e * kTilePad + s] = src[e * n + s];
  }
  __syncthreads();
  if (RECON) {
    if ((int)threadIdx.x < ns) {
      int s = threadIdx.x;
      cx<double> u[9];
#pragma unroll
      for (int e = 0; e < 6; ++e) u[e] = {tile[(2 * e) * kTilePad + s], tile[(2 * e + 1) * kTilePad + s]};
      recon_row2(u);
#pragma unroll
      for (int e = 6; e < 9; ++e) {
        tile[(2 * e) * kTilePad + s] = u[e].re;
        tile[(2 * e + 1) * kTilePad + s] = u[e].im;
      }
    }
    __syncthreads();
  }
  for (int j = threadIdx.x; j < ns * wa; j += blockDim.x) dst[j] = tile[(j % wa) * kTilePad + j / wa];
}

void launch_shift_storage(int kind, int to_soa, const void* src, void* dst, long vh, hipStream_t st) {
  // kind 0: half spinor [Vh][24] <-> [24][Vh]; kind 1: gauge [8][Vh][18] <-> [8][12][Vh]
  int wa = kind == 0 ? 24 : 18, w = kind == 0 ? 24 : 12, nf = kind == 0 ? 1 : 8;
  dim3 grid(nblk(vh, kTile), nf);
  const double* s = static_cast<const double*>(src);
  double* d = static_cast<double*>(dst);
  if (to_soa)
    hipLaunchKernelGGL(aos_to_soa_kernel, grid, dim3(256), 0, st, s, d, vh, wa, w, vh * wa, vh * w);
  else if (kind == 0)
    hipLaunchKernelGGL(soa_to_aos_kernel<false>, grid, dim3(256), 0, st, s, d, vh, w, wa, vh * w, vh * wa);
  else
    hipLaunchKernelGGL(soa_to_aos_kernel<true>, grid, dim3(256), 0, st, s, d, vh, w, wa, vh * w, vh * wa);
}

// ------------------------------------------- lexicographic <-> even-odd ---
// lexicographic site l = ((t*Lz + z)*Ly + y)*Lx + x  <->  even-odd (parity, l/2) with
// parity = (x+y+z+t) mod 2, the half-site order of include/qcu_point.cuh:20-93.  One
// thread per 16-byte complex of a site record (NC complex per site), so consecutive
// lanes move consecutive bytes of the lexicographic field; blockIdx.y picks the field.
// The launcher guarantees vol < 2^31, the range of FastDiv.
template <bool TO_EO, int NC>
__global__ void lex_eo_kernel(const double2* src, double2* dst, long vol, FastDiv dlx, FastDiv dly, FastDiv dlz,
                              long field) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= vol * NC) return;
  uint32_t l = (uint32_t)(j / NC);
  int k = (int)(j - (long)l * NC);
  uint32_t r = dlx.div(l), x = l - r * dlx.d;
  uint32_t r2 = dly.div(r), y = r - r2 * dly.d;
  uint32_t t = dlz.div(r2), z = r2 - t * dlz.d;
  uint32_t parity = (x + y + z + t) & 1u;
  long e = (long)parity * (vol / 2) + (l >> 1);
  src += blockIdx.y * field;
  dst += blockIdx.y * field;
  if (TO_EO) dst[e * NC + k] = src[(long)l * NC + k];
  else dst[(long)l * NC + k] = src[e * NC + k];
}

void launch_lex_eo(int to_eo, const void* src, void* dst, const int L[4], int nc, int nfields, hipStream_t st) {
  long vol = (long)L[0] * L[1] * L[2] * L[3];
  long n = vol * nc;
  dim3 grid(nblk(n, 256), nfields);
  const double2* s = static_cast<const double2*>(src);
  double2* d = static_cast<double2*>(dst);
  FastDiv dlx((uint32_t)L[0]), dly((uint32_t)L[1]), dlz((uint32_t)L[2]);
#define QCU_LEX(E, N)                                                                                   \
  if (to_eo == E && nc == N) {                                                                          \
    hipLaunchKernelGGL((lex_eo_kernel<E, N>), grid, dim3(256), 0, st, s, d, vol, dlx, dly, dlz, n);     \
    return;                                                                                             \
  }
  QCU_LEX(1, 12) QCU_LEX(0, 12) QCU_LEX(1, 9) QCU_LEX(0, 9)
#undef QCU_LEX
}

template <int PI, int PO>
__global__ void spinor_convert_kernel(Vec in, Vec out, long vh, int accumulate) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (i >= vh) return;
  using TI = typename Real<PI>::T;
  using TO = typename Real<PO>::T;
  cx<TI> a[12];
  SpinorIO<PI>::load(static_cast<const char*>(in.p) + r * in.stride, vh, i, a);
  cx<TO> b[12];
  char* o = static_cast<char*>(out.p) + r * out.stride;
  if (accumulate) {
    SpinorIO<PO>::load(o, vh, i, b);
#pragma unroll
    for (int c = 0; c < 12; ++c) b[c] = {b[c].re + TO(a[c].re), b[c].im + TO(a[c].im)};
  } else {
#pragma unroll
    for (int c = 0; c < 12; ++c) b[c] = {TO(a[c].re), TO(a[c].im)};
  }
  SpinorIO<PO>::store(o, vh, i, b);
}

void launch_spinor_convert(int pin, Vec in, int pout, Vec out, long vh, int nrhs, int accumulate, hipStream_t st) {
  dim3 grid(nblk(vh, 256), nrhs);
#define QCU_SC(A, B)                                                                                       \
  if (pin == A && pout == B) {                                                                             \
    hipLaunchKernelGGL((spinor_convert_kernel<A, B>), grid, dim3(256), 0, st, in, out, vh, accumulate); \
    return;                                                                                                \
  }
  QCU_SC(F64, F64) QCU_SC(F64, F32) QCU_SC(F64, F16)
  QCU_SC(F32, F64) QCU_SC(F32, F32) QCU_SC(F32, F16)
  QCU_SC(F16, F64) QCU_SC(F16, F32) QCU_SC(F16, F16)
#undef QCU_SC
}


// ------------------------------------------------------------ fingerprint ---
// A 64-bit fingerprint of a buffer: the sum (mod 2^64) over its 64-bit words w_i of
// mix(w_i ^ (i * 0x9E3779B97F4A7C15)), mix being the splitmix64 finaliser (Steele,
// Lea and Flood, OOPSLA 2014).  A sum of integers does not depend on the order it is
// taken in, so the result is the same for any launch shape, and a change of any one
// word changes it except with probability about 2^-64.  Used by the drop-in entry
// points to notice a gauge the caller changed in place.
__device__ __forceinline__ uint64_t fp_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
    lo = __shfl_xor(lo, off, 64);
    hi = __shfl_xor(hi, off, 64);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// partials[block] = sum over the block's grid-strided words
__global__ void __launch_bounds__(256) fingerprint_kernel(const uint64_t* __restrict__ w, long n,
                                                          uint64_t* __restrict__ partials) {
  __shared__ uint64_t red[4];
  constexpr uint64_t K = 0x9E3779B97F4A7C15ull;
  uint64_t h = 0;
  const long stride = (long)gridDim.x * blockDim.x;
  const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const ulonglong2* w2 = reinterpret_cast<const ulonglong2*>(w);  // 16-byte loads
  const long n2 = n / 2;
#pragma unroll 4
  for (long i = t0; i < n2; i += stride) {
    const ulonglong2 q = w2[i];
    h += fp_mix(q.x ^ ((uint64_t)(2 * i) * K)) + fp_mix(q.y ^ ((uint64_t)(2 * i + 1) * K));
  }
  if ((n & 1) && t0 == 0) h += fp_mix(w[n - 1] ^ ((uint64_t)(n - 1) * K));
  h = wave_sum_u64(h);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) fingerprint_sum_kernel(const uint64_t* __restrict__ partials, int n,
                                                              uint64_t* __restrict__ out) {
  __shared__ uint64_t red[4];
  uint64_t h = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) h += partials[i];
  h = wave_sum_u64(h);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

void launch_fingerprint(const void* buf, size_t bytes, uint64_t* scratch, hipStream_t st) {
  const long n = (long)(bytes / 8);
  hipLaunchKernelGGL(fingerprint_kernel, dim3(kFingerprintBlocks), dim3(256), 0, st,
                     static_cast<const uint64_t*>(buf), n, scratch + 1);
  hipLaunchKernelGGL(fingerprint_sum_kernel, dim3(1), dim3(256), 0, st, scratch + 1, kFingerprintBlocks, scratch);
}

}  // namespace qcu
