// qcu_fields.h -- device-side complex arithmetic and the per-precision loads and
// stores of spinor, link and clover sites (layouts in qcu_common.h).
#pragma once

#include "qcu_common.h"

namespace qcu {

template <typename T>
struct cx {
  T re, im;
};

// Fused multiply-adds are written out (the library is built without implicit
// contraction), so a kernel's rounding depends only on its source.
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {fmx(a.re, b.re, -(a.im * b.im)), fmx(a.re, b.im, a.im * b.re)};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> mulc(cx<T> a, cx<T> b) {
  return {fmx(a.re, b.re, a.im * b.im), fmx(a.im, b.re, -(a.re * b.im))};
}
// c + a b
template <typename T>
__device__ __forceinline__ cx<T> cmad(cx<T> a, cx<T> b, cx<T> c) {
  return {fmx(a.re, b.re, fmx(-a.im, b.im, c.re)), fmx(a.re, b.im, fmx(a.im, b.re, c.im))};
}
// c + conj(a) b
template <typename T>
__device__ __forceinline__ cx<T> cmadc(cx<T> a, cx<T> b, cx<T> c) {
  return {fmx(a.re, b.re, fmx(a.im, b.im, c.re)), fmx(a.re, b.im, fmx(-a.im, b.re, c.im))};
}
// y + s a (real s)
template <typename T>
__device__ __forceinline__ cx<T> axpy(T s, cx<T> a, cx<T> y) { return {fmx(s, a.re, y.re), fmx(s, a.im, y.im)}; }
// multiply by i^k (k = 0..3)
template <typename T>
__device__ __forceinline__ cx<T> iph(cx<T> a, int k) {
  k &= 3;
  return k == 0 ? a : k == 1 ? cx<T>{-a.im, a.re} : k == 2 ? cx<T>{-a.re, -a.im} : cx<T>{a.im, -a.re};
}
template <typename T>
__device__ __forceinline__ T norm2(cx<T> a) { return fmx(a.re, a.re, a.im * a.im); }

// streaming loads (nontemporal hint) of the HIP vector types through clang vectors
typedef double v2d_t __attribute__((ext_vector_type(2)));
typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef float v2f_t __attribute__((ext_vector_type(2)));
typedef unsigned v4u_t __attribute__((ext_vector_type(4)));
typedef unsigned v2u_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_nt(const double2* p) {
  v2d_t v = __builtin_nontemporal_load(reinterpret_cast<const v2d_t*>(p));
  return double2{v.x, v.y};
}
__device__ __forceinline__ float4 ld_nt(const float4* p) {
  v4f_t v = __builtin_nontemporal_load(reinterpret_cast<const v4f_t*>(p));
  return float4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ float2 ld_nt(const float2* p) {
  v2f_t v = __builtin_nontemporal_load(reinterpret_cast<const v2f_t*>(p));
  return float2{v.x, v.y};
}
__device__ __forceinline__ uint4 ld_nt(const uint4* p) {
  v4u_t v = __builtin_nontemporal_load(reinterpret_cast<const v4u_t*>(p));
  return uint4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ uint2 ld_nt(const uint2* p) {
  v2u_t v = __builtin_nontemporal_load(reinterpret_cast<const v2u_t*>(p));
  return uint2{v.x, v.y};
}
__device__ __forceinline__ unsigned ld_nt(const unsigned* p) { return __builtin_nontemporal_load(p); }
// nontemporal (streaming) vector stores
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
  __builtin_nontemporal_store(v2d_t{v.x, v.y}, reinterpret_cast<v2d_t*>(p));
}
__device__ __forceinline__ void st_nt(float4* p, float4 v) {
  __builtin_nontemporal_store(v4f_t{v.x, v.y, v.z, v.w}, reinterpret_cast<v4f_t*>(p));
}
__device__ __forceinline__ void st_nt(uint4* p, uint4 v) {
  __builtin_nontemporal_store(v4u_t{v.x, v.y, v.z, v.w}, reinterpret_cast<v4u_t*>(p));
}
__device__ __forceinline__ void st_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }
// NTS = 1: nontemporal store (nt_store() in qcu_common.h says when); NTS = 0: ordinary
template <int NTS, typename V>
__device__ __forceinline__ void stp(V* p, V v) {
  if (NTS) st_nt(p, v);
  else *p = v;
}
// NT = 1: nontemporal (streaming) load, which keeps data read once out of the caches'
// way; NT = 0: ordinary load, which lets a field that fits the 256 MB Infinity Cache
// be served from there on the next launch.  Which is faster depends on the lattice, so
// the hop's launch shape carries the choice.
template <int NT, typename V>
__device__ __forceinline__ V ldp(const V* p) {
  if (NT) return ld_nt(p);
  return *p;
}

template <int P>
struct Real { using T = float; };
template <>
struct Real<F64> { using T = double; };

// ---------------------------------------------------------------- spinors ---
// Spinor fields are stored in blocks of kSiteBlock consecutive half sites: block b
// holds word w of its sites at [(b * W + w) * kSiteBlock + site % kSiteBlock], W words
// per site (F64: 12 double2, F32: 6 float4, F16: 3 uint4; the F16 per-site norms
// follow as one float plane).  A wave's load of one word is a contiguous 1 KB run as
// with plain planes, and a block's words sit together, so the streaming kernels open
// far fewer DRAM rows than with 6-24 planes a field apart (a CG update measured
// 5.2 TB/s against 4.6 TB/s on 32^3x64 with 12 right-hand sides).
template <int W, typename V>
__device__ __forceinline__ V* site_block(V* base, long i) {
  return base + (i >> kSiteBlockLog) * (long)(W * kSiteBlock) + (i & (kSiteBlock - 1));
}
template <int W, typename V>
__device__ __forceinline__ const V* site_block(const V* base, long i) {
  return base + (i >> kSiteBlockLog) * (long)(W * kSiteBlock) + (i & (kSiteBlock - 1));
}

template <int P>
struct SpinorIO;

template <>
struct SpinorIO<F64> {
  using T = double;
  static __device__ __forceinline__ const double2* at(const void* base, long i) {
    return site_block<12>(static_cast<const double2*>(base), i);
  }
  static __device__ __forceinline__ void load(const void* base, long vh, long i, cx<T> v[12]) {
    const double2* b = at(base, i);
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      double2 q = ld_nt(b + c * kSiteBlock);
      v[c] = {q.x, q.y};
    }
  }
  static __device__ __forceinline__ void load_cached(const void* base, long vh, long i, cx<T> v[12]) {
    const double2* b = at(base, i);
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      double2 q = b[c * kSiteBlock];
      v[c] = {q.x, q.y};
    }
  }
  template <int NT = 1>
  static __device__ __forceinline__ void load_half(const void* base, long vh, long i, int h, cx<T> v[6]) {
    const double2* b = at(base, i) + 6 * h * kSiteBlock;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double2 q = ldp<NT>(b + c * kSiteBlock);
      v[c] = {q.x, q.y};
    }
  }
  template <int NTS = 0>
  static __device__ __forceinline__ void store(void* base, long vh, long i, const cx<T> v[12]) {
    double2* b = site_block<12>(static_cast<double2*>(base), i);
#pragma unroll
    for (int c = 0; c < 12; ++c) stp<NTS>(b + c * kSiteBlock, double2{v[c].re, v[c].im});
  }
  template <int NTS = 0>
  static __device__ __forceinline__ void store_half(void* base, long vh, long i, int h, const cx<T> v[6]) {
    double2* b = site_block<12>(static_cast<double2*>(base), i) + 6 * h * kSiteBlock;
#pragma unroll
    for (int c = 0; c < 6; ++c) stp<NTS>(b + c * kSiteBlock, double2{v[c].re, v[c].im});
  }
};

template <>
struct SpinorIO<F32> {
  using T = float;
  static __device__ __forceinline__ const float4* at(const void* base, long i) {
    return site_block<6>(static_cast<const float4*>(base), i);
  }
  static __device__ __forceinline__ void load_cached(const void* base, long vh, long i, cx<T> v[12]) {
    const float4* b = at(base, i);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float4 q = b[k * kSiteBlock];
      v[2 * k] = {q.x, q.y};
      v[2 * k + 1] = {q.z, q.w};
    }
  }
  // streaming (BLAS) loads of a site's own spinor: nontemporal
  static __device__ __forceinline__ void load(const void* base, long vh, long i, cx<T> v[12]) {
    const float4* b = at(base, i);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float4 q = ld_nt(b + k * kSiteBlock);
      v[2 * k] = {q.x, q.y};
      v[2 * k + 1] = {q.z, q.w};
    }
  }
  template <int NT = 0>
  static __device__ __forceinline__ void load_half(const void* base, long vh, long i, int h, cx<T> v[6]) {
    const float4* b = at(base, i) + 3 * h * kSiteBlock;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float4 q = ldp<NT>(b + k * kSiteBlock);
      v[2 * k] = {q.x, q.y};
      v[2 * k + 1] = {q.z, q.w};
    }
  }
  template <int NTS = 0>
  static __device__ __forceinline__ void store(void* base, long vh, long i, const cx<T> v[12]) {
    float4* b = site_block<6>(static_cast<float4*>(base), i);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      stp<NTS>(b + k * kSiteBlock, float4{v[2 * k].re, v[2 * k].im, v[2 * k + 1].re, v[2 * k + 1].im});
  }
};

__device__ __forceinline__ float h2f(unsigned short h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ unsigned short f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }
__device__ __forceinline__ void unpack8(uint4 q, float f[8]) {
  unsigned w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[2 * k] = h2f((unsigned short)(w[k] & 0xffff));
    f[2 * k + 1] = h2f((unsigned short)(w[k] >> 16));
  }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
  unsigned w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = (unsigned)f2h(f[2 * k]) | ((unsigned)f2h(f[2 * k + 1]) << 16);
  return uint4{w[0], w[1], w[2], w[3]};
}

// half precision with one float scale per site: value = half * norm,
// norm = max |real component| of the site
template <>
struct SpinorIO<F16> {
  using T = float;
  static __device__ __forceinline__ const float* norms(const void* base, long vh) {
    return reinterpret_cast<const float*>(static_cast<const uint4*>(base) + 3 * padded_sites(vh));
  }
  static __device__ __forceinline__ const uint4* at(const void* base, long i) {
    return site_block<3>(static_cast<const uint4*>(base), i);
  }
  static __device__ __forceinline__ void load_cached(const void* base, long vh, long i, cx<T> v[12]) {
    const uint4* b = at(base, i);
    float nrm = norms(base, vh)[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float f[8];
      unpack8(b[k * kSiteBlock], f);
#pragma unroll
      for (int j = 0; j < 4; ++j) v[4 * k + j] = {f[2 * j] * nrm, f[2 * j + 1] * nrm};
    }
  }
  // streaming (BLAS) loads of a site's own spinor: nontemporal
  static __device__ __forceinline__ void load(const void* base, long vh, long i, cx<T> v[12]) {
    const uint4* b = at(base, i);
    float nrm = __uint_as_float(ld_nt(reinterpret_cast<const unsigned*>(norms(base, vh) + i)));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float f[8];
      unpack8(ld_nt(b + k * kSiteBlock), f);
#pragma unroll
      for (int j = 0; j < 4; ++j) v[4 * k + j] = {f[2 * j] * nrm, f[2 * j + 1] * nrm};
    }
  }
  // components 0..5 are word 0 and the first half of word 1; 6..11 the rest
  template <int NT = 0>
  static __device__ __forceinline__ void load_half(const void* base, long vh, long i, int h, cx<T> v[6]) {
    const uint4* b = at(base, i) + h * kSiteBlock;
    const unsigned* np = reinterpret_cast<const unsigned*>(norms(base, vh) + i);
    float nrm = __uint_as_float(ldp<NT>(np));
    float f[16];
    unpack8(ldp<NT>(b), f);
    unpack8(ldp<NT>(b + kSiteBlock), f + 8);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int o = h ? 2 * c + 4 : 2 * c;
      v[c] = {f[o] * nrm, f[o + 1] * nrm};
    }
  }
  template <int NTS = 0>
  static __device__ __forceinline__ void store(void* base, long vh, long i, const cx<T> v[12]) {
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 12; ++c) m = fmaxf(m, fmaxf(fabsf(v[c].re), fabsf(v[c].im)));
    // 1 / m overflows float for m below 2^-128 (denormal norms): such a site is scaled
    // by 2^64 first, exactly, so its halves stay in [-1, 1]; other sites see s = 1
    const float s = m < 0x1p-100f ? 0x1p64f : 1.0f;
    const float inv = m > 0.f ? 1.0f / (m * s) : 0.f;
    uint4* b = site_block<3>(static_cast<uint4*>(base), i);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float f[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        f[2 * j] = (v[4 * k + j].re * s) * inv;
        f[2 * j + 1] = (v[4 * k + j].im * s) * inv;
      }
      stp<NTS>(b + k * kSiteBlock, pack8(f));
    }
    stp<NTS>(reinterpret_cast<float*>(static_cast<uint4*>(base) + 3 * padded_sites(vh)) + i, m);
  }
};

// ------------------------------------------------------------------ links ---
// row 2 of an SU(3) link from rows 0 and 1: conj(row0 x row1)
template <typename T>
__device__ __forceinline__ void recon_row2(cx<T> u[9]) {
  u[6] = conj(u[1] * u[5] - u[2] * u[4]);
  u[7] = conj(u[2] * u[3] - u[0] * u[5]);
  u[8] = conj(u[0] * u[4] - u[1] * u[3]);
}

template <int P, int R>
struct LinkIO;

// Packed links: field f = mu * 2 + parity holds NE words per site, either in blocks of
// kSiteBlock sites like the spinors (word e of site i at link_word(...)[e *
// link_stride]) or, for FP64 with all 18 reals, in one plane per word.  Measured on
// MI355X, blocks made the FP32 and FP16 hops and the FP64 recon-12 hop 2-18 % faster
// and the FP64 recon-18 hop 4-11 % slower, so each form keeps the faster layout.
template <int P, int R>
struct LinkLayout {
  static constexpr bool blocked = !(P == F64 && R == 18);
  static constexpr int NE = P == F64 ? (R == 12 ? 6 : 9) : (R == 12 ? 3 : 9);  // words per link
};
template <int P, int R>
__host__ __device__ inline long link_stride(long vh) {
  return LinkLayout<P, R>::blocked ? (long)kSiteBlock : vh;
}
template <int P, int R, typename V>
__device__ __forceinline__ V* link_word(V* g, long vh, int f, long i) {
  constexpr int NE = LinkLayout<P, R>::NE;
  if (LinkLayout<P, R>::blocked)
    return g + (long)f * NE * padded_sites(vh) + (i >> kSiteBlockLog) * (long)(NE * kSiteBlock) +
           (i & (kSiteBlock - 1));
  return g + (long)f * NE * padded_sites(vh) + i;
}

template <int R>
struct LinkIO<F64, R> {
  using T = double;
  static constexpr int NE = R == 12 ? 6 : 9;
  template <int NT = 1>
  static __device__ __forceinline__ void load(const void* g, long vh, int mu, int par, long i, cx<T> u[9]) {
    const double2* b = link_word<F64, R>(static_cast<const double2*>(g), vh, mu * 2 + par, i);
    const long ws = link_stride<F64, R>(vh);
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      double2 q = ldp<NT>(b + e * ws);
      u[e] = {q.x, q.y};
    }
    if (R == 12) recon_row2(u);
  }
};

template <int R>
struct LinkIO<F32, R> {
  using T = float;
  template <int NT = 1>
  static __device__ __forceinline__ void load(const void* g, long vh, int mu, int par, long i, cx<T> u[9]) {
    if (R == 12) {
      const float4* b = link_word<F32, R>(static_cast<const float4*>(g), vh, mu * 2 + par, i);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float4 q = ldp<NT>(b + k * kSiteBlock);
        u[2 * k] = {q.x, q.y};
        u[2 * k + 1] = {q.z, q.w};
      }
      recon_row2(u);
    } else {
      const float2* b = link_word<F32, R>(static_cast<const float2*>(g), vh, mu * 2 + par, i);
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        float2 q = ldp<NT>(b + e * kSiteBlock);
        u[e] = {q.x, q.y};
      }
    }
  }
};

template <int R>
struct LinkIO<F16, R> {
  using T = float;
  template <int NT = 1>
  static __device__ __forceinline__ void load(const void* g, long vh, int mu, int par, long i, cx<T> u[9]) {
    if (R == 12) {
      const uint2* b = link_word<F16, R>(static_cast<const uint2*>(g), vh, mu * 2 + par, i);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        uint2 q = ldp<NT>(b + k * kSiteBlock);
        u[2 * k] = {h2f(q.x & 0xffff), h2f(q.x >> 16)};
        u[2 * k + 1] = {h2f(q.y & 0xffff), h2f(q.y >> 16)};
      }
      recon_row2(u);
    } else {
      const unsigned* b = link_word<F16, R>(static_cast<const unsigned*>(g), vh, mu * 2 + par, i);
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        unsigned q = ldp<NT>(b + e * kSiteBlock);
        u[e] = {h2f(q & 0xffff), h2f(q >> 16)};
      }
    }
  }
};

// ----------------------------------------------------------------- clover ---
// 72 reals per site: block b at reals [36b, 36b+36): d0..d5, then the 15 strictly
// lower entries (1,0),(2,0),(2,1),(3,0),...,(5,4) as (re, im) pairs.  Stored per
// parity in blocks of kSiteBlock sites like the spinors (site_block), 36 double2,
// 18 float4 or 9 uint4 words per site.
template <int P>
struct CloverIO;

template <>
struct CloverIO<F64> {
  using T = double;
  static __device__ __forceinline__ void load(const void* c, long vh, long i, T a[72]) {
    const double2* b = site_block<36>(static_cast<const double2*>(c), i);
#pragma unroll
    for (int k = 0; k < 36; ++k) {
      double2 q = ld_nt(b + k * kSiteBlock);
      a[2 * k] = q.x;
      a[2 * k + 1] = q.y;
    }
  }
  // chiral block b (36 reals) only
  template <int NT = 1>
  static __device__ __forceinline__ void load_block(const void* c, long vh, long i, int b, T m[36]) {
    const double2* p = site_block<36>(static_cast<const double2*>(c), i) + 18 * b * kSiteBlock;
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      double2 q = ldp<NT>(p + k * kSiteBlock);
      m[2 * k] = q.x;
      m[2 * k + 1] = q.y;
    }
  }
};
template <>
struct CloverIO<F32> {
  using T = float;
  static __device__ __forceinline__ void load(const void* c, long vh, long i, T a[72]) {
    const float4* b = site_block<18>(static_cast<const float4*>(c), i);
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      float4 q = ld_nt(b + k * kSiteBlock);
      a[4 * k] = q.x;
      a[4 * k + 1] = q.y;
      a[4 * k + 2] = q.z;
      a[4 * k + 3] = q.w;
    }
  }
  template <int NT = 1>
  static __device__ __forceinline__ void load_block(const void* c, long vh, long i, int b, T m[36]) {
    const float4* p = site_block<18>(static_cast<const float4*>(c), i) + 9 * b * kSiteBlock;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      float4 q = ldp<NT>(p + k * kSiteBlock);
      m[4 * k] = q.x;
      m[4 * k + 1] = q.y;
      m[4 * k + 2] = q.z;
      m[4 * k + 3] = q.w;
    }
  }
};
template <>
struct CloverIO<F16> {
  using T = float;
  static __device__ __forceinline__ void load(const void* c, long vh, long i, T a[72]) {
    const uint4* b = site_block<9>(static_cast<const uint4*>(c), i);
#pragma unroll
    for (int k = 0; k < 9; ++k) unpack8(ld_nt(b + k * kSiteBlock), a + 8 * k);
  }
  // block 0 is halves 0..35 (words 0..4, first half of word 4), block 1 is 36..71
  template <int NT = 1>
  static __device__ __forceinline__ void load_block(const void* c, long vh, long i, int b, T m[36]) {
    const uint4* p = site_block<9>(static_cast<const uint4*>(c), i) + 4 * b * kSiteBlock;
    float f[40];
#pragma unroll
    for (int k = 0; k < 5; ++k) unpack8(ldp<NT>(p + k * kSiteBlock), f + 8 * k);
#pragma unroll
    for (int k = 0; k < 36; ++k) m[k] = b ? f[k + 4] : f[k];
  }
};

// y = M x for one packed Hermitian 6x6 chiral block (6 real diagonals, then the 15
// complex entries below the diagonal row by row)
template <typename T>
__device__ __forceinline__ void clover_mul_block(const T m[36], const cx<T> x[6], cx<T> y[6]) {
#pragma unroll
  for (int r = 0; r < 6; ++r) y[r] = m[r] * x[r];
  int k = 6;
#pragma unroll
  for (int r = 1; r < 6; ++r) {
#pragma unroll
    for (int c = 0; c < r; ++c) {
      cx<T> l = {m[k], m[k + 1]};
      k += 2;
      y[r] = cmad(l, x[c], y[r]);
      y[c] = cmadc(l, x[r], y[c]);
    }
  }
}

// out = blockdiag(M0, M1) v for one site, M packed as above
template <typename T>
__device__ __forceinline__ void clover_mul(const T a[72], const cx<T> v[12], cx<T> out[12]) {
#pragma unroll
  for (int b = 0; b < 2; ++b) clover_mul_block(a + 36 * b, v + 6 * b, out + 6 * b);
}

// XCD column groups: the hardware hands workgroup b to XCD x = b mod 8.  Give that XCD
// W consecutive 64-site chunks of every Q-chunk sub-plane of the walk, in one of
// R = 8 W / Q ranges of the NP sub-planes, in walk order.  Each XCD then sweeps a
// column W chunks wide: the sub-planes a site's t and z neighbours live in are a few
// steps away on the same XCD, so its L2 holds them, while the column's edge rows are
// fetched by two XCDs.  W = 1 is plain round-robin.  A bijection on [0, NP Q) when
// Q % W == 0, 8 W % Q == 0 and NP % R == 0 (the launcher checks).
__device__ __forceinline__ unsigned colgroup_block(unsigned b, int w, int q, int r, long np) {
  const unsigned x = b & 7u, k = b >> 3;
  const unsigned per = (unsigned)(q / w);  // XCDs per range
  const unsigned rr = x / per, gx = x - rr * per;
  const unsigned pl = k / (unsigned)w, i = k - pl * (unsigned)w;
  const unsigned P = rr * (unsigned)(np / r) + pl;
  return P * (unsigned)q + gx * (unsigned)w + i;
}

// XCD-aware block order: the hardware hands workgroup b to XCD (b mod 8); remap so
// that each XCD walks one contiguous range of sites (its L2 then sees neighbours).
__device__ __forceinline__ unsigned remap_block(unsigned b, unsigned nb, int on) {
  if (!on) return b;
  unsigned x = b & 7u, k = b >> 3, q = nb >> 3, r = nb & 7u;
  return x * q + (x < r ? x : r) + k;
}

}  // namespace qcu
