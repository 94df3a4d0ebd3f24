// qcu_common.h -- shared host/device definitions of the MI355X qcu library.
//
// Internal field layouts in HBM (Vh = half-lattice volume, site index
// i = ((t*Lz + z)*Ly + y)*(Lx/2) + xh, the reference's even-odd order,
// include/qcu_point.cuh:70-72):
//   spinor  F64: double2 [12][Vh]            (component c = spin*3 + colour)
//           F32: float4  [6][Vh]             (components 2k, 2k+1)
//           F16: uint4   [3][Vh] + float norm[Vh]   (8 halves per uint4, scaled by norm)
//   gauge   [mu 4][parity 2][entries][Vh] planes; entries = 9 complex (recon-18) or
//           6 complex (recon-12, rows 0-1; row 2 rebuilt in registers)
//           F64: double2 per complex; F32: float2 (r18) / float4 pairs (r12);
//           F16: half2 (r18) / 4-half uint2 pairs (r12)
//   clover  [parity 2][36 double2][Vh] per precision (F32: [18] float4, F16: [9] uint4):
//           two Hermitian 6x6 chiral blocks, each packed as 6 real diagonals
//           followed by the 15 complex strictly-lower entries (row-major).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace qcu {

enum Prec : int { F64 = 0, F32 = 1, F16 = 2 };

inline const char* prec_name(int p) { return p == F64 ? "f64" : p == F32 ? "f32" : "f16"; }

// 32-bit division by a runtime constant through a multiply-high
// (round-up method: q = (umulhi(n, m) + n) >> s, exact for n < 2^31).
struct FastDiv {
  uint32_t d, m, s;
  FastDiv() : d(1), m(0), s(0) {}
  explicit FastDiv(uint32_t div) : d(div) {
    uint32_t l = 0;
    while ((1ull << l) < div) ++l;
    s = l;
    m = (uint32_t)((((1ull << l) - div) << 32) / div + 1);
    if (div == 1) { m = 0; s = 0; }
  }
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t hi = __umulhi(n, m);
#else
    uint32_t hi = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
    return (hi + n) >> s;
  }
};

// Lattice geometry of one even-odd half lattice.
struct Geom {
  int L[4];       // Lx, Ly, Lz, Lt
  int lxh;        // Lx / 2
  long vh;        // half volume
  long sy, sz, st;  // index strides of y, z, t in half-site units
  FastDiv dx, dy, dz;
};

// Order in which the hopping kernel walks the sites.  tile_z = 0: storage order.
// Otherwise (y, z, t) is cut into tile_y x tile_z x tile_t tiles walked one after
// another (y tiles fastest, then z, then t).  A tile is a stack of sub-planes (all x,
// tile_y rows of y) run with t fastest, then z, so a site's t-neighbours are one
// sub-plane away and its z-neighbours tile_t sub-planes away instead of a whole
// t-slice: the neighbour spinors of a batch of right-hand sides stay in the XCD's L2.
// A sub-plane is a run of consecutive storage sites whose length is a multiple of 64,
// so a wave still covers one aligned 64-site storage chunk.
struct Walk {
  int tile_y, tile_z, tile_t, ny, ntz;
  FastDiv dsp, dtt, dtz, dny, dntz;
};

inline Geom make_geom(const int L[4]) {
  Geom g;
  for (int i = 0; i < 4; ++i) g.L[i] = L[i];
  g.lxh = L[0] / 2;
  g.vh = (long)L[0] * L[1] * L[2] * L[3] / 2;
  g.sy = g.lxh;
  g.sz = g.sy * L[1];
  g.st = g.sz * L[2];
  g.dx = FastDiv((uint32_t)g.lxh);
  g.dy = FastDiv((uint32_t)L[1]);
  g.dz = FastDiv((uint32_t)L[2]);
  return g;
}

// Spinor fields are stored in blocks of kSiteBlock half sites (SpinorIO), so a field
// spans the half volume rounded up to whole blocks; the sites past vh are never read.
constexpr int kSiteBlockLog = 6;
constexpr int kSiteBlock = 1 << kSiteBlockLog;
__host__ __device__ inline long padded_sites(long vh) { return (vh + kSiteBlock - 1) & ~(long)(kSiteBlock - 1); }
inline size_t spinor_bytes(int prec, long vh) {
  const size_t n = (size_t)padded_sites(vh);
  return prec == F64 ? n * 192 : prec == F32 ? n * 96 : n * 52;
}
// 1: a launch writing `nrhs` fields of `prec` stores them nontemporally, and a hop
// reads its epilogue's x that way too.  Streamed fields that overflow about half the
// 256 MB Infinity Cache would only evict what the next kernel reads; smaller ones are
// served from it to that kernel.  Measured on 32^3x64: nontemporal stores make the
// FP64 hop 5-7 % and the 12-RHS FP32 hop 4 % faster, and the 1-RHS FP32 hop (100 MB
// written) 6 % slower; on 16^3x32 the FP64 hop (13 MB) 7 % slower.
constexpr size_t kNtStoreBytes = (size_t)128 << 20;
inline int nt_store(int prec, long vh, int nrhs) {
  return (size_t)nrhs * spinor_bytes(prec, vh) > kNtStoreBytes ? 1 : 0;
}
inline int link_entries(int recon) { return recon == 12 ? 6 : 9; }
inline size_t gauge_bytes(int prec, int recon, long vh) {
  size_t cplx = (size_t)8 * link_entries(recon) * padded_sites(vh);  // 4 mu x 2 parities, site blocks
  return cplx * (prec == F64 ? 16 : prec == F32 ? 8 : 4);
}
inline size_t clover_bytes(int prec, long vh) {  // both parities, site blocks
  return (size_t)2 * padded_sites(vh) * 72 * (prec == F64 ? 8 : prec == F32 ? 4 : 2);
}

// Epilogue of the hopping kernel: what is written for out(parity p), where
// h = D^s_{p,1-p} in and x is a field of parity p.
enum Epi : int {
  EPI_PLAIN = 0,     // out = a * h
  EPI_XPAY = 1,      // out = x + a * h
  EPI_CLOV = 2,      // out = C (a * h)            (C = A^-1 of parity p)
  EPI_CLOV_XPAY = 3  // out = C x + a * h          (C = A of parity p)
};

constexpr int kMaxRhs = 16;

// One hopping-kernel launch.  Fields of a batch of right-hand sides are strided in
// memory (rhs r at base + r * stride bytes); clover / coefficient may differ per rhs.
struct HopArgs {
  const void* in;
  void* out;
  const void* x;
  const void* gauge;
  long in_stride, out_stride, x_stride;  // bytes between right-hand sides
  const void* clov[kMaxRhs];             // packed clover of parity p per rhs (may repeat)
  double coef[kMaxRhs];
  double* partials;                      // |out|^2 per block per rhs, or null
  Geom geo;
  int parity;
  int dagger;                            // 0: s = +1, 1: s = -1
  double tbc;                            // t boundary sign
  int nrhs;
  int xcd_remap;                         // contiguous site ranges per XCD
  int sg;                                // sites per right-hand side group (set by the launcher)
  int ngroups;                           // blocks per site range (set by the launcher)
  int gmajor;                            // 1: right-hand side group is the slowest block index
  int rpw;                               // right-hand sides per wave (1, 2 or 4)
  int spt;                               // site blocks walked by each block (sites per thread)
  int cached;                            // 1: ordinary loads of links / clover / x, 0: streaming
  long nsb;                              // site blocks of the lattice (set by the launcher)
  int clov_shared;                       // every right-hand side uses clov[0]
  int nt_store;                          // nontemporal stores of out and loads of x (set by the launcher)
  Walk walk;                             // site traversal order
  int lds;                               // 1: hop_lds_kernel (every right-hand side per block, links in LDS),
                                         // 2: and a shared clover in LDS
  int cgw, cgq, cgr;                     // XCD column groups (xcd_remap >= 2, set by the launcher): W chunks
  long cgnp;                             // per XCD of each Q-chunk sub-plane, R ranges of the NP sub-planes
};

// walk with (tile_z, tile_t) tiles, or storage order when the tiles do not divide the
// lattice or a plane is not a multiple of 64 sites
inline Walk make_walk(const Geom& g, int ty, int tz, int tt) {
  Walk w;
  w.tile_y = w.tile_z = w.tile_t = w.ny = w.ntz = 0;
  if (ty <= 0) ty = g.L[1];
  const long sub = (long)g.lxh * ty;
  if (tz <= 0 || tt <= 0 || g.L[1] % ty != 0 || sub % 64 != 0 || g.L[2] % tz != 0 || g.L[3] % tt != 0) return w;
  w.tile_y = ty;
  w.tile_z = tz;
  w.tile_t = tt;
  w.ny = g.L[1] / ty;
  w.ntz = g.L[2] / tz;
  w.dsp = FastDiv((uint32_t)sub);
  w.dtt = FastDiv((uint32_t)tt);
  w.dtz = FastDiv((uint32_t)tz);
  w.dny = FastDiv((uint32_t)w.ny);
  w.dntz = FastDiv((uint32_t)w.ntz);
  return w;
}

}  // namespace qcu
