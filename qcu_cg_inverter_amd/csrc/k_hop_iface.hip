// k_hop_iface.hip -- the hopping term read and written in the interface layouts, the
// kernel behind dslashQcu.
//
// The reference's dslashQcu returns mpiDslashNaive (src/qcu_wilson_dslash_neo.cu:
// 941-1206): the half spinor [Vh][4][3] and the gauge [mu][parity][Vh][3][3] are read
// as complex128 records straight from the caller's buffers, every call, and only rows
// 0-1 of each link (row 2 = conj(row0 x row1), :73-86).  This kernel does the same in
// one pass with no conversion kernels and no copy of the gauge, so a gauge the caller
// changed in place is always the one applied.
//
// A thread owns one output site; the arithmetic is hop_dir of k_hop.inc, in the same
// leg order, so the result is bit-identical to the SoA hop with reconstruct-12.  The
// records a wave needs for one leg (64 neighbour spinors of 192 B, or 64 links of
// which 96 B are read) are gathered as consecutive 16-byte words -- lane l loads word
// f = 64 k + l of the list, record f / W, word f % W -- so each load instruction
// reads a contiguous run wherever the neighbours of consecutive sites are consecutive
// (all but the lattice-edge rows).  The words go through the wave's LDS tile, padded
// to an odd number of 16-byte words per record so that lanes reading their own record
// hit distinct banks.  The output goes back the same way: each lane writes its 12
// words to the tile and the wave stores 768 consecutive words.
#include "k_hop.inc"

namespace qcu {

constexpr int kIfaceSpPad = 13;  // LDS words per spinor record (12 + 1)
constexpr int kIfaceLkPad = 7;   // LDS words per link record (6 + 1)

// a wave's LDS accesses are processed in order; this keeps the compiler from moving
// the wave's own loads and stores across the point where lanes trade records
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// keeps the compiler from computing a lane's tile offsets once and holding them in
// registers across all eight legs (they are cheaper to recompute than to keep)
__device__ __forceinline__ int fresh(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// leg L = 2 mu + (backward): neighbour spinor records I[L][.], link records J[L][.]
template <int MU, bool FWD, int DAG>
__device__ __forceinline__ void iface_leg(const double2* __restrict__ in, const double2* __restrict__ gauge,
                                          long vh, int parity, double bs, double2* T, const int* I, const int* J,
                                          int lane_in, cx<double> acc[12]) {
  const int lane = fresh(lane_in);
  const int lpar = FWD ? parity : 1 - parity;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const int f = k * 64 + lane, j = f / 12, w = f - 12 * j;
    T[j * kIfaceSpPad + w] = in[(long)I[j] * 12 + w];
  }
  wave_sync();
  cx<double> psi[12];
#pragma unroll
  for (int w = 0; w < 12; ++w) {
    const double2 q = T[lane * kIfaceSpPad + w];
    psi[w] = {q.x, q.y};
  }
  wave_sync();
  const double2* gb = gauge + (long)(MU * 2 + lpar) * vh * 9;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int f = k * 64 + lane, j = f / 6, w = f - 6 * j;
    T[j * kIfaceLkPad + w] = gb[(long)J[j] * 9 + w];
  }
  wave_sync();
  cx<double> u[9];
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double2 q = T[lane * kIfaceLkPad + e];
    u[e] = {q.x, q.y};
  }
  recon_row2(u);
  wave_sync();  // the next leg overwrites T
  hop_dir<MU, FWD, DAG>(acc, u, psi, bs);
}

// one wave per block; wave b covers half sites [64 b, 64 b + 64)
template <int DAG>
__global__ void __launch_bounds__(64) hop_iface_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                       const double2* __restrict__ gauge, const Geom g, int parity,
                                                       double tbc, double coef) {
  __shared__ double2 T[64 * kIfaceSpPad];
  __shared__ int I[8][64], J[8][64];
  const int lane = threadIdx.x;
  const long w0 = (long)blockIdx.x * 64;
  const int nlive = (int)(g.vh - w0 < 64 ? g.vh - w0 : 64);  // grid = ceil(vh / 64): nlive >= 1
  const long site = w0 + (lane < nlive ? lane : nlive - 1);   // past the end: a valid site, not stored
  double bs_tf, bs_tb;
  {
    int c[4];
    SiteNb n;
    decode_site(g, site, parity, c, n);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      I[2 * mu][lane] = (int)n.f[mu];
      J[2 * mu][lane] = (int)site;  // U_mu(x)
      I[2 * mu + 1][lane] = (int)n.b[mu];
      J[2 * mu + 1][lane] = (int)n.b[mu];  // U_mu(x - mu)
    }
    bs_tf = n.tf ? tbc : 1.0;
    bs_tb = n.tb ? tbc : 1.0;
  }
  wave_sync();
  cx<double> acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = {0.0, 0.0};
  const double one = 1.0;
  iface_leg<0, true, DAG>(in, gauge, g.vh, parity, one, T, I[0], J[0], lane, acc);
  iface_leg<0, false, DAG>(in, gauge, g.vh, parity, one, T, I[1], J[1], lane, acc);
  iface_leg<1, true, DAG>(in, gauge, g.vh, parity, one, T, I[2], J[2], lane, acc);
  iface_leg<1, false, DAG>(in, gauge, g.vh, parity, one, T, I[3], J[3], lane, acc);
  iface_leg<2, true, DAG>(in, gauge, g.vh, parity, one, T, I[4], J[4], lane, acc);
  iface_leg<2, false, DAG>(in, gauge, g.vh, parity, one, T, I[5], J[5], lane, acc);
  iface_leg<3, true, DAG>(in, gauge, g.vh, parity, bs_tf, T, I[6], J[6], lane, acc);
  iface_leg<3, false, DAG>(in, gauge, g.vh, parity, bs_tb, T, I[7], J[7], lane, acc);
  const int ln = fresh(lane);
#pragma unroll
  for (int w = 0; w < 12; ++w) T[ln * kIfaceSpPad + w] = double2{coef * acc[w].re, coef * acc[w].im};
  wave_sync();
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const int f = k * 64 + ln, j = f / 12, w = f - 12 * j;
    if (j < nlive) out[(w0 + j) * 12 + w] = T[j * kIfaceSpPad + w];
  }
}

void launch_hop_iface(const void* in, void* out, const void* gauge_aos, const Geom& g, int parity, int dagger,
                      double tbc, double coef, hipStream_t st) {
  const dim3 grid((unsigned)((g.vh + 63) / 64));
  const double2* i = static_cast<const double2*>(in);
  double2* o = static_cast<double2*>(out);
  const double2* u = static_cast<const double2*>(gauge_aos);
  if (dagger) hipLaunchKernelGGL(hop_iface_kernel<1>, grid, dim3(64), 0, st, i, o, u, g, parity, tbc, coef);
  else hipLaunchKernelGGL(hop_iface_kernel<0>, grid, dim3(64), 0, st, i, o, u, g, parity, tbc, coef);
}

}  // namespace qcu
