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
// A block of four waves owns 64 output sites; wave mu computes the two legs of
// direction mu (forward, then backward) for all 64, with hop_dir's arithmetic
// (hop_dir_t), and leaves their link products in LDS; then wave r adds spin row r of
// every site up over the eight legs in hop_dir's order, so the result is bit-identical
// to the SoA hop with reconstruct-12.  One call of 16^3x32 is only 1,024 such 64-site
// chunks -- one wave per SIMD if a wave did all eight legs -- so splitting the legs
// over four waves is what lets the memory latency of one wave overlap another's.
//
// The 64 neighbour spinors of 192 B a wave needs for one leg are gathered as
// consecutive 16-byte words -- lane l loads word f = 64 k + l of the list, record
// f / 12, word f % 12 -- so each load instruction reads a contiguous run wherever the
// neighbours of consecutive sites are consecutive (all but the lattice-edge rows), and
// the words reach their lanes through the wave's LDS tile, padded to 13 words per record
// so that lanes reading their own record hit distinct banks.  Each lane loads the 96 B
// of its own link's rows 0-1 directly: that leaves the block 52 KB of LDS, three blocks
// per CU instead of two with a link tile as well (20.9 against 21.4 us per 16^3x32
// call; loading the spinors directly too took 22.2 us).  The output leaves through LDS
// the same way: the block stores 768 consecutive words (each wave storing its own spin
// row's words took 21.6 us).
#include "k_hop.inc"

namespace qcu {

constexpr int kIfaceSpPad = 13;  // LDS words per spinor record (12 + 1)
// one wave's LDS region: its spinor tile (13 KB), later its legs' link products
constexpr int kIfaceWave = 64 * kIfaceSpPad;

// a wave's LDS accesses are processed in order; this keeps the compiler from moving
// the wave's own loads and stores across the point where lanes trade records
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the value v of lane j
__device__ __forceinline__ int from_lane(int v, int j) { return __builtin_amdgcn_ds_bpermute(j << 2, v); }

// link products t of leg (MU, FWD) for the wave's 64 sites: lane l's neighbour spinor
// is record `sp` of `in`, its link record `lk` of the gauge's (mu, link parity) field
template <int MU, bool FWD, int DAG>
__device__ __forceinline__ void iface_leg_t(const double2* __restrict__ in, const double2* __restrict__ gauge,
                                            long vh, int parity, double bs, int sp, int lk, double2* W, int lane,
                                            cx<double> t[2][3]) {
  const int lpar = FWD ? parity : 1 - parity;
  const double2* gb = gauge + (long)(MU * 2 + lpar) * vh * 9;
  double2* TS = W;
  double2 sw[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const int f = k * 64 + lane, j = f / 12, w = f - 12 * j;
    sw[k] = in[(long)from_lane(sp, j) * 12 + w];
  }
  // the lane's own link rows 0-1, six 16-byte words of one record
  cx<double> psi[12], u[9];
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double2 q = gb[(long)lk * 9 + e];
    u[e] = {q.x, q.y};
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const int f = k * 64 + lane, j = f / 12, w = f - 12 * j;
    TS[j * kIfaceSpPad + w] = sw[k];
  }
  wave_sync();
#pragma unroll
  for (int w = 0; w < 12; ++w) {
    const double2 q = TS[lane * kIfaceSpPad + w];
    psi[w] = {q.x, q.y};
  }
  wave_sync();  // the tile is free again
  recon_row2(u);
  hop_dir_t<MU, FWD, DAG>(u, psi, bs, t);
}

// wave MU: both legs of direction MU; their products go to the wave's region as
// [leg (0 forward, 1 backward)][6 = 3 r + i][64 lanes]
template <int MU, int DAG>
__device__ __forceinline__ void iface_wave(const double2* __restrict__ in, const double2* __restrict__ gauge,
                                           const Geom& g, int parity, double tbc, long site, int lane, double2* W) {
  int c[4];
  SiteNb n;
  decode_site(g, site, parity, c, n);
  const double bf = (MU == 3 && n.tf) ? tbc : 1.0, bb = (MU == 3 && n.tb) ? tbc : 1.0;
  cx<double> tf[2][3], tb[2][3];
  iface_leg_t<MU, true, DAG>(in, gauge, g.vh, parity, bf, (int)n.f[MU], (int)site, W, lane, tf);
  iface_leg_t<MU, false, DAG>(in, gauge, g.vh, parity, bb, (int)n.b[MU], (int)n.b[MU], W, lane, tb);
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      W[(3 * r + i) * 64 + lane] = double2{tf[r][i].re, tf[r][i].im};
      W[(6 + 3 * r + i) * 64 + lane] = double2{tb[r][i].re, tb[r][i].im};
    }
}

// o += the spin-row-R terms of leg L (direction L / 2, forward when L is even), whose
// products wave L / 2 left in S
template <int L, int DAG, int R>
__device__ __forceinline__ void iface_add_leg(const double2* S, int lane, cx<double> o[3]) {
  constexpr int MU = L / 2;
  constexpr bool FWD = (L & 1) == 0;
  constexpr int src = R < 2 ? R : GCOL[MU][R];
  const double2* P = S + MU * kIfaceWave + (L & 1) * 6 * 64;
  cx<double> t[2][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double2 q = P[(3 * src + i) * 64 + lane];
    t[src][i] = {q.x, q.y};
    t[1 - src][i] = {0.0, 0.0};
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = o[i] + leg_row<MU, FWD, DAG, R>(t, i);
}

template <int DAG, int R>
__device__ __forceinline__ void iface_row(const double2* S, int lane, cx<double> o[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = {0.0, 0.0};
  iface_add_leg<0, DAG, R>(S, lane, o);
  iface_add_leg<1, DAG, R>(S, lane, o);
  iface_add_leg<2, DAG, R>(S, lane, o);
  iface_add_leg<3, DAG, R>(S, lane, o);
  iface_add_leg<4, DAG, R>(S, lane, o);
  iface_add_leg<5, DAG, R>(S, lane, o);
  iface_add_leg<6, DAG, R>(S, lane, o);
  iface_add_leg<7, DAG, R>(S, lane, o);
}

// block b covers half sites [64 b, 64 b + 64); 256 threads, wave w = direction w, then
// spin row w
template <int DAG>
__global__ void __launch_bounds__(256, 3) hop_iface_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                          const double2* __restrict__ gauge, const Geom g,
                                                          int parity, double tbc, double coef) {
  __shared__ double2 S[4 * kIfaceWave];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long w0 = (long)blockIdx.x * 64;
  const int nlive = (int)(g.vh - w0 < 64 ? g.vh - w0 : 64);  // grid = ceil(vh / 64): nlive >= 1
  const long site = w0 + (lane < nlive ? lane : nlive - 1);   // past the end: a valid site, not stored
  double2* W = S + wv * kIfaceWave;
  switch (wv) {
    case 0: iface_wave<0, DAG>(in, gauge, g, parity, tbc, site, lane, W); break;
    case 1: iface_wave<1, DAG>(in, gauge, g, parity, tbc, site, lane, W); break;
    case 2: iface_wave<2, DAG>(in, gauge, g, parity, tbc, site, lane, W); break;
    default: iface_wave<3, DAG>(in, gauge, g, parity, tbc, site, lane, W); break;
  }
  __syncthreads();
  cx<double> o[3];
  switch (wv) {
    case 0: iface_row<DAG, 0>(S, lane, o); break;
    case 1: iface_row<DAG, 1>(S, lane, o); break;
    case 2: iface_row<DAG, 2>(S, lane, o); break;
    default: iface_row<DAG, 3>(S, lane, o); break;
  }
  __syncthreads();  // every wave has read the products; S takes the output tile
#pragma unroll
  for (int i = 0; i < 3; ++i) S[lane * kIfaceSpPad + 3 * wv + i] = double2{coef * o[i].re, coef * o[i].im};
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int f = k * 256 + (int)threadIdx.x, j = f / 12, w = f - 12 * j;
    if (j < nlive) out[(w0 + j) * 12 + w] = S[j * kIfaceSpPad + w];
  }
}

void launch_hop_iface(const void* in, void* out, const void* gauge_aos, const Geom& g, int parity, int dagger,
                      double tbc, double coef, hipStream_t st) {
  const dim3 grid((unsigned)((g.vh + 63) / 64));
  const double2* i = static_cast<const double2*>(in);
  double2* o = static_cast<double2*>(out);
  const double2* u = static_cast<const double2*>(gauge_aos);
  if (dagger) hipLaunchKernelGGL(hop_iface_kernel<1>, grid, dim3(256), 0, st, i, o, u, g, parity, tbc, coef);
  else hipLaunchKernelGGL(hop_iface_kernel<0>, grid, dim3(256), 0, st, i, o, u, g, parity, tbc, coef);
}

}  // namespace qcu
