// qcu_kernels.h -- host-side launchers of the HIP kernels (all asynchronous on the
// given stream).  Fields are one half lattice in the internal layout unless noted;
// a batch of right-hand sides is strided by `stride` bytes.
#pragma once

#include "qcu_common.h"

namespace qcu {

struct Vec {
  void* p;
  long stride;  // bytes between right-hand sides
};
inline Vec vec(void* p, long stride = 0) { return Vec{p, stride}; }

// number of 16-site reduction partials of one half lattice
__host__ __device__ inline long n_partials(long vh) { return (vh + 15) / 16; }

// ------------------------------------------------------------------ hop -----
// block threads = sg * rb; the rb groups share sites (and link loads); a.rpw
// right-hand sides per wave (1, 2 or 4), rpt per thread (1..4)
void launch_hop(int prec, int recon, int epi, const HopArgs& a, int block, int rb, int rpt, hipStream_t st);

// the hop on the interface layouts (complex128 half spinors [Vh][4][3], gauge
// [4][2][Vh][3][3] of which rows 0-1 are read): out = coef * D^s_{p,1-p} in, FP64,
// straight from and to the caller's buffers (k_hop_iface.hip)
void launch_hop_iface(const void* in, void* out, const void* gauge_aos, const Geom& g, int parity, int dagger,
                      double tbc, double coef, hipStream_t st);

// --------------------------------------------------------------- layout -----
// interface spinor (complex128 [Vh][4][3]) <-> internal layout of `prec`
void launch_spinor_from_aos(int prec, const void* aos, long aos_stride, Vec out, long vh, int nrhs, hipStream_t st);
void launch_spinor_to_aos(int prec, Vec in, void* aos, long aos_stride, long vh, int nrhs, hipStream_t st);
// interface gauge (complex128 [4][2][Vh][3][3]) -> packed planes
void launch_gauge_pack(int prec, int recon, const void* aos, void* out, long vh, hipStream_t st);
// rebuild row 2 of every interface-layout link from rows 0 and 1 (in place)
void launch_gauge_reconstruct(void* aos, long vh, hipStream_t st);
// reference storage shifts: kind 0 half spinor [Vh][24] <-> [24][Vh] doubles; kind 1
// gauge [8][Vh][18] -> [8][12][Vh] (back with row 2 rebuilt)
void launch_shift_storage(int kind, int to_soa, const void* src, void* dst, long vh, hipStream_t st);
// lexicographic <-> even-odd, nc (12 or 9) complex per site, nfields fields of V sites
void launch_lex_eo(int to_eo, const void* src, void* dst, const int L[4], int nc, int nfields, hipStream_t st);
// 64-bit fingerprint of `bytes` (a multiple of 8) at buf, order-independent; scratch
// holds kFingerprintBlocks + 1 words and the result lands in scratch[0]
constexpr int kFingerprintBlocks = 1024;
void launch_fingerprint(const void* buf, size_t bytes, uint64_t* scratch, hipStream_t st);
// out = in (accumulate = 0) or out += in (accumulate = 1), any precisions
void launch_spinor_convert(int pin, Vec in, int pout, Vec out, long vh, int nrhs, int accumulate, hipStream_t st);

// ----------------------------------------------------------------- blas -----
// per-right-hand-side CG scalars kept in device memory
struct CgScal {
  double alpha, beta, gamma, tt, ss, target2, ss_ref;
  int first, converged, active, breakdown, iter, pad;
};

void launch_zero(int prec, Vec v, long vh, int nrhs, hipStream_t st);
// dst = src, bytes a multiple of 16 (reference kernel for bandwidth and counters)
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
// partials[r][w] = sum over wave w of |v_r|^2
void launch_norm2(int prec, Vec v, long vh, int nrhs, double* partials, hipStream_t st);
// out[r] = sum_w partials[r][w] in a fixed order (one block per right-hand side)
void launch_sum_partials(const double* partials, long nw, int nrhs, double* out, hipStream_t st);
// y = a x + b y (F64 or same-precision fields)
void launch_axpby(int prec, double a, Vec x, double b, Vec y, long vh, int nrhs, hipStream_t st);
// out = C in with C packed clover of one parity (same precision as the fields)
void launch_clover_apply(int prec, const void* clov, Vec in, Vec out, long vh, int nrhs, hipStream_t st);

// CGLS steps (see solver.cpp).  All skip right-hand sides with active == 0.
void launch_cg_beta(const double* part_rr, long nw, CgScal* sc, int nrhs, hipStream_t st);
void launch_cg_alpha(const double* part_tt, long nw, CgScal* sc, int nrhs, hipStream_t st);
void launch_cg_ss(const double* part_ss, long nw, CgScal* sc, int nrhs, hipStream_t st);
// x += alpha p with the alpha and p of the previous iteration (pending until here), then
// p = r + beta p
void launch_cg_xpay(int prec, int xprec, Vec r, Vec p, Vec x, const CgScal* sc, long vh, int nrhs, hipStream_t st);
// s -= alpha t ; partials of |s|^2
void launch_cg_update(int prec, Vec t, Vec s, const CgScal* sc, double* part_ss, long vh, int nrhs, hipStream_t st);
// x += alpha p for one right-hand side (a pending iterate update, applied by the host)
void launch_cg_xflush(int prec, int xprec, Vec x, Vec p, double alpha, long vh, hipStream_t st);

// --------------------------------------------------------------- clover -----
// G[pl] = Q - Q^dag for the 6 planes at every site of `parity`: double2 [6][9][Vh]
void launch_clover_leaves(const void* gauge_f64_r18, void* G, const Geom& g, int parity, hipStream_t st);
// A = 1 + ck T and A^-1 for one parity, packed F64 (either output may be null);
// nbad counts blocks that needed the pivoted fallback instead of Cholesky
void launch_clover_assemble(const void* G, void* A, void* Ainv, long vh, double ck, int* nbad, hipStream_t st);
// packed F64 -> packed F32 / F16 (one parity)
void launch_clover_convert(int prec, const void* src, void* dst, long vh, hipStream_t st);
// packed F64 -> reference layout complex128 [Vh][2][6][6] (one parity)
void launch_clover_unpack(const void* packed, void* out, long vh, hipStream_t st);

// ------------------------------------------------------------ measurement --
// S site-legs (a multiple of 262144) of the 12-right-hand-side link product Y = U H
// (6x6 by 6x24, FP64) on VALU (mfma = 0) or v_mfma_f64_16x16x4f64 (mfma = 1), operands
// in registers; out receives 262144 partial sums of Y (k_linkprod.hip)
void launch_linkprod(int mfma, long site_legs, double* out, hipStream_t st);

// ------------------------------------------------------------ generators ---
// U_mu(x) = SU3-project(exp(i eps H)) into the interface gauge layout
void launch_gauge_random(void* aos, long vh, uint64_t seed, double eps, hipStream_t st);
// N(0,1) per real of n complex128 values, counter = element index
void launch_source_gauss(void* aos, long n_complex, uint64_t seed, hipStream_t st);

}  // namespace qcu
