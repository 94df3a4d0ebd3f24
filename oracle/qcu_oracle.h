/*
 * qcu_oracle.h -- CPU FP64 restatement of the qcu even-odd Wilson / Wilson-clover
 * arithmetic (reference: YellowBicycleee/Qcu_Cg_Inverter, cg_inverter/).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP library
 * (libqcu.so): only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  The product path never links or calls it.
 *
 * Field layouts are the reference's interface formats (complex128 = two doubles):
 *   gauge        [mu 4][parity 2][Vh][row 3][col 3]        include/qcu_point.cuh:65-67
 *   half spinor  [Vh][spin 4][colour 3]                    include/qcu_point.cuh:70-72
 *   full spinor  [parity 2][Vh][spin 4][colour 3]          src/qcu_clover_dslash.cu:1002-1003
 *   clover       [parity 2][Vh][block 2][6][6]              include/qcu_point.cuh:91-93
 * Half-site index: ((t*Lz + z)*Ly + y)*(Lx/2) + xh, full x = 2*xh + ((parity+y+z+t)&1).
 *
 * Pinning status per function (tests/test_oracle_*.py):
 *   orc_neighbor            brute-force lexicographic coordinates
 *   orc_hop                 dense textbook matrix (numpy), free-field plane wave,
 *                           gauge covariance, s=+1/s=-1 adjointness
 *   orc_clover_T            U=1 => T=0, hermiticity, chiral blocks, gauge covariance,
 *                           textbook sigma.F build (numpy) on 2^4 and 4^4, single-link
 *                           perturbation support
 *   orc_invert6_gj/_chol    A*A^-1 = 1, numpy.linalg.inv
 *   orc_full_op/orc_schur_* dense (A - kappa D) from the textbook hop; Schur identity
 *   orc_cg_eo / orc_cg_ref  numpy.linalg.solve of the dense operator on 4^4
 *   orc_gauge_random etc.   unitarity / det = 1, seeded determinism, moments
 *   orc_shift_*, orc_lex_eo coordinate-coded known answers (numpy lexicographic grid),
 *                           round trips, SU(3) row-2 closed form
 */
#ifndef QCU_ORACLE_H
#define QCU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry ------------------------------------------------------------ */
/* coordinates of half-site idx of the given parity */
void orc_coords(const int L[4], int parity, long idx, int c[4]);
/* half-site index and parity of full coordinates c (taken modulo L) */
long orc_index(const int L[4], const int c[4], int* parity);
/* neighbour of (parity, idx) one step along mu (0..3), forward if fwd else backward.
 * Restates Point::move, include/qcu_point.cuh:20-63.  Returns the neighbour's
 * half-site index; its parity is 1-parity.  *wrapped = 1 if the step crossed the
 * lattice edge in direction mu. */
long orc_neighbor(const int L[4], int parity, long idx, int mu, int fwd, int* wrapped);

/* ---- layouts (complex128 fields as doubles) ------------------------------ */
/* one half spinor [Vh][24] <-> the reference's coalesced [24][Vh]; dir 0 = to it
 * (src/qcu_shift_storage.cu:74-116) */
void orc_shift_vector(double* dst, const double* src, int dir, long vh);
/* gauge [8][Vh][18] -> coalesced rows 0-1 [8][12][Vh] (dir 0,
 * src/qcu_shift_storage.cu:123-161); dir 1 back, rebuilding row 2 */
void orc_shift_gauge(double* dst, const double* src, int dir, long vh);
/* nf fields of nc complex per site: lexicographic [nf][V][nc] <-> even-odd
 * [nf][2][Vh][nc] (to_eo selects) */
void orc_lex_eo(double* dst, const double* src, const int L[4], int nc, int nf, int to_eo);

/* ---- SU(3) ---------------------------------------------------------------- */
/* row2 = conj(row0 x row1), src/qcu_wilson_dslash_neo.cu:73-78 */
void orc_reconstruct12(double* u /* 9 complex */);

/* ---- hopping term --------------------------------------------------------- */
/* out(parity) = D^s_{p,1-p} in(1-parity):
 *   sum_mu U_mu(x)(1 - s g_mu) in(x+mu) + U_mu(x-mu)^dag (1 + s g_mu) in(x-mu)
 * src/qcu_wilson_dslash_neo.cu:941-1206 (s = +1 is flag 0, s = -1 is dagger,
 * src/qcu_clover_dslash.cu:570-772).  tbc multiplies every hop that crosses the
 * t edge (+1 periodic, -1 antiperiodic).  recon = 12 rebuilds row 2 of each link
 * as the reference does; recon = 18 uses all nine stored entries. */
void orc_hop(double* out, const double* in, const double* gauge, const int L[4],
             int parity, int s, double tbc, int recon);

/* ---- clover ----------------------------------------------------------------*/
/* T(x) = -(1/8) sum_{mu<nu} (g_mu g_nu) (x) (Q_munu - Q_munu^dag) for both parities,
 * reference layout [2][Vh][72 complex].  src/qcu_clover_dslash.cu:329-458,210-267 */
void orc_clover_T(double* T, const double* gauge, const int L[4], int recon);
/* A = 1 + ck*T (ck = kappa*c_sw; the reference is ck = 1), src/qcu_clover_dslash.cu:536-540 */
void orc_clover_A(double* A, const double* T, long nsites, double ck);
/* 6x6 complex Gauss-Jordan inverse as in src/qcu_clover_dslash.cu:42-100 */
void orc_invert6_gj(const double* m, double* inv);
/* 6x6 Hermitian positive-definite inverse by Cholesky; returns 0 on success,
 * -1 if a pivot is not positive */
int orc_invert6_chol(const double* m, double* inv);
/* per-site inverse of both chiral blocks of every site (Gauss-Jordan if chol==0) */
int orc_clover_invert(double* Ainv, const double* A, long nsites, int chol);
/* out = blockdiag(M0, M1) in for every site of one half lattice,
 * src/qcu_clover_dslash.cu:461-506 */
void orc_clover_apply(double* out, const double* in, const double* M, long vh);

/* ---- operators ------------------------------------------------------------ */
/* full lattice: out = A in - kappa D^s in (A = NULL means A = 1, Wilson). */
void orc_full_op(double* out, const double* in, const double* gauge, const double* A,
                 const int L[4], double kappa, int s, double tbc, int recon);
/* even-odd Schur operator on even sites:
 *   out_e = A_ee in_e - kappa^2 D^s_eo A_oo^-1 D^s_oe in_e
 * Aee / Ainv_oo are half-lattice clover arrays (NULL = identity). */
void orc_schur_op(double* out, const double* in, const double* gauge,
                  const double* Aee, const double* Ainv_oo, const int L[4],
                  double kappa, int s, double tbc, int recon);

typedef struct {
  int iterations;
  double true_residual;   /* ||b - M x|| / ||b|| of the full system, FP64 */
  double normal_residual; /* last |r|/|b^| of the normal equations */
  int converged;
} OrcSolveStats;

/* Even-odd CG on Mhat^dag Mhat, FP64, true adjoint.  b, x full spinors.
 * A / Ainv full-lattice clover arrays (NULL => Wilson). */
int orc_cg_eo(double* x, const double* b, const double* gauge, const double* A,
              const double* Ainv, const int L[4], double kappa, double tbc, int recon,
              double tol, int max_iter, OrcSolveStats* st);

/* The reference's full-lattice CG, src/cg_inverter.cu:181-369 with its operator
 * compositions src/qcu_clover_dslash.cu:902-1049: solves M'M x = M'A^-1 b with
 * M = 1 - A^-1 D and M' = 1 - A^-1 D^dag (kappa = 1, periodic), stopping when
 * |<r_old, r_old>| < stop_rr.  Exact CG only when A = 1 (Wilson). */
int orc_cg_ref(double* x, const double* b, const double* gauge, const double* Ainv,
               const int L[4], double stop_rr, int max_iter, int* iters);

/* ---- seeded inputs (counter-based; the HIP library implements the same) -- */
uint64_t orc_rng_u64(uint64_t seed, uint64_t ctr);
double orc_rng_gauss(uint64_t seed, uint64_t ctr);
/* U_mu(x) = SU3-project(exp(i eps H)), H = sum_a h_a lambda_a / 2, h_a ~ N(0,1) */
void orc_gauge_random(double* gauge, const int L[4], uint64_t seed, double eps);
void orc_source_gauss(double* v, long nsites, uint64_t seed);
void orc_source_point(double* v, const int L[4], const int c[4], int spin, int colour);

/* ---- BLAS helpers --------------------------------------------------------- */
double orc_norm2(const double* v, long n_complex);
void orc_cdot(const double* x, const double* y, long n_complex, double* re, double* im);

void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
