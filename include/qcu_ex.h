/*
 * qcu_ex.h -- extended C API of the MI355X qcu library (libqcu.so).
 *
 * The reference has only the five void entry points of qcu.h; everything here is
 * new and has no counterpart in the reference.  It adds resident gauge handles,
 * kappa / c_sw / t boundary, mixed precision, batched right-hand sides, solve
 * statistics, int error codes and the device-memory helpers that a caller without its
 * own allocator (tests, benchmarks, ctypes) needs.  Every function returns
 * QCU_SUCCESS (0) or an error code, and never exits the process.
 *
 * Field pointers are HIP device pointers in the qcu.h interface layouts
 * (complex128): gauge [4][2][Vh][3][3], half spinor [Vh][4][3], full spinor
 * [2][Vh][4][3], clover [2][Vh][2][6][6].
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "qcu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  QCU_SUCCESS = 0,
  QCU_ERROR_INVALID_VALUE = 1,
  QCU_ERROR_HIP = 2,
  QCU_ERROR_OUT_OF_MEMORY = 3,
  QCU_ERROR_NOT_CONVERGED = 4,
  QCU_ERROR_BREAKDOWN = 5,
  QCU_ERROR_NOT_FINITE = 6,
  QCU_ERROR_STATE = 7,
  QCU_ERROR_ADMISSION = 8
};

enum { QCU_PREC_DOUBLE = 0, QCU_PREC_SINGLE = 1, QCU_PREC_HALF = 2 };

typedef struct qcuGauge_st* qcuGauge;

typedef struct {
  double kappa;    /* hopping parameter */
  double csw;      /* clover coefficient; 0 = Wilson (A = 1) */
  int t_boundary;  /* +1 periodic, -1 antiperiodic in t */
  int recon;       /* link storage: 12 (two rows, third rebuilt) or 18 */
} qcuOperator;

typedef struct {
  double tol;      /* stop when ||b - M x|| / ||b|| <= tol (FP64, full lattice) */
  int max_iter;    /* iteration cap (<= 0: 10 * volume) */
  int precision;   /* inner precision: QCU_PREC_DOUBLE = plain FP64 CG */
  double delta;    /* reliable update when |r| < delta * |r| at the last one; outside
                      (0, 1), as by default: 1e-3 for an FP32 inner loop, 0.1 for FP16 */
  int graph;       /* 1: replay captured hipGraphs of `chunk` iterations */
  int chunk;       /* iterations between host convergence checks */
} qcuSolveParams;

typedef struct {
  int iterations;        /* CG iterations (max over right-hand sides in a batch) */
  int reliable_updates;  /* FP64 residual recomputations */
  int converged;         /* 1 if every right-hand side reached tol */
  int breakdown;         /* 1 if CG met p^dag A p <= 0 or a non-finite value */
  double true_residual;  /* max over right-hand sides of ||b - M x|| / ||b||, FP64 */
  double seconds;        /* wall time of the solve */
} qcuSolveStats;

/* ---- library --------------------------------------------------------------- */
int qcuInit(int device);
int qcuFinalize(void);
int qcuGetLastError(void);
const char* qcuGetLastErrorString(void);
void qcuClearLastError(void);
const char* qcuVersion(void);

/* ---- device memory helpers ------------------------------------------------- */
int qcuMalloc(void** ptr, size_t bytes);
int qcuFree(void* ptr);
int qcuMemcpyHtoD(void* dst, const void* src, size_t bytes);
int qcuMemcpyDtoH(void* dst, const void* src, size_t bytes);
int qcuMemcpyDtoD(void* dst, const void* src, size_t bytes);
int qcuMemset(void* dst, int value, size_t bytes);
int qcuSynchronize(void);
int qcuMemInfo(size_t* free_bytes, size_t* total_bytes);

/* ---- resident gauge fields ------------------------------------------------- */
/* copy and pack an interface-layout gauge (device pointer) */
int qcuGaugeCreate(qcuGauge* g, const void* gauge, const int lattice_size[4]);
/* seeded near-unit SU(3) gauge: U = SU3-project(exp(i eps H)) per link */
int qcuGaugeCreateRandom(qcuGauge* g, const int lattice_size[4], uint64_t seed, double eps);
int qcuGaugeDestroy(qcuGauge g);
/* interface-layout copy of the gauge into a device buffer */
int qcuGaugeExport(qcuGauge g, void* gauge);
/* A = 1 + kappa*csw*T and A^-1, complex128 [2][Vh][2][6][6] (either may be NULL);
 * *nfallback (may be NULL) = blocks inverted by the pivoted fallback */
int qcuCloverExport(qcuGauge g, const qcuOperator* op, void* A, void* Ainv, int* nfallback);

/* ---- sources ---------------------------------------------------------------- */
/* N(0,1) per real of a full spinor of `lattice_size`, element k from counter k */
int qcuSourceGaussian(void* v, const int lattice_size[4], uint64_t seed);
/* unit point source: 1 at full coordinates `coords` (x,y,z,t), spin, colour; 0 elsewhere */
int qcuSourcePoint(void* v, const int lattice_size[4], const int coords[4], int spin, int colour);

/* ---- layout converters ------------------------------------------------------ */
/* Lexicographic <-> even-odd site order, complex128, for a caller whose fields are
 * lexicographic (the reference takes even-odd only; its outside caller converts):
 *   QCU_FIELD_SPINOR  [V][4][3]     <-> full spinor [2][Vh][4][3]
 *   QCU_FIELD_GAUGE   [4][V][3][3]  <-> gauge [4][2][Vh][3][3]
 * Lexicographic site l = ((t*Lz+z)*Ly+y)*Lx+x goes to parity (x+y+z+t)%2, half-site
 * index l/2 (include/qcu_point.cuh:20-93).  dst and src must not overlap.  The
 * reference's AoS <-> coalesced SoA shifts are in qcu_shift_storage.h. */
enum { QCU_FIELD_SPINOR = 0, QCU_FIELD_GAUGE = 1 };
int qcuLexToEvenOdd(void* eo, const void* lex, const int lattice_size[4], int field);
int qcuEvenOddToLex(void* lex, const void* eo, const int lattice_size[4], int field);

/* ---- operators (interface layouts) --------------------------------------- */
/* out(parity) = D^s in(1-parity), s = -1 if dagger; precision selects the kernel
 * (inputs are rounded to it); recon 12 or 18 */
int qcuHop(qcuGauge g, void* out, const void* in, int parity, int dagger, int t_boundary, int recon,
           int precision);
/* the same for nrhs half spinors in one launch (each link loaded once per site for
 * all of them); out[r] and in[r] are separate device buffers */
int qcuHopBatch(qcuGauge g, int nrhs, void* const* out, const void* const* in, int parity, int dagger,
                int t_boundary, int recon, int precision);
/* qcuHop without device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and result as qcuHop, bit for
 * bit, with the two internal staging fields in a static device array of 32 MiB that
 * each device the library runs on holds until the process ends: a half volume whose two
 * fields of `precision` exceed it (FP64 above 87,360 half sites; 16^3x32 fits) is
 * rejected with QCU_ERROR_INVALID_VALUE.  It allocates nothing itself, so the gauge's
 * packed links of this precision and recon must exist already (one qcuHop, qcuHopBatch
 * or solve with them makes them), else QCU_ERROR_STATE.  Calls that overlap on one
 * device share the array and race. */
int qcuHopNoAlloc(qcuGauge g, void* out, const void* in, int parity, int dagger, int t_boundary, int recon,
                  int precision);
/* full lattice: out = (A - kappa D_s) in */
int qcuApplyM(qcuGauge g, const qcuOperator* op, void* out, const void* in, int dagger);
/* even half spinors: out = Mhat in = A_ee in - kappa^2 D_eo A_oo^-1 D_oe in (s = +1 or
 * -1 if dagger), computed in `precision` */
int qcuApplyMhat(qcuGauge g, const qcuOperator* op, void* out, const void* in, int dagger, int precision);

/* Full-lattice compositions of the reference's clover operator path (reference
 * src/qcu_clover_dslash.cu), for a caller that builds its own solver from them.  With
 * M = A - kappa D (A = 1 for c_sw = 0) and s = -1 if dagger:
 *   QCU_COMP_AINV_D            out = kappa A^-1 D_s in      cloverDslashOneRound :990-1011
 *   QCU_COMP_ONE_MINUS_AINV_D  out = in - kappa A^-1 D_s in newFullCloverDslashOneRound :1013-1042
 *   QCU_COMP_AINV              out = A^-1 in                invertCloverDslash :902-918
 *   QCU_COMP_A                 out = A in                   preCloverDslash :927-946
 *   QCU_COMP_NORMAL            out = M' M in, M = 1 - kappa A^-1 D_+, M' = 1 - kappa A^-1 D_-
 *                                                           MmV_one_round :1044-1049
 * The reference's invertCloverDslash and preCloverDslash apply the clover in place on
 * fermion_out; here they read `in` and write `out` (pass the same pointer for in place).
 * A = 1 + kappa c_sw T; the reference's normalisation is kappa = c_sw = 1.  FP64. */
enum {
  QCU_COMP_AINV_D = 0,
  QCU_COMP_ONE_MINUS_AINV_D = 1,
  QCU_COMP_AINV = 2,
  QCU_COMP_A = 3,
  QCU_COMP_NORMAL = 4
};
int qcuApplyComposite(qcuGauge g, const qcuOperator* op, int kind, void* out, const void* in, int dagger);

/* ---- solver ----------------------------------------------------------------- */
void qcuDefaultSolveParams(qcuSolveParams* sp);
/* parameters cg_inverter uses */
int qcuSetDefaultSolve(const qcuSolveParams* sp);
/* x <- M^-1 b for one full spinor */
int qcuSolve(qcuGauge g, const qcuOperator* op, const qcuSolveParams* sp, void* x, const void* b,
             qcuSolveStats* stats);
/* nrhs full spinors solved together (links loaded once per site for all of them);
 * stats may be an array of nrhs entries or NULL */
int qcuSolveBatch(qcuGauge g, const qcuOperator* op, const qcuSolveParams* sp, int nrhs, void* const* x,
                  const void* const* b, qcuSolveStats* stats);

/* ---- measurement ------------------------------------------------------------ */
/* Times `iters` back-to-back even-odd hop applications (alternating parities) on
 * internal-layout fields of `precision`, nrhs right-hand sides per launch, with
 * hipEvents on the library stream.  ms = average per launch. */
int qcuBenchHop(qcuGauge g, int precision, int recon, int nrhs, int iters, double* ms);
/* device copy bandwidth: ms per copy of `bytes` (read + write) */
int qcuBenchCopy(size_t bytes, int iters, double* ms);
/* The link product of a 12-right-hand-side FP64 hop leg, Y = U H (a site's link as a
 * 6x6 real matrix times its 24 projected colour vectors as 6x24), over site_legs
 * site-legs (a multiple of 262144; a 32^3x64 launch has 8 per site, 8388608) with
 * operands in registers: mfma = 0 on VALU FMAs, 1 on v_mfma_f64_16x16x4f64.
 * ms = average per run; checksum = sum of every entry of Y (the same products both ways). */
int qcuBenchLinkProduct(int mfma, long site_legs, int iters, double* ms, double* checksum);

/* ---- serving ------------------------------------------------------------------ */
/* A request loop with continuous batching.  A worker thread keeps one batch of
 * right-hand-side slots per (gauge, Wilson or clover) on its own stream; requests
 * that arrive while others run are admitted as slots and the HBM budget free up,
 * right-hand sides of different requests (and different kappa) on one gauge share
 * each link load, and batches on different gauges overlap on the device.  A
 * right-hand side gives the same bits alone or in any mix. */
typedef struct qcuServer_st* qcuServer;

typedef struct {
  size_t hbm_budget;  /* device bytes for batches and clover fields (0: 3/4 of free memory) */
  int max_batch;      /* right-hand-side slots per batch, 1..16 */
  int precision;      /* inner precision of every solve (QCU_PREC_*) */
  int chunk;          /* CG iterations per scheduling round */
  int graph;          /* replay captured hipGraphs */
} qcuServerConfig;

typedef struct {
  qcuGauge gauge;          /* resident gauge (must outlive the request) */
  double kappa, csw, tol;  /* csw = 0: Wilson */
  int t_boundary, recon;
  int nrhs;                /* 1..16 */
  void* const* x;          /* nrhs solution buffers (full spinors) or NULL: residuals only */
  const void* const* b;    /* nrhs sources (read again when each right-hand side finishes,
                              so unchanged until the request is done) or NULL: Gaussian
                              sources, seeds seed + r */
  uint64_t seed;
  int max_iter;            /* <= 0: 10 * volume */
} qcuRequest;

typedef struct {
  int done;                /* every right-hand side finished */
  int converged;           /* every right-hand side reached tol */
  int iterations;          /* max over right-hand sides */
  double true_residual;    /* max over right-hand sides, FP64 */
  double queue_seconds;    /* submit -> first right-hand side admitted */
  double latency_seconds;  /* submit -> last right-hand side finished */
} qcuRequestStatus;

void qcuDefaultServerConfig(qcuServerConfig* cfg);
int qcuServerCreate(qcuServer* s, const qcuServerConfig* cfg);
/* non-blocking; *id identifies the request */
int qcuServerSubmit(qcuServer s, const qcuRequest* req, uint64_t* id);
int qcuServerPoll(qcuServer s, uint64_t id, qcuRequestStatus* st);
/* blocks until request id is done */
int qcuServerWait(qcuServer s, uint64_t id, qcuRequestStatus* st);
/* blocks until every submitted request is done */
int qcuServerDrain(qcuServer s);
/* device bytes currently held, peak, and batches in flight */
int qcuServerUsage(qcuServer s, size_t* bytes, size_t* peak_bytes, int* batches);
int qcuServerDestroy(qcuServer s);

/* ---- tuning ------------------------------------------------------------------ */
/* Times every launch shape of the hopping kernel for (precision, recon, nrhs) on this
 * gauge's volume and keeps the fastest for later launches; the eight fastest are
 * re-timed with each epilogue that reads x or a clover field, and solves use the
 * winner per epilogue.  shape (may be NULL; the plain epilogue's winner)
 * receives 12 ints: threads per block, wave groups per block sharing sites, right-hand
 * sides per thread, XCD-contiguous block order (0/1), the z and t extents of the
 * site-walk tile (0 = storage order), right-hand side groups as the slowest block
 * index (1) or the fastest (0), the y extent of the walk tile (0 = all of y),
 * right-hand sides per wave (1, 2 or 4; a wave then covers 64, 32 or 16 sites),
 * site blocks walked by each block (sites per thread, 1..64), the load policy of
 * links, clover and x (0 streaming, 1 cached), and the kernel (0: the above; 1: every
 * right-hand side in one block of 1, 2 or 4 waves per 64-site chunk, whose links are
 * loaded once into LDS -- wave groups, right-hand sides per wave and per thread and
 * group order then do not apply; 2: as 1, and a clover field shared by every
 * right-hand side is staged in LDS too).
 * A tile that does not divide the lattice falls back to storage order.  QCU_TUNE=1 makes solves tune on first use;
 * QCU_TUNE_CACHE=<file> persists results across processes. */
int qcuTuneHop(qcuGauge g, int precision, int recon, int nrhs, int* shape, double* ms);
/* Sets the launch shape (the 12 ints above, for every epilogue) used for (precision, recon, nrhs) on this
 * gauge's volume, e.g. one tuned earlier.  Every valid shape gives the same bits. */
int qcuSetHopShape(qcuGauge g, int precision, int recon, int nrhs, const int* shape);
/* The launch shape (the 12 ints above) the plain hop uses for (precision, recon, nrhs) on this gauge's volume:
 * one tuned, pinned or read from QCU_TUNE_CACHE (whose lines with a shape qcuSetHopShape would refuse are
 * skipped), else the default.  A tile extent of -1 stands for the whole lattice in that direction. */
int qcuGetHopShape(qcuGauge g, int precision, int recon, int nrhs, int* shape);

#ifdef __cplusplus
}
#endif
