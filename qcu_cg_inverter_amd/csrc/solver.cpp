// solver.cpp -- even-odd preconditioned CG on the normal equations, FP64 or mixed
// precision with FP64 reliable updates, for a batch of right-hand sides.
//
// Replaces the reference's cg_inverter (src/cg_inverter.cu:260-369), which runs CG on
// the full lattice through M' M with M' = 1 - A^-1 D^dag (not the adjoint of
// M = 1 - A^-1 D, src/qcu_clover_dslash.cu:1044-1049) and a host round trip per
// reduction.  Here, with M = A - kappa D and even sites "e" = parity 0:
//   Mhat   = A_ee - kappa^2 D_eo A_oo^-1 D_oe,   Mhat^dag uses s = -1 in both hops
//   b^     = b_e + kappa D_eo A_oo^-1 b_o
//   x_o    = A_oo^-1 (b_o + kappa D_oe x_e)
// and CGLS (CG on Mhat^dag Mhat x = Mhat^dag b^ that carries s = b^ - Mhat x):
//   r = Mhat^dag s ; beta = |r|^2/gamma ; gamma = |r|^2 ; p = r + beta p
//   t = Mhat p ; alpha = gamma/|t|^2 ; x += alpha p ; s -= alpha t
// so the stopping test |s| <= tol |b| is on the residual of the system itself.  The
// step x += alpha p is carried out by the next iteration's p update, which reads p
// anyway (launch order: r, beta, [x += alpha p, p = r + beta p], t, alpha, s).
// Mixed precision: the loop runs in FP32 or FP16 (+ per-site norm); when |s| drops
// below delta times its value at the last update (or reaches the target) the FP32
// iterate is added to the FP64 solution and s is recomputed in FP64 (reliable
// updates, Clark et al., Comput. Phys. Commun. 181 (2010) 1517).
// Scalars stay on the device; the host looks every `chunk` iterations, and the
// chunk can be a captured hipGraph.  The slot machinery is in cg_batch.cpp.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cg_batch.h"
#include "qcu_host.h"

namespace qcu {

struct OpView {
  GaugeField* g;
  qcuOperator op;
  bool clover;
  CloverField* cf;
  double tbc() const { return (double)op.t_boundary; }
};

static void check_op(const qcuOperator& op) {
  QREQ(std::isfinite(op.kappa) && std::isfinite(op.csw), "kappa and csw must be finite");
  QREQ(op.recon == 12 || op.recon == 18, "recon must be 12 or 18");
  QREQ(op.t_boundary == 1 || op.t_boundary == -1, "t_boundary must be +1 or -1");
}

static OpView make_view(GaugeField& g, const qcuOperator& op, int prec) {
  check_op(op);
  OpView o;
  o.g = &g;
  o.op = op;
  o.clover = op.csw != 0.0;
  o.cf = nullptr;
  if (o.clover) o.cf = g.clover(op.kappa * op.csw, prec, true);
  g.get(prec, op.recon);
  g.get(F64, op.recon);
  return o;
}

// out = Mhat in (Mhat^dag if dag) on even fields; tmp holds the odd intermediate
static void mhat(OpView& o, int prec, int dag, Vec in, Vec out, Vec tmp, hipStream_t st) {
  const double k2 = o.op.kappa * o.op.kappa;
  if (o.clover) {
    hop(*o.g, prec, o.op.recon, EPI_CLOV, 1, dag, o.tbc(), in, tmp, vec(nullptr), o.cf->ainv(prec, 1), 1.0, 1,
        nullptr, st);
    hop(*o.g, prec, o.op.recon, EPI_CLOV_XPAY, 0, dag, o.tbc(), tmp, out, in, o.cf->a(prec, 0), -k2, 1, nullptr, st);
  } else {
    hop(*o.g, prec, o.op.recon, EPI_PLAIN, 1, dag, o.tbc(), in, tmp, vec(nullptr), nullptr, 1.0, 1, nullptr, st);
    hop(*o.g, prec, o.op.recon, EPI_XPAY, 0, dag, o.tbc(), tmp, out, in, nullptr, -k2, 1, nullptr, st);
  }
}

// out_p = A_pp in_p - kappa D^s in_(1-p) for one parity p (FP64)
static void full_op_parity(OpView& o, int p, int dag, Vec in_p, Vec in_q, Vec out, hipStream_t st) {
  if (o.clover)
    hop(*o.g, F64, o.op.recon, EPI_CLOV_XPAY, p, dag, o.tbc(), in_q, out, in_p, o.cf->a(F64, p), -o.op.kappa, 1,
        nullptr, st);
  else
    hop(*o.g, F64, o.op.recon, EPI_XPAY, p, dag, o.tbc(), in_q, out, in_p, nullptr, -o.op.kappa, 1, nullptr, st);
}

// A batch object kept between calls so repeated solves of the same shape reuse
// their device buffers and captured graphs' allocations.
struct CachedBatch {
  GaugeField* g = nullptr;
  int prec = -1, recon = 0, tbc = 0, clover = -1, cap = 0, chunk = 0, graph = -1;
  double delta = 0.0;
  std::unique_ptr<CgBatch> batch;
};
static CachedBatch g_cached;

void drop_cached_batches(GaugeField* g) {
  if (!g || g_cached.g == g) {
    g_cached.batch.reset();
    g_cached.g = nullptr;
  }
}

static CgBatch& batch_for(GaugeField& g, int prec, const qcuOperator& op, int nrhs, const qcuSolveParams& sp) {
  CachedBatch& c = g_cached;
  const int clov = op.csw != 0.0;
  if (!c.batch || c.g != &g || c.prec != prec || c.recon != op.recon || c.tbc != op.t_boundary || c.clover != clov ||
      c.cap != nrhs || c.chunk != sp.chunk || c.graph != sp.graph || c.delta != sp.delta) {
    c.batch.reset();
    c.batch.reset(new CgBatch(g, prec, op.recon, op.t_boundary, clov != 0, nrhs, sp, stream()));
    c.g = &g;
    c.prec = prec;
    c.recon = op.recon;
    c.tbc = op.t_boundary;
    c.clover = clov;
    c.cap = nrhs;
    c.chunk = sp.chunk;
    c.graph = sp.graph;
    c.delta = sp.delta;
  }
  return *c.batch;
}

void solve_batch(GaugeField& g, const qcuOperator& op, const qcuSolveParams& spin, int nrhs, void* const* x,
                 const void* const* b, qcuSolveStats* stats) {
  qcuSolveParams sp = spin;
  QREQ(sp.tol > 0.0 && std::isfinite(sp.tol), "tol must be positive");
  QREQ(sp.precision >= 0 && sp.precision <= 2, "bad precision");
  if (sp.chunk < 1) sp.chunk = 1;
  sp.delta = reliable_delta(sp.precision, sp.delta);
  for (int r = 0; r < nrhs; ++r) QREQ(x[r] && b[r], "NULL spinor");
  check_op(op);
  const int P = sp.precision;
  maybe_autotune(g, P, op.recon, nrhs);  // QCU_TUNE=1: before any graph is captured
  CgBatch& batch = batch_for(g, P, op, nrhs, sp);
  batch.take_finished();
  std::vector<SlotResult> res;
  try {
    std::vector<SlotSpec> specs(nrhs);
    for (int r = 0; r < nrhs; ++r) {
      SlotSpec& s = specs[r];
      s.b = b[r];
      s.seed = 0;
      s.x = x[r];
      s.kappa = op.kappa;
      s.csw = op.csw;
      s.tol = sp.tol;
      s.max_iter = sp.max_iter;
      s.tag = (uint64_t)r;
    }
    batch.add_slots(specs.data(), nrhs);
    while (batch.active() > 0) {
      batch.launch_chunk();
      batch.harvest();
    }
    res = batch.take_finished();
  } catch (...) {
    // slots added before the failure would stay active in the cached batch and make
    // every later solve of this shape fail with "batch is full": drop the batch
    drop_cached_batches(&g);
    throw;
  }
  std::sort(res.begin(), res.end(), [](const SlotResult& a, const SlotResult& c) { return a.tag < c.tag; });
  bool all_conv = true, breakdown = false;
  qcuSolveStats agg;
  memset(&agg, 0, sizeof agg);
  for (const SlotResult& r : res) {
    all_conv &= r.stats.converged != 0;
    breakdown |= r.stats.breakdown != 0;
    agg.iterations = std::max(agg.iterations, r.stats.iterations);
    agg.reliable_updates += r.stats.reliable_updates;
    agg.true_residual = std::max(agg.true_residual, r.stats.true_residual);
    agg.seconds = std::max(agg.seconds, r.stats.seconds);
    if (stats && nrhs > 1) stats[r.tag] = r.stats;
  }
  agg.converged = all_conv;
  agg.breakdown = breakdown;
  if (stats && nrhs == 1) *stats = agg;
  if (breakdown) throw Error(QCU_ERROR_BREAKDOWN, "CG breakdown (p^dag A p <= 0 or non-finite value)");
  if (!all_conv) throw Error(QCU_ERROR_NOT_CONVERGED, "CG did not reach the tolerance within max_iter");
}

void apply_mhat_iface(GaugeField& g, const qcuOperator& op, void* out, const void* in, int dagger, int prec) {
  OpView o = make_view(g, op, prec);
  long vh = g.geo.vh;
  Vec a = vec(ctx().buf("mh_a", spinor_bytes(prec, vh))->p);
  Vec b = vec(ctx().buf("mh_b", spinor_bytes(prec, vh))->p);
  Vec t = vec(ctx().buf("mh_t", spinor_bytes(prec, vh))->p);
  launch_spinor_from_aos(prec, in, 0, a, vh, 1, stream());
  mhat(o, prec, dagger, a, b, t, stream());
  launch_spinor_to_aos(prec, b, out, 0, vh, 1, stream());
  QHIP(hipGetLastError());
}

// out_p = kappa A_pp^-1 D^s in_(1-p) for both parities (A = 1 for Wilson)
static void ainv_d(OpView& o, int dag, const Vec in[2], const Vec out[2], hipStream_t st) {
  for (int p = 0; p < 2; ++p) {
    if (o.clover)
      hop(*o.g, F64, o.op.recon, EPI_CLOV, p, dag, o.tbc(), in[1 - p], out[p], vec(nullptr), o.cf->ainv(F64, p),
          o.op.kappa, 1, nullptr, st);
    else
      hop(*o.g, F64, o.op.recon, EPI_PLAIN, p, dag, o.tbc(), in[1 - p], out[p], vec(nullptr), nullptr, o.op.kappa, 1,
          nullptr, st);
  }
}

// out = in - kappa A^-1 D^s in (both parities; out must not alias in)
static void one_minus_ainv_d(OpView& o, int dag, const Vec in[2], const Vec out[2], hipStream_t st) {
  ainv_d(o, dag, in, out, st);
  for (int p = 0; p < 2; ++p) launch_axpby(F64, 1.0, in[p], -1.0, out[p], o.g->geo.vh, 1, st);
}

void apply_composite_iface(GaugeField& g, const qcuOperator& op, int kind, void* out, const void* in, int dagger) {
  QREQ(kind >= QCU_COMP_AINV_D && kind <= QCU_COMP_NORMAL, "qcuApplyComposite: unknown kind");
  OpView o = make_view(g, op, F64);
  const long vh = g.geo.vh;
  const size_t fb = spinor_bytes(F64, vh);
  Vec vi[2] = {vec(ctx().buf("cp_i0", fb)->p), vec(ctx().buf("cp_i1", fb)->p)};
  Vec vo[2] = {vec(ctx().buf("cp_o0", fb)->p), vec(ctx().buf("cp_o1", fb)->p)};
  hipStream_t st = stream();
  for (int p = 0; p < 2; ++p)
    launch_spinor_from_aos(F64, static_cast<const char*>(in) + p * vh * 192, 0, vi[p], vh, 1, st);
  const int dag = dagger ? 1 : 0;
  switch (kind) {
    case QCU_COMP_AINV_D: ainv_d(o, dag, vi, vo, st); break;
    case QCU_COMP_ONE_MINUS_AINV_D: one_minus_ainv_d(o, dag, vi, vo, st); break;
    case QCU_COMP_AINV:
    case QCU_COMP_A:
      for (int p = 0; p < 2; ++p) {
        if (o.clover)
          launch_clover_apply(F64, kind == QCU_COMP_AINV ? o.cf->ainv(F64, p) : o.cf->a(F64, p), vi[p], vo[p], vh, 1,
                              st);
        else
          QHIP(hipMemcpyAsync(vo[p].p, vi[p].p, fb, hipMemcpyDeviceToDevice, st));
      }
      break;
    default: {  // QCU_COMP_NORMAL: M' M in
      Vec vt[2] = {vec(ctx().buf("cp_t0", fb)->p), vec(ctx().buf("cp_t1", fb)->p)};
      one_minus_ainv_d(o, 0, vi, vt, st);
      one_minus_ainv_d(o, 1, vt, vo, st);
    }
  }
  for (int p = 0; p < 2; ++p)
    launch_spinor_to_aos(F64, vo[p], static_cast<char*>(out) + p * vh * 192, 0, vh, 1, st);
  QHIP(hipGetLastError());
}

void apply_m_iface(GaugeField& g, const qcuOperator& op, void* out, const void* in, int dagger) {
  OpView o = make_view(g, op, F64);
  long vh = g.geo.vh;
  Vec ie = vec(ctx().buf("m_ie", spinor_bytes(F64, vh))->p);
  Vec io = vec(ctx().buf("m_io", spinor_bytes(F64, vh))->p);
  Vec oe = vec(ctx().buf("m_oe", spinor_bytes(F64, vh))->p);
  Vec oo = vec(ctx().buf("m_oo", spinor_bytes(F64, vh))->p);
  launch_spinor_from_aos(F64, in, 0, ie, vh, 1, stream());
  launch_spinor_from_aos(F64, static_cast<const char*>(in) + vh * 192, 0, io, vh, 1, stream());
  full_op_parity(o, 0, dagger, ie, io, oe, stream());
  full_op_parity(o, 1, dagger, io, ie, oo, stream());
  launch_spinor_to_aos(F64, oe, out, 0, vh, 1, stream());
  launch_spinor_to_aos(F64, oo, static_cast<char*>(out) + vh * 192, 0, vh, 1, stream());
  QHIP(hipGetLastError());
}

}  // namespace qcu
