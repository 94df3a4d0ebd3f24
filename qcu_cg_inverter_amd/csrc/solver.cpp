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
// so the stopping test |s| <= tol |b| is on the residual of the system itself.
// Mixed precision: the loop runs in FP32 or FP16 (+ per-site norm); when |s| drops
// below delta times its value at the last update (or reaches the target) the FP32
// iterate is added to the FP64 solution and s is recomputed in FP64 (reliable
// updates, Clark et al., Comput. Phys. Commun. 181 (2010) 1517).
// Scalars stay on the device; the host looks every `chunk` iterations, and the
// chunk can be a captured hipGraph.  Every per-right-hand-side decision uses only
// that right-hand side's own numbers, so a solve gives the same bits alone or batched.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

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
  if (o.clover) o.cf = g.clover(op.kappa * op.csw, prec);
  g.get(prec, op.recon);
  g.get(F64, op.recon);
  return o;
}

// out = Mhat in (Mhat^dag if dag) on even fields; tmp holds the odd intermediate
static void mhat(OpView& o, int prec, int dag, Vec in, Vec out, Vec tmp, int nrhs, double* partials, hipStream_t st) {
  const double k2 = o.op.kappa * o.op.kappa;
  if (o.clover) {
    hop(*o.g, prec, o.op.recon, EPI_CLOV, 1, dag, o.tbc(), in, tmp, vec(nullptr), o.cf->ainv(prec, 1), 1.0, nrhs,
        nullptr, st);
    hop(*o.g, prec, o.op.recon, EPI_CLOV_XPAY, 0, dag, o.tbc(), tmp, out, in, o.cf->a(prec, 0), -k2, nrhs, partials,
        st);
  } else {
    hop(*o.g, prec, o.op.recon, EPI_PLAIN, 1, dag, o.tbc(), in, tmp, vec(nullptr), nullptr, 1.0, nrhs, nullptr, st);
    hop(*o.g, prec, o.op.recon, EPI_XPAY, 0, dag, o.tbc(), tmp, out, in, nullptr, -k2, nrhs, partials, st);
  }
}

// out_p = A_pp in_p - kappa D^s in_(1-p) for one parity p (FP64)
static void full_op_parity(OpView& o, int p, int dag, Vec in_p, Vec in_q, Vec out, int nrhs, hipStream_t st) {
  if (o.clover)
    hop(*o.g, F64, o.op.recon, EPI_CLOV_XPAY, p, dag, o.tbc(), in_q, out, in_p, o.cf->a(F64, p), -o.op.kappa, nrhs,
        nullptr, st);
  else
    hop(*o.g, F64, o.op.recon, EPI_XPAY, p, dag, o.tbc(), in_q, out, in_p, nullptr, -o.op.kappa, nrhs, nullptr, st);
}

static Vec at(Vec v, int r) { return vec(static_cast<char*>(v.p) + (long)r * v.stride, v.stride); }

// sums of |v_r|^2 to the host (synchronises)
static void norms_host(int prec, Vec v, long vh, int nrhs, double* out) {
  long nw = n_partials(vh);
  DevBuf* part = ctx().buf("norm_part", sizeof(double) * nw * nrhs);
  DevBuf* sums = ctx().buf("norm_sums", sizeof(double) * nrhs);
  launch_norm2(prec, v, vh, nrhs, static_cast<double*>(part->p), stream());
  launch_sum_partials(static_cast<double*>(part->p), nw, nrhs, static_cast<double*>(sums->p), stream());
  QHIP(hipMemcpyAsync(out, sums->p, sizeof(double) * nrhs, hipMemcpyDeviceToHost, stream()));
  QHIP(hipStreamSynchronize(stream()));
}

static CgScal* pinned_scalars() {
  static CgScal* h = nullptr;
  if (!h) QHIP(hipHostMalloc((void**)&h, sizeof(CgScal) * kMaxRhs, 0));
  return h;
}

void solve_batch(GaugeField& g, const qcuOperator& op, const qcuSolveParams& spin, int nrhs, void* const* x,
                 const void* const* b, qcuSolveStats* stats) {
  auto t_start = std::chrono::steady_clock::now();
  qcuSolveParams sp = spin;
  QREQ(sp.tol > 0.0 && std::isfinite(sp.tol), "tol must be positive");
  QREQ(sp.precision >= 0 && sp.precision <= 2, "bad precision");
  if (sp.chunk < 1) sp.chunk = 1;
  if (sp.delta <= 0.0 || sp.delta >= 1.0) sp.delta = 0.1;
  for (int r = 0; r < nrhs; ++r) QREQ(x[r] && b[r], "NULL spinor");
  const long vh = g.geo.vh;
  const long nw = n_partials(vh);
  const int P = sp.precision;
  const int XP = P == F64 ? F64 : F32;
  if (sp.max_iter <= 0) sp.max_iter = (int)std::min<long>(10 * 2 * vh, 1000000);
  OpView o = make_view(g, op, P);
  if (P != F64 && o.clover) g.clover(op.kappa * op.csw, F64);
  maybe_autotune(g, P, op.recon, nrhs);  // QCU_TUNE=1: before any graph is captured
  hipStream_t st = stream();

  const long fb = (long)spinor_bytes(F64, vh), pb = (long)spinor_bytes(P, vh), xb = (long)spinor_bytes(XP, vh);
  auto f64 = [&](const char* n) { return vec(ctx().buf(n, (size_t)fb * nrhs)->p, fb); };
  auto slo = [&](const char* n) { return vec(ctx().buf(n, (size_t)pb * nrhs)->p, pb); };
  Vec be = f64("cg_be"), bo = f64("cg_bo"), bh = f64("cg_bh"), xe = f64("cg_xe"), shi = f64("cg_shi"),
      t64 = f64("cg_t64");
  Vec S = P == F64 ? shi : slo("cg_s");
  Vec R = slo("cg_r"), Pv = slo("cg_p"), T = slo("cg_t"), TO = slo("cg_to");
  Vec X = P == F64 ? xe : vec(ctx().buf("cg_xlo", (size_t)xb * nrhs)->p, xb);
  double* part = static_cast<double*>(ctx().buf("cg_part", sizeof(double) * 3 * nw * nrhs)->p);
  double *part_rr = part, *part_tt = part + nw * nrhs, *part_ss = part + 2 * nw * nrhs;
  CgScal* dsc = static_cast<CgScal*>(ctx().buf("cg_scal", sizeof(CgScal) * kMaxRhs)->p);
  CgScal* hsc = pinned_scalars();

  // sources into the internal layout and |b|^2
  for (int r = 0; r < nrhs; ++r) {
    launch_spinor_from_aos(F64, b[r], 0, at(be, r), vh, 1, st);
    launch_spinor_from_aos(F64, static_cast<const char*>(b[r]) + vh * 192, 0, at(bo, r), vh, 1, st);
  }
  QHIP(hipGetLastError());
  std::vector<double> nbe(nrhs), nbo(nrhs), bnorm2(nrhs), ss0(nrhs);
  norms_host(F64, be, vh, nrhs, nbe.data());
  norms_host(F64, bo, vh, nrhs, nbo.data());
  for (int r = 0; r < nrhs; ++r) {
    bnorm2[r] = nbe[r] + nbo[r];
    if (!std::isfinite(bnorm2[r])) throw Error(QCU_ERROR_NOT_FINITE, "source contains NaN or Inf");
  }

  // b^ = b_e + kappa D_eo A_oo^-1 b_o
  if (o.clover) {
    launch_clover_apply(F64, o.cf->ainv(F64, 1), bo, t64, vh, nrhs, st);
    hop(g, F64, op.recon, EPI_XPAY, 0, 0, o.tbc(), t64, bh, be, nullptr, op.kappa, nrhs, nullptr, st);
  } else {
    hop(g, F64, op.recon, EPI_XPAY, 0, 0, o.tbc(), bo, bh, be, nullptr, op.kappa, nrhs, nullptr, st);
  }
  launch_zero(F64, xe, vh, nrhs, st);
  QHIP(hipMemcpyAsync(shi.p, bh.p, (size_t)fb * nrhs, hipMemcpyDeviceToDevice, st));
  if (P != F64) {
    launch_spinor_convert(F64, shi, P, S, vh, nrhs, 0, st);
    launch_zero(XP, X, vh, nrhs, st);
  }
  launch_zero(P, Pv, vh, nrhs, st);  // scratch may hold stale, non-finite bits
  QHIP(hipGetLastError());
  norms_host(F64, shi, vh, nrhs, ss0.data());

  std::vector<int> upd(nrhs, 0), done(nrhs, 0);
  for (int r = 0; r < nrhs; ++r) {
    CgScal& s = hsc[r];
    memset(&s, 0, sizeof s);
    s.ss = ss0[r];
    s.ss_ref = ss0[r];
    s.target2 = sp.tol * sp.tol * bnorm2[r];
    s.first = 1;
    s.converged = ss0[r] <= s.target2;
    s.active = !s.converged;
    done[r] = s.converged;
  }
  QHIP(hipMemcpyAsync(dsc, hsc, sizeof(CgScal) * nrhs, hipMemcpyHostToDevice, st));

  auto iteration = [&](hipStream_t s) {
    mhat(o, P, 1, S, R, TO, nrhs, part_rr, s);
    launch_cg_beta(part_rr, nw, dsc, nrhs, s);
    launch_cg_xpay(P, R, Pv, dsc, vh, nrhs, s);
    mhat(o, P, 0, Pv, T, TO, nrhs, part_tt, s);
    launch_cg_alpha(part_tt, nw, dsc, nrhs, s);
    launch_cg_update(P, XP, X, Pv, T, S, dsc, part_ss, vh, nrhs, s);
    launch_cg_ss(part_ss, nw, dsc, nrhs, s);
  };

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool any_active = false;
  for (int r = 0; r < nrhs; ++r) any_active |= hsc[r].active != 0;
  if (sp.graph && any_active) {
    QHIP(hipStreamSynchronize(st));
    QHIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < sp.chunk; ++i) iteration(st);
    QHIP(hipStreamEndCapture(st, &graph));
    QHIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }

  const char* dbg_env = getenv("QCU_DEBUG");
  const bool debug = dbg_env && dbg_env[0] && dbg_env[0] != '0';
  // s = b^ - Mhat x in FP64 for right-hand side r; returns |s|^2
  auto reliable_update = [&](int r) {
    if (P != F64) {
      launch_spinor_convert(XP, at(X, r), F64, at(xe, r), vh, 1, 1, st);
      launch_zero(XP, at(X, r), vh, 1, st);
    }
    mhat(o, F64, 0, at(xe, r), at(shi, r), at(t64, r), 1, nullptr, st);
    launch_axpby(F64, 1.0, at(bh, r), -1.0, at(shi, r), vh, 1, st);
    double ss = 0.0;
    norms_host(F64, at(shi, r), vh, 1, &ss);
    if (P != F64) launch_spinor_convert(F64, at(shi, r), P, at(S, r), vh, 1, 0, st);
    upd[r]++;
    if (debug) fprintf(stderr, "qcu cg reliable update r=%d |s|^2=%.6e\n", r, ss);
    return ss;
  };

  int breakdown = 0;
  while (any_active) {
    if (exec) {
      QHIP(hipGraphLaunch(exec, st));
    } else {
      for (int i = 0; i < sp.chunk; ++i) iteration(st);
    }
    QHIP(hipGetLastError());
    QHIP(hipMemcpyAsync(hsc, dsc, sizeof(CgScal) * nrhs, hipMemcpyDeviceToHost, st));
    QHIP(hipStreamSynchronize(st));
    if (debug) {
      for (int r = 0; r < nrhs; ++r) {
        const CgScal& s = hsc[r];
        fprintf(stderr, "qcu cg r=%d it=%d ss=%.6e gamma=%.6e tt=%.6e alpha=%.6e beta=%.6e conv=%d act=%d brk=%d\n", r,
                s.iter, s.ss, s.gamma, s.tt, s.alpha, s.beta, s.converged, s.active, s.breakdown);
      }
    }
    bool changed = false;
    any_active = false;
    for (int r = 0; r < nrhs; ++r) {
      CgScal& s = hsc[r];
      if (!s.active) continue;
      changed = true;
      if (s.breakdown || !std::isfinite(s.ss)) {
        breakdown = 1;
        s.active = 0;
        s.converged = 0;
        continue;
      }
      bool cap = s.iter >= sp.max_iter;
      bool need = s.converged || cap || (P != F64 && s.ss < sp.delta * sp.delta * s.ss_ref);
      if (need) {
        double ss = reliable_update(r);
        s.ss = ss;
        s.ss_ref = ss;
        s.converged = ss <= s.target2;
        if (!std::isfinite(ss)) {
          breakdown = 1;
          s.converged = 0;
        }
        if (s.converged || cap || !std::isfinite(ss)) s.active = 0;
      }
      done[r] = s.converged;
      any_active |= s.active != 0;
    }
    if (changed) QHIP(hipMemcpyAsync(dsc, hsc, sizeof(CgScal) * nrhs, hipMemcpyHostToDevice, st));
  }
  if (exec) QHIP(hipGraphExecDestroy(exec));
  if (graph) QHIP(hipGraphDestroy(graph));
  if (P != F64) launch_spinor_convert(XP, X, F64, xe, vh, nrhs, 1, st);

  // x_o = A_oo^-1 (b_o + kappa D_oe x_e), into shi (no longer needed)
  Vec xo = shi;
  if (o.clover) {
    hop(g, F64, op.recon, EPI_XPAY, 1, 0, o.tbc(), xe, t64, bo, nullptr, op.kappa, nrhs, nullptr, st);
    launch_clover_apply(F64, o.cf->ainv(F64, 1), t64, xo, vh, nrhs, st);
  } else {
    hop(g, F64, op.recon, EPI_XPAY, 1, 0, o.tbc(), xe, xo, bo, nullptr, op.kappa, nrhs, nullptr, st);
  }
  // true residual of the full system in FP64: r = b - M x
  Vec re = bh, ro = t64;  // reuse
  full_op_parity(o, 0, 0, xe, xo, re, nrhs, st);
  full_op_parity(o, 1, 0, xo, xe, ro, nrhs, st);
  launch_axpby(F64, 1.0, be, -1.0, re, vh, nrhs, st);
  launch_axpby(F64, 1.0, bo, -1.0, ro, vh, nrhs, st);
  for (int r = 0; r < nrhs; ++r) {
    launch_spinor_to_aos(F64, at(xe, r), x[r], 0, vh, 1, st);
    launch_spinor_to_aos(F64, at(xo, r), static_cast<char*>(x[r]) + vh * 192, 0, vh, 1, st);
  }
  QHIP(hipGetLastError());
  std::vector<double> nre(nrhs), nro(nrhs);
  norms_host(F64, re, vh, nrhs, nre.data());
  norms_host(F64, ro, vh, nrhs, nro.data());
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();

  bool all_conv = true;
  double worst = 0.0;
  int maxit = 0, nupd = 0;
  for (int r = 0; r < nrhs; ++r) {
    double tr = bnorm2[r] > 0 ? std::sqrt((nre[r] + nro[r]) / bnorm2[r]) : 0.0;
    worst = std::max(worst, tr);
    all_conv &= done[r] != 0;
    maxit = std::max(maxit, hsc[r].iter);
    nupd += upd[r];
    if (stats && nrhs > 1) {
      stats[r].iterations = hsc[r].iter;
      stats[r].reliable_updates = upd[r];
      stats[r].converged = done[r];
      stats[r].breakdown = breakdown;
      stats[r].true_residual = tr;
      stats[r].seconds = secs;
    }
  }
  if (stats && nrhs == 1) {
    stats->iterations = maxit;
    stats->reliable_updates = nupd;
    stats->converged = all_conv;
    stats->breakdown = breakdown;
    stats->true_residual = worst;
    stats->seconds = secs;
  }
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
  mhat(o, prec, dagger, a, b, t, 1, nullptr, stream());
  launch_spinor_to_aos(prec, b, out, 0, vh, 1, stream());
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
  full_op_parity(o, 0, dagger, ie, io, oe, 1, stream());
  full_op_parity(o, 1, dagger, io, ie, oo, 1, stream());
  launch_spinor_to_aos(F64, oe, out, 0, vh, 1, stream());
  launch_spinor_to_aos(F64, oo, static_cast<char*>(out) + vh * 192, 0, vh, 1, stream());
  QHIP(hipGetLastError());
}

}  // namespace qcu
