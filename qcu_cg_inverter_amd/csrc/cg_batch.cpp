// cg_batch.cpp -- slot-based batched even-odd CGLS (algorithm in solver.cpp).
#include "cg_batch.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace qcu {

size_t CgBatch::bytes_for(long vh, int prec, int cap) {
  size_t fb = spinor_bytes(F64, vh), pb = spinor_bytes(prec, vh), xb = spinor_bytes(iterate_precision(prec), vh);
  size_t s = 4 * fb * cap + (prec == F64 ? 4 : 5) * pb * cap + (prec == F64 ? 0 : xb * cap);
  s += sizeof(double) * (4 * n_partials(vh) * cap + cap) + sizeof(CgScal) * cap + 2 * (size_t)vh * 192;
  return s;
}

CgBatch::CgBatch(GaugeField& g, int prec, int recon, int t_boundary, bool clover, int capacity,
                 const qcuSolveParams& sp, hipStream_t st)
    : g_(g), prec_(prec), xprec_(iterate_precision(prec)), recon_(recon), tbc_(t_boundary), clover_(clover),
      cap_(capacity), sp_(sp), st_(st) {
  QREQ(capacity >= 1 && capacity <= kMaxRhs, "batch capacity must be 1..16");
  QREQ(prec >= 0 && prec <= 2, "bad precision");
  QREQ(recon == 12 || recon == 18, "recon must be 12 or 18");
  QREQ(t_boundary == 1 || t_boundary == -1, "t_boundary must be +1 or -1");
  if (sp_.chunk < 1) sp_.chunk = 1;
  sp_.delta = reliable_delta(prec, sp_.delta);
  vh_ = g.geo.vh;
  nw_ = n_partials(vh_);
  fb_ = (long)spinor_bytes(F64, vh_);
  pb_ = (long)spinor_bytes(prec, vh_);
  xb_ = (long)spinor_bytes(xprec_, vh_);
  // the source itself is not kept: it is read again (or regenerated from its seed)
  // when the slot finishes
  f64_.alloc((size_t)4 * fb_ * cap_);
  char* f = f64_.c();
  bh_ = vec(f, fb_);
  xe_ = vec(f + 1 * fb_ * cap_, fb_);
  shi_ = vec(f + 2 * fb_ * cap_, fb_);
  t64_ = vec(f + 3 * fb_ * cap_, fb_);
  int nslo = prec == F64 ? 4 : 5;
  slo_.alloc((size_t)nslo * pb_ * cap_);
  char* s = slo_.c();
  R_ = vec(s, pb_);
  P_ = vec(s + 1 * pb_ * cap_, pb_);
  T_ = vec(s + 2 * pb_ * cap_, pb_);
  TO_ = vec(s + 3 * pb_ * cap_, pb_);
  S_ = prec == F64 ? shi_ : vec(s + 4 * pb_ * cap_, pb_);
  if (prec == F64) {
    X_ = xe_;
  } else {
    xlo_.alloc((size_t)xb_ * cap_);
    X_ = vec(xlo_.p, xb_);
  }
  part_.alloc(sizeof(double) * (4 * nw_ * cap_ + cap_));
  double* pp = static_cast<double*>(part_.p);
  part_rr_ = pp;
  part_tt_ = pp + nw_ * cap_;
  part_ss_ = pp + 2 * nw_ * cap_;
  sums_ = pp + 4 * nw_ * cap_;
  scal_.alloc(sizeof(CgScal) * cap_);
  dsc_ = static_cast<CgScal*>(scal_.p);
  QHIP(hipHostMalloc((void**)&hsc_, sizeof(CgScal) * cap_, 0));
  g.get(prec, recon);
  g.get(F64, recon);
  QHIP(hipStreamSynchronize(ctx().stream));  // packed fields were made on the context stream
  const char* e = getenv("QCU_DEBUG");
  debug_ = e && e[0] && e[0] != '0';
}

CgBatch::~CgBatch() {
  clear_graphs();
  if (hsc_) (void)hipHostFree(hsc_);
}

void CgBatch::clear_graphs() {
  for (auto& gr : graphs_) {
    if (gr.second) (void)hipGraphExecDestroy(gr.second);
    if (gr.first) (void)hipGraphDestroy(gr.first);
  }
  graphs_.clear();
}

void CgBatch::check_operator(const SlotSpec& s) const {
  QREQ(std::isfinite(s.kappa) && std::isfinite(s.csw), "kappa and csw must be finite");
  QREQ((s.csw != 0.0) == clover_, "right-hand side does not match the batch's operator type");
  QREQ(s.tol > 0.0 && std::isfinite(s.tol), "tol must be positive");
}

Vec CgBatch::at(Vec v, int r) const {
  const Vec base[] = {bh_, xe_, shi_, t64_, S_, R_, P_, T_, TO_, X_};
  bool ok = false;
  for (const Vec& b : base) ok |= b.p == v.p && b.stride == v.stride;
  if (!ok || r < 0 || r >= cap_) throw Error(QCU_ERROR_STATE, "internal: batch slot view out of range");
  return vec(static_cast<char*>(v.p) + (long)r * v.stride, v.stride);
}

double CgBatch::norm2_slot(Vec v, int r) {
  double* part = part_ss_ + nw_ * cap_;  // spare partials region
  launch_norm2(F64, at(v, r), vh_, 1, part, st_);
  launch_sum_partials(part, nw_, 1, sums_, st_);
  double out = 0.0;
  QHIP(hipMemcpyAsync(&out, sums_, sizeof(double), hipMemcpyDeviceToHost, st_));
  QHIP(hipStreamSynchronize(st_));
  return out;
}

// out = Mhat in (Mhat^dag if dag) for slots [r0, r0 + n), each with its own kappa and
// clover; tmp holds the odd intermediate
void CgBatch::mhat(int prec, int dag, Vec in, Vec out, Vec tmp, int r0, int n, double* partials) {
  if (r0 < 0 || n < 1 || r0 + n > n_) throw Error(QCU_ERROR_STATE, "internal: batch slot range out of range");
  const void* c1[kMaxRhs];
  const void* c2[kMaxRhs];
  double k1[kMaxRhs], k2[kMaxRhs];
  for (int j = 0; j < n; ++j) {
    const Slot& sl = slots_[r0 + j];
    c1[j] = sl.cf ? sl.cf->ainv(prec, 1) : nullptr;
    c2[j] = sl.cf ? sl.cf->a(prec, 0) : nullptr;
    k1[j] = 1.0;
    k2[j] = -sl.spec.kappa * sl.spec.kappa;
  }
  hop_v(g_, prec, recon_, clover_ ? EPI_CLOV : EPI_PLAIN, 1, dag, (double)tbc_, at(in, r0), at(tmp, r0), vec(nullptr),
        c1, k1, n, nullptr, st_);
  hop_v(g_, prec, recon_, clover_ ? EPI_CLOV_XPAY : EPI_XPAY, 0, dag, (double)tbc_, at(tmp, r0), at(out, r0),
        at(in, r0), c2, k2, n, partials ? partials + (long)r0 * nw_ : nullptr, st_);
}

void CgBatch::add(const SlotSpec& spec) { add_slots(&spec, 1); }

// The m new slots take [n_, n_ + m); their FP64 set-up runs batched over that run (a
// batched hop gives every slot the bits of its own launch), with two synchronisations
// in all.
void CgBatch::add_slots(const SlotSpec* specs, int m) {
  QREQ(m >= 1 && n_ + m <= cap_, "batch is full");
  for (int j = 0; j < m; ++j) check_operator(specs[j]);
  const int r0 = n_, r1 = n_ + m - 1;
  for (int j = 0; j < m; ++j) {
    Slot sl;
    sl.spec = specs[j];
    sl.updates = 0;
    sl.cf = nullptr;
    if (clover_) {
      sl.cf = g_.clover(sl.spec.kappa * sl.spec.csw, F64);
      if (prec_ != F64) g_.clover(sl.spec.kappa * sl.spec.csw, prec_);
    }
    if (sl.spec.max_iter <= 0) sl.spec.max_iter = (int)std::min<long>(20 * vh_, 1000000);
    slots_.push_back(sl);
  }
  n_ += m;
  if (clover_) QHIP(hipStreamSynchronize(ctx().stream));  // clover fields built there
  // b_e in shi_, b_o in t64_ (both scratch until the iteration starts)
  for (int r = r0; r <= r1; ++r) {
    const char* b = static_cast<const char*>(source_aos(slots_[r]));
    launch_spinor_from_aos(F64, b, 0, at(shi_, r), vh_, 1, st_);
    launch_spinor_from_aos(F64, b + vh_ * 192, 0, at(t64_, r), vh_, 1, st_);
  }
  QHIP(hipGetLastError());
  std::vector<double> ne(m), no(m);
  norm2_run(shi_, r0, r1, ne.data());
  norm2_run(t64_, r0, r1, no.data());
  QHIP(hipStreamSynchronize(st_));
  for (int j = 0; j < m; ++j) {
    slots_[r0 + j].bnorm2 = ne[j] + no[j];
    if (!std::isfinite(slots_[r0 + j].bnorm2)) {
      slots_.resize(r0);
      n_ = r0;
      throw Error(QCU_ERROR_NOT_FINITE, "source contains NaN or Inf");
    }
  }
  // b^ = b_e + kappa D_eo A_oo^-1 b_o, A_oo^-1 b_o in place over runs sharing a clover
  const void* none[kMaxRhs];
  double kap[kMaxRhs];
  for (int j = 0; j < m; ++j) {
    none[j] = nullptr;
    kap[j] = slots_[r0 + j].spec.kappa;
  }
  if (clover_) {
    for (int lo = r0; lo <= r1;) {
      int hi = lo;
      while (hi < r1 && slots_[hi + 1].cf == slots_[lo].cf) ++hi;
      launch_clover_apply(F64, slots_[lo].cf->ainv(F64, 1), at(t64_, lo), at(t64_, lo), vh_, hi - lo + 1, st_);
      lo = hi + 1;
    }
  }
  hop_v(g_, F64, recon_, EPI_XPAY, 0, 0, (double)tbc_, at(t64_, r0), at(bh_, r0), at(shi_, r0), none, kap, m,
        nullptr, st_);
  launch_zero(F64, at(xe_, r0), vh_, m, st_);
  for (int r = r0; r <= r1; ++r)
    QHIP(hipMemcpyAsync(at(shi_, r).p, at(bh_, r).p, fb_, hipMemcpyDeviceToDevice, st_));
  if (prec_ != F64) {
    launch_spinor_convert(F64, at(shi_, r0), prec_, at(S_, r0), vh_, m, 0, st_);
    launch_zero(xprec_, at(X_, r0), vh_, m, st_);
  }
  launch_zero(prec_, at(P_, r0), vh_, m, st_);  // scratch may hold stale, non-finite bits
  QHIP(hipGetLastError());
  std::vector<double> ss0(m);
  norm2_run(shi_, r0, r1, ss0.data());
  QHIP(hipStreamSynchronize(st_));
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<char> done(n_, 0);
  bool any = false;
  for (int j = 0; j < m; ++j) {
    Slot& sl = slots_[r0 + j];
    sl.t0 = t0;
    CgScal& c = hsc_[r0 + j];
    memset(&c, 0, sizeof c);
    c.ss = ss0[j];
    c.ss_ref = ss0[j];
    c.target2 = sl.spec.tol * sl.spec.tol * sl.bnorm2;
    c.first = 1;
    c.converged = ss0[j] <= c.target2;
    c.active = 1;
    done[r0 + j] = c.converged ? 1 : 0;
    any = any || c.converged;
  }
  clear_graphs();
  if (any) finish_slots(done);  // compaction moves slots within [r0, n_) only
  if (n_ > r0) QHIP(hipMemcpyAsync(dsc_ + r0, hsc_ + r0, sizeof(CgScal) * (n_ - r0), hipMemcpyHostToDevice, st_));
}

void CgBatch::launch_chunk() {
  const int n = n_;
  if (n == 0) return;
  auto iteration = [&]() {
    mhat(prec_, 1, S_, R_, TO_, 0, n, part_rr_);
    launch_cg_beta(part_rr_, nw_, dsc_, n, st_);
    launch_cg_xpay(prec_, xprec_, R_, P_, X_, dsc_, vh_, n, st_);
    mhat(prec_, 0, P_, T_, TO_, 0, n, part_tt_);
    launch_cg_alpha(part_tt_, nw_, dsc_, n, st_);
    launch_cg_update(prec_, T_, S_, dsc_, part_ss_, vh_, n, st_);
    launch_cg_ss(part_ss_, nw_, dsc_, n, st_);
  };
  if (!sp_.graph) {
    for (int i = 0; i < sp_.chunk; ++i) iteration();
    QHIP(hipGetLastError());
    return;
  }
  if ((int)graphs_.size() <= n) graphs_.resize(n + 1, {nullptr, nullptr});
  auto& gr = graphs_[n];
  if (!gr.second) {
    QHIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < sp_.chunk; ++i) iteration();
    QHIP(hipStreamEndCapture(st_, &gr.first));
    QHIP(hipGraphInstantiate(&gr.second, gr.first, nullptr, nullptr, 0));
  }
  QHIP(hipGraphLaunch(gr.second, st_));
}

// The iterate update x += alpha p of a chunk's last iteration is applied by the next
// iteration's p update; before the host reads x it applies it here (with the host's
// copy of alpha, which it then clears, so the device does not apply it again).
void CgBatch::flush_x(int r) {
  CgScal& c = hsc_[r];
  if (c.alpha != 0.0) launch_cg_xflush(prec_, xprec_, at(X_, r), at(P_, r), c.alpha, vh_, st_);
  c.alpha = 0.0;
}

// s = b^ - Mhat x in FP64 for slot r (the low-precision iterate folded into x first)
double CgBatch::reliable_update(int r) {
  flush_x(r);
  if (prec_ != F64) {
    launch_spinor_convert(xprec_, at(X_, r), F64, at(xe_, r), vh_, 1, 1, st_);
    launch_zero(xprec_, at(X_, r), vh_, 1, st_);
  }
  mhat(F64, 0, xe_, shi_, t64_, r, 1, nullptr);
  launch_axpby(F64, 1.0, at(bh_, r), -1.0, at(shi_, r), vh_, 1, st_);
  double ss = norm2_slot(shi_, r);
  if (prec_ != F64) launch_spinor_convert(F64, at(shi_, r), prec_, at(S_, r), vh_, 1, 0, st_);
  slots_[r].updates++;
  if (debug_) fprintf(stderr, "qcu cg reliable update tag=%llu |s|^2=%.6e\n", (unsigned long long)slots_[r].spec.tag, ss);
  return ss;
}

void CgBatch::harvest() {
  const int n = n_;
  if (n == 0) return;
  QHIP(hipMemcpyAsync(hsc_, dsc_, sizeof(CgScal) * n, hipMemcpyDeviceToHost, st_));
  QHIP(hipStreamSynchronize(st_));
  if (debug_)
    for (int r = 0; r < n; ++r)
      fprintf(stderr, "qcu cg tag=%llu it=%d ss=%.6e gamma=%.6e tt=%.6e alpha=%.6e beta=%.6e conv=%d brk=%d\n",
              (unsigned long long)slots_[r].spec.tag, hsc_[r].iter, hsc_[r].ss, hsc_[r].gamma, hsc_[r].tt,
              hsc_[r].alpha, hsc_[r].beta, hsc_[r].converged, hsc_[r].breakdown);
  std::vector<char> brk(n, 0), upd(n, 0), cap(n, 0);
  int lo = n, hi = -1;
  for (int r = 0; r < n; ++r) {
    const CgScal& s = hsc_[r];
    if (s.breakdown || !std::isfinite(s.ss)) {
      brk[r] = 1;
      continue;
    }
    cap[r] = s.iter >= slots_[r].spec.max_iter;
    if (s.converged || cap[r] || (prec_ != F64 && s.ss < sp_.delta * sp_.delta * s.ss_ref)) {
      upd[r] = 1;
      lo = std::min(lo, r);
      hi = std::max(hi, r);
    }
  }
  if (hi >= lo) {
    std::vector<double> ss = reliable_updates(lo, hi, upd);
    for (int r = lo; r <= hi; ++r) {
      if (!upd[r]) continue;
      CgScal& s = hsc_[r];
      const bool seemed_converged = s.converged;
      s.ss = ss[r - lo];
      s.ss_ref = s.ss;
      s.converged = s.ss <= s.target2;
      // The iterated residual reached the target but the true one did not: restart
      // from the true residual (p = M^dag s).  Keeping the old direction there can
      // stall an ill-conditioned solve for good (kappa = 1, the reference's
      // normalisation, on 4^4: 40000 iterations at 3e-11 against 9349 to 1e-12).
      if (seemed_converged && !s.converged) s.first = 1;
      if (!std::isfinite(s.ss)) brk[r] = 1;
    }
  }
  std::vector<char> done(n, 0);
  for (int r = 0; r < n; ++r) {
    if (brk[r]) done[r] = 2;
    else if (upd[r] && (hsc_[r].converged || cap[r])) done[r] = 1;
  }
  finish_slots(done);
  if (n_ > 0) QHIP(hipMemcpyAsync(dsc_, hsc_, sizeof(CgScal) * n_, hipMemcpyHostToDevice, st_));
}

// s = b^ - Mhat x in FP64 for the slots r in [lo, hi] with upd[r] set, the
// low-precision iterate of each folded into x first; returns |s|^2 per slot of the
// range.  With a low-precision loop the FP64 residual fields are scratch between
// updates, so the operator runs once over the whole range (the per-slot arithmetic
// is the same as alone); an FP64 loop updates its live residual slot by slot.
std::vector<double> CgBatch::reliable_updates(int lo, int hi, const std::vector<char>& upd) {
  const int n = hi - lo + 1;
  std::vector<double> out(n, 0.0);
  if (prec_ == F64) {
    for (int r = lo; r <= hi; ++r)
      if (upd[r]) out[r - lo] = reliable_update(r);
    return out;
  }
  for (int r = lo; r <= hi; ++r) {
    if (!upd[r]) continue;
    flush_x(r);
    launch_spinor_convert(xprec_, at(X_, r), F64, at(xe_, r), vh_, 1, 1, st_);
    launch_zero(xprec_, at(X_, r), vh_, 1, st_);
  }
  mhat(F64, 0, xe_, shi_, t64_, lo, n, nullptr);
  launch_axpby(F64, 1.0, at(bh_, lo), -1.0, at(shi_, lo), vh_, n, st_);
  double* part = part_ss_ + nw_ * cap_;
  launch_norm2(F64, at(shi_, lo), vh_, n, part, st_);
  launch_sum_partials(part, nw_, n, sums_, st_);
  QHIP(hipMemcpyAsync(out.data(), sums_, sizeof(double) * n, hipMemcpyDeviceToHost, st_));
  for (int r = lo; r <= hi; ++r) {
    if (!upd[r]) continue;
    launch_spinor_convert(F64, at(shi_, r), prec_, at(S_, r), vh_, 1, 0, st_);
    slots_[r].updates++;
  }
  QHIP(hipStreamSynchronize(st_));
  if (debug_)
    for (int r = lo; r <= hi; ++r)
      if (upd[r])
        fprintf(stderr, "qcu cg reliable update tag=%llu |s|^2=%.6e\n", (unsigned long long)slots_[r].spec.tag,
                out[r - lo]);
  return out;
}

// the slot's interface-layout source: the caller's buffer, or regenerated from its
// seed into the batch's staging buffer
const void* CgBatch::source_aos(const Slot& sl) {
  if (sl.spec.b) return sl.spec.b;
  aos_.ensure((size_t)2 * vh_ * 192);
  launch_source_gauss(aos_.p, 2 * vh_ * 12, sl.spec.seed, st_);
  return aos_.p;
}

// |v|^2 of slots [lo, hi] of field v into out[0 .. hi - lo] (enqueued; read after a
// synchronisation); the per-slot sums are those of norm2_slot
void CgBatch::norm2_run(Vec v, int lo, int hi, double* out) {
  const int m = hi - lo + 1;
  double* part = part_ss_ + nw_ * cap_;  // spare partials region
  launch_norm2(F64, at(v, lo), vh_, m, part, st_);
  launch_sum_partials(part, nw_, m, sums_, st_);
  QHIP(hipMemcpyAsync(out, sums_, sizeof(double) * m, hipMemcpyDeviceToHost, st_));
}

// x_o = A_oo^-1 (b_o + kappa D_oe x_e) and the true residual |b - M x|^2 of slots
// [lo, hi], each with its own kappa, clover and source, the FP64 operators batched over
// the run (a batched hop gives every slot the bits of its own launch).  Uses the slots'
// bh_, shi_ and t64_ as scratch (b^ is no longer needed).
void CgBatch::finish_run(int lo, int hi, std::vector<double>& nr) {
  const int m = hi - lo + 1;
  for (int r = lo; r <= hi; ++r) flush_x(r);
  if (prec_ != F64) {
    launch_spinor_convert(xprec_, at(X_, lo), F64, at(xe_, lo), vh_, m, 1, st_);
    launch_zero(xprec_, at(X_, lo), vh_, m, st_);
  }
  const void* none[kMaxRhs];
  const void* a_o[kMaxRhs];
  const void* a_e[kMaxRhs];
  double kap[kMaxRhs], mkap[kMaxRhs];
  for (int j = 0; j < m; ++j) {
    const Slot& sl = slots_[lo + j];
    none[j] = nullptr;
    a_o[j] = clover_ ? sl.cf->a(F64, 1) : nullptr;
    a_e[j] = clover_ ? sl.cf->a(F64, 0) : nullptr;
    kap[j] = sl.spec.kappa;
    mkap[j] = -sl.spec.kappa;
  }
  Vec xo = at(shi_, lo), tmp = at(bh_, lo), bq = at(t64_, lo);
  for (int r = lo; r <= hi; ++r)  // b_o
    launch_spinor_from_aos(F64, static_cast<const char*>(source_aos(slots_[r])) + vh_ * 192, 0, at(t64_, r), vh_,
                           1, st_);
  // x_o = A_oo^-1 (b_o + kappa D_oe x_e)
  if (clover_) {
    hop_v(g_, F64, recon_, EPI_XPAY, 1, 0, (double)tbc_, at(xe_, lo), tmp, bq, none, kap, m, nullptr, st_);
    for (int r = lo; r <= hi; ++r)
      launch_clover_apply(F64, slots_[r].cf->ainv(F64, 1), at(bh_, r), at(shi_, r), vh_, 1, st_);
  } else {
    hop_v(g_, F64, recon_, EPI_XPAY, 1, 0, (double)tbc_, at(xe_, lo), xo, bq, none, kap, m, nullptr, st_);
  }
  // r = b - M x with M = A - kappa D, odd then even parity
  std::vector<double> ro(m), re(m);
  hop_v(g_, F64, recon_, clover_ ? EPI_CLOV_XPAY : EPI_XPAY, 1, 0, (double)tbc_, at(xe_, lo), tmp, xo,
        clover_ ? a_o : none, mkap, m, nullptr, st_);
  launch_axpby(F64, 1.0, bq, -1.0, tmp, vh_, m, st_);
  norm2_run(bh_, lo, hi, ro.data());
  for (int r = lo; r <= hi; ++r)  // b_e
    launch_spinor_from_aos(F64, source_aos(slots_[r]), 0, at(t64_, r), vh_, 1, st_);
  hop_v(g_, F64, recon_, clover_ ? EPI_CLOV_XPAY : EPI_XPAY, 0, 0, (double)tbc_, xo, tmp, at(xe_, lo),
        clover_ ? a_e : none, mkap, m, nullptr, st_);
  launch_axpby(F64, 1.0, bq, -1.0, tmp, vh_, m, st_);
  for (int r = lo; r <= hi; ++r) {
    const Slot& sl = slots_[r];
    if (sl.spec.x) {
      launch_spinor_to_aos(F64, at(xe_, r), sl.spec.x, 0, vh_, 1, st_);
      launch_spinor_to_aos(F64, at(shi_, r), static_cast<char*>(sl.spec.x) + vh_ * 192, 0, vh_, 1, st_);
    }
  }
  norm2_run(bh_, lo, hi, re.data());
  QHIP(hipGetLastError());
  QHIP(hipStreamSynchronize(st_));
  for (int r = lo; r <= hi; ++r) nr[r] = re[r - lo] + ro[r - lo];  // |r_e|^2 + |r_o|^2
}

void CgBatch::finish_slots(const std::vector<char>& done) {
  const int n = n_;
  bool any = false;
  for (int r = 0; r < n; ++r) any = any || done[r];
  if (!any) return;  // the captured chunks stay valid
  std::vector<double> nr(n, 0.0);
  for (int lo = 0; lo < n;) {
    if (!done[lo]) {
      ++lo;
      continue;
    }
    int hi = lo;
    while (hi + 1 < n && done[hi + 1]) ++hi;
    finish_run(lo, hi, nr);
    lo = hi + 1;
  }
  // stats, then compaction top down, so the slot moved into a finished one has
  // already been dealt with
  for (int r = n - 1; r >= 0; --r) {
    if (!done[r]) continue;
    const Slot& sl = slots_[r];
    const bool breakdown = done[r] == 2;
    SlotResult res;
    res.tag = sl.spec.tag;
    res.stats.iterations = hsc_[r].iter;
    res.stats.reliable_updates = sl.updates;
    res.stats.breakdown = breakdown ? 1 : 0;
    res.stats.converged = (!breakdown && hsc_[r].converged) ? 1 : 0;
    res.stats.true_residual = sl.bnorm2 > 0 ? std::sqrt(nr[r] / sl.bnorm2) : 0.0;
    res.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - sl.t0).count();
    finished_.push_back(res);
    const int last = n_ - 1;
    if (r != last) move_slot(last, r);
    slots_.pop_back();
    n_--;
  }
  clear_graphs();
}

void CgBatch::move_slot(int from, int to) {
  Vec keep[] = {bh_, xe_, shi_, P_};
  for (const Vec& v : keep)
    QHIP(hipMemcpyAsync(at(v, to).p, at(v, from).p, v.stride, hipMemcpyDeviceToDevice, st_));
  if (prec_ != F64) {
    QHIP(hipMemcpyAsync(at(S_, to).p, at(S_, from).p, S_.stride, hipMemcpyDeviceToDevice, st_));
    QHIP(hipMemcpyAsync(at(X_, to).p, at(X_, from).p, X_.stride, hipMemcpyDeviceToDevice, st_));
  }
  slots_[to] = slots_[from];
  hsc_[to] = hsc_[from];
}

std::vector<SlotResult> CgBatch::take_finished() {
  std::vector<SlotResult> out;
  out.swap(finished_);
  return out;
}

}  // namespace qcu
