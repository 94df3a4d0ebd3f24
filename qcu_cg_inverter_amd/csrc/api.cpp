// api.cpp -- libqcu.so C ABI: the five drop-in entry points of qcu.h and the
// extended API of qcu_ex.h, on top of the resident-field runtime.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "qcu_host.h"
#include "qcu_shift_storage.h"

namespace qcu {

// ------------------------------------------------------------ errors ------
static int g_last_code = QCU_SUCCESS;
static std::string g_last_msg;

void check_hip(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return;
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) from %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  throw Error(e == hipErrorOutOfMemory ? QCU_ERROR_OUT_OF_MEMORY : QCU_ERROR_HIP, buf);
}

static void record(int code, const std::string& msg) {
  g_last_code = code;
  g_last_msg = msg;
}
void set_last_error(int code, const std::string& msg) { record(code, msg); }

template <typename F>
static int guarded(F&& f) {
  try {
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    f();
    return QCU_SUCCESS;
  } catch (const Error& e) {
    record(e.code, e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    record(QCU_ERROR_OUT_OF_MEMORY, "host allocation failed");
    return QCU_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    record(QCU_ERROR_STATE, e.what());
    return QCU_ERROR_STATE;
  }
}

// drop-in entry points report like the reference's checkCudaErrors: message on
// stderr, then exit(-1) unless QCU_EXIT_ON_ERROR=0
template <typename F>
static void dropin(const char* name, F&& f) {
  int rc = guarded(f);
  if (rc == QCU_SUCCESS) return;
  fprintf(stderr, "qcu: %s failed (error %d): %s\n", name, rc, g_last_msg.c_str());
  const char* env = getenv("QCU_EXIT_ON_ERROR");
  if (!(env && strcmp(env, "0") == 0)) exit(-1);
}

static bool quiet() {
  const char* e = getenv("QCU_QUIET");
  return e && strcmp(e, "0") != 0;
}

// ------------------------------------------------------------ buffers -----
DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}
void DevBuf::release() {
  if (p) QHIP(hipFree(p));
  p = nullptr;
  n = 0;
}
void DevBuf::alloc(size_t bytes) {
  release();
  if (bytes == 0) return;
  QHIP(hipMalloc(&p, bytes));
  n = bytes;
}
void DevBuf::ensure(size_t bytes) {
  if (bytes > n) alloc(bytes);
}

Context& ctx() {
  static Context c;
  return c;
}

DevBuf* Context::buf(const std::string& name, size_t bytes) {
  auto& slot = scratch[name];
  if (!slot) slot.reset(new DevBuf);
  if (slot->n < bytes) {
    QHIP(hipStreamSynchronize(stream));  // the old buffer may still be in use
    slot->alloc(bytes);
  }
  return slot.get();
}

void ensure_init() {
  Context& c = ctx();
  if (c.inited) return;
  int n = 0;
  QHIP(hipGetDeviceCount(&n));
  if (n <= 0) throw Error(QCU_ERROR_HIP, "no HIP device");
  QHIP(hipSetDevice(c.device));
  QHIP(hipStreamCreate(&c.stream));  // default priority, blocking w.r.t. the null stream
  noalloc_setup(c.device);           // device address of qcuHopNoAlloc's static scratch
  qcuDefaultSolveParams(&c.default_solve);
  c.default_solve.precision = QCU_PREC_DOUBLE;
  c.default_solve.tol = 1e-12;
  c.inited = true;
  load_tune_cache();
}

void validate_lattice(const int L[4]) {
  for (int i = 0; i < 4; ++i) QREQ(L[i] > 0, "lattice sizes must be positive");
  QREQ(L[0] % 2 == 0, "Lx must be even");
  double v = (double)L[0] * L[1] * L[2] * L[3];
  QREQ(v / 2 < 2147483647.0, "half volume must fit in 31 bits");
}

// --------------------------------------------------------- gauge fields ---
size_t GaugeField::bytes() const {
  size_t s = aos.n;
  for (auto& a : packed)
    for (auto& b : a) s += b.n;
  for (auto& kv : clovers)
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 2; ++q) s += kv.second->A[p][q].n + kv.second->Ainv[p][q].n;
  return s;
}

const void* GaugeField::get(int prec, int recon) {
  DevBuf& b = packed[prec][recon == 12 ? 1 : 0];
  if (!b.p) {
    b.alloc(gauge_bytes(prec, recon, geo.vh));
    launch_gauge_pack(prec, recon, aos.p, b.p, geo.vh, stream());
    QHIP(hipGetLastError());
  }
  return b.p;
}

CloverField* GaugeField::clover(double ck, int prec, bool full_inverse) {
  auto it = clovers.find(ck);
  CloverField* cf;
  if (it == clovers.end()) {
    std::unique_ptr<CloverField> c(new CloverField);
    c->ck = ck;
    c->vh = geo.vh;
    cf = c.get();
    clovers[ck] = std::move(c);
  } else {
    cf = it->second.get();
  }
  const size_t half64 = clover_bytes(F64, geo.vh) / 2;
  // FP64 parts: A on both parities, A^-1 on the odd one, A^-1 on the even one if asked
  const bool need_a = !cf->A[F64][0].p, need_ie = full_inverse && !cf->Ainv[F64][0].p;
  if (need_a || need_ie) {
    const void* g18 = get(F64, 18);
    DevBuf* G = ctx().buf("clover_G", (size_t)6 * 9 * 16 * geo.vh);
    DevBuf* nb = ctx().buf("clover_nbad", sizeof(int));
    QHIP(hipMemsetAsync(nb->p, 0, sizeof(int), stream()));
    for (int p = 0; p < 2; ++p) {
      const bool want_a = need_a, want_i = p == 1 ? need_a : (need_ie || (need_a && full_inverse));
      if (!want_a && !want_i) continue;
      if (want_a) cf->A[F64][p].alloc(half64);
      if (want_i) cf->Ainv[F64][p].alloc(half64);
      launch_clover_leaves(g18, G->p, geo, p, stream());
      launch_clover_assemble(G->p, want_a ? cf->A[F64][p].p : nullptr, want_i ? cf->Ainv[F64][p].p : nullptr,
                             geo.vh, ck, static_cast<int*>(nb->p), stream());
    }
    QHIP(hipGetLastError());
    int nbad = 0;
    QHIP(hipMemcpyAsync(&nbad, nb->p, sizeof(int), hipMemcpyDeviceToHost, stream()));
    QHIP(hipStreamSynchronize(stream()));
    cf->nbad += nbad;
  }
  // inner-precision parts: A_ee and A^-1_oo
  if (prec != F64 && !cf->A[prec][0].p) {
    const size_t half = clover_bytes(prec, geo.vh) / 2;
    cf->A[prec][0].alloc(half);
    cf->Ainv[prec][1].alloc(half);
    launch_clover_convert(prec, cf->a(F64, 0), cf->A[prec][0].p, geo.vh, stream());
    launch_clover_convert(prec, cf->ainv(F64, 1), cf->Ainv[prec][1].p, geo.vh, stream());
    QHIP(hipGetLastError());
  }
  return cf;
}

static std::unique_ptr<GaugeField> new_gauge(const int L[4]) {
  validate_lattice(L);
  std::unique_ptr<GaugeField> g(new GaugeField);
  for (int i = 0; i < 4; ++i) g->L[i] = L[i];
  g->geo = make_geom(L);
  g->aos.alloc((size_t)8 * g->geo.vh * 9 * 16);
  return g;
}

std::unique_ptr<GaugeField> make_gauge_from_aos(const void* aos_device, const int L[4]) {
  auto g = new_gauge(L);
  QHIP(hipMemcpyAsync(g->aos.p, aos_device, g->aos.n, hipMemcpyDeviceToDevice, stream()));
  return g;
}

std::unique_ptr<GaugeField> make_gauge_random(const int L[4], uint64_t seed, double eps) {
  auto g = new_gauge(L);
  launch_gauge_random(g->aos.p, g->geo.vh, seed, eps, stream());
  QHIP(hipGetLastError());
  return g;
}

// ------------------------------------------------------------- tuning -----
// "hop:prec:recon:vh:nrhs" for the plain epilogue, with ":e<epi>" for the others
std::string tune_key(int prec, int recon, long vh, int nrhs, int epi) {
  char b[96];
  if (epi == EPI_PLAIN) snprintf(b, sizeof b, "hop:%d:%d:%ld:%d", prec, recon, vh, nrhs);
  else snprintf(b, sizeof b, "hop:%d:%d:%ld:%d:e%d", prec, recon, vh, nrhs, epi);
  return b;
}
// the shape tuned for this epilogue, else the one tuned for the plain epilogue, else
// the default
Tune tuned_hop(int prec, int recon, long vh, int nrhs, int epi) {
  auto it = ctx().tune.find(tune_key(prec, recon, vh, nrhs, epi));
  if (it != ctx().tune.end()) return it->second;
  it = ctx().tune.find(tune_key(prec, recon, vh, nrhs, EPI_PLAIN));
  if (it != ctx().tune.end()) return it->second;
  Tune t;
  if (nrhs > 1) {  // every right-hand side in one block per 64-site chunk, its links
    // staged in LDS once (hop_lds_kernel; on 32^3x64 with 12 right-hand sides 22-29 %
    // faster than the best shape sharing link loads within a wave), 4 waves per block,
    // blocks round-robin over the XCDs
    t.lds = 2;  // a clover shared by every right-hand side staged in LDS as well
    t.block = nrhs >= 4 ? 256 : 128;
    t.xcd = 0;
    t.tz = -1;  // tiles of all of z by 8 t-planes (ignored where they do not fit)
    t.tt = 8;
  }
  return t;
}
void set_tuned_hop(int prec, int recon, long vh, int nrhs, const Tune& t) {
  for (int e = EPI_PLAIN; e <= EPI_CLOV_XPAY; ++e) ctx().tune.erase(tune_key(prec, recon, vh, nrhs, e));
  ctx().tune[tune_key(prec, recon, vh, nrhs, EPI_PLAIN)] = t;
}

void hop(GaugeField& g, int prec, int recon, int epi, int parity, int dagger, double tbc, Vec in, Vec out, Vec x,
         const void* clov, double coef, int nrhs, double* partials, hipStream_t st) {
  QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range");
  const void* cl[kMaxRhs];
  double co[kMaxRhs];
  for (int r = 0; r < nrhs; ++r) {
    cl[r] = clov;
    co[r] = coef;
  }
  hop_v(g, prec, recon, epi, parity, dagger, tbc, in, out, x, cl, co, nrhs, partials, st);
}

void hop_v(GaugeField& g, int prec, int recon, int epi, int parity, int dagger, double tbc, Vec in, Vec out, Vec x,
           const void* const* clovs, const double* coefs, int nrhs, double* partials, hipStream_t st) {
  QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range");
  HopArgs a;
  memset(&a, 0, sizeof a);
  a.in = in.p;
  a.out = out.p;
  a.x = x.p;
  a.in_stride = in.stride;
  a.out_stride = out.stride;
  a.x_stride = x.stride;
  a.gauge = g.get(prec, recon);
  a.clov_shared = 1;
  for (int r = 0; r < nrhs; ++r) {
    a.clov[r] = clovs[r];
    a.coef[r] = coefs[r];
    if (clovs[r] != clovs[0]) a.clov_shared = 0;
  }
  a.partials = partials;
  a.geo = g.geo;
  a.parity = parity;
  a.dagger = dagger;
  a.tbc = tbc;
  a.nrhs = nrhs;
  Tune t = tuned_hop(prec, recon, g.geo.vh, nrhs, epi);
  a.xcd_remap = t.xcd;
  a.gmajor = t.gm;
  a.rpw = t.rpw;
  a.spt = t.spt;
  a.cached = t.cached;
  a.lds = t.lds;
  a.walk = make_walk(g.geo, t.ty, t.tz < 0 ? g.geo.L[2] : t.tz, t.tt);
  launch_hop(prec, recon, epi, a, t.block, t.rb, t.rpt, st);
}

}  // namespace qcu

using namespace qcu;

// =================================================================== qcu.h ===
static QcuParam g_param;

// fingerprint of the caller's interface gauge (launch_fingerprint), on the host
static uint64_t gauge_fingerprint(const void* aos, long vh) {
  DevBuf* s = ctx().buf("fingerprint", sizeof(uint64_t) * (kFingerprintBlocks + 1));
  launch_fingerprint(aos, (size_t)8 * vh * 9 * 16, static_cast<uint64_t*>(s->p), stream());
  QHIP(hipGetLastError());
  uint64_t h = 0;
  QHIP(hipMemcpyAsync(&h, s->p, sizeof h, hipMemcpyDeviceToHost, stream()));
  QHIP(hipStreamSynchronize(stream()));
  return h;
}

// 1: the drop-in calls trust that a gauge pointer they have seen still holds the same
// links (the copy is remade only by initGridSize / loadQcuGauge or a new pointer)
static bool dropin_trust_pointer() {
  const char* e = getenv("QCU_DROPIN_GAUGE_CACHE");
  return e && strcmp(e, "0") != 0;
}

static void require_aligned(const void* p, const char* what) {
  QREQ((reinterpret_cast<uintptr_t>(p) & 15) == 0, std::string(what) + " must be 16-byte aligned (complex128)");
}

// pinned host word the fingerprint of a deferred check lands in
static uint64_t* fingerprint_host() {
  static uint64_t* h = nullptr;
  if (!h) QHIP(hipHostMalloc(reinterpret_cast<void**>(&h), sizeof(uint64_t), hipHostMallocDefault));
  return h;
}

// remake the drop-in copy from the caller's gauge, whose fingerprint is fp
static void dropin_rebuild(void* gauge, const QcuParam* param, bool same_geom, uint64_t fp) {
  Context::Dropin& d = ctx().dropin;
  if (d.g) {
    QHIP(hipStreamSynchronize(stream()));
    drop_cached_batches(d.g.get());
  }
  if (same_geom) {  // reuse the allocation: new links, no packed copies, no clover
    GaugeField& g = *d.g;
    for (auto& a : g.packed)
      for (auto& b : a) b.release();
    g.clovers.clear();
    QHIP(hipMemcpyAsync(g.aos.p, gauge, g.aos.n, hipMemcpyDeviceToDevice, stream()));
  } else {
    d.g.reset();
    d.g = make_gauge_from_aos(gauge, param->lattice_size);
  }
  launch_gauge_reconstruct(d.g->aos.p, d.g->geo.vh, stream());
  QHIP(hipGetLastError());
  d.ptr = gauge;
  d.fp = fp;
}

// The packed copy and clover fields behind fullDslashQcu and cg_inverter.  The
// reference reads the caller's gauge on every call (src/dslash.cu:73,
// src/cg_inverter.cu:312-339), so a caller may change links in place between calls
// (an HMC update, smearing) or pass a reused address.  The copy is therefore checked
// against a fingerprint of the caller's buffer on every call -- one read of the gauge
// -- and remade, with its clover fields and solver state dropped, when it differs.
// QCU_DROPIN_GAUGE_CACHE=1 skips the check.  One copy is kept: a new pointer or
// geometry replaces it.  With `deferred` (fullDslashQcu), a copy that exists for this
// pointer and geometry is returned at once and the fingerprint is only queued: the
// caller runs its work, synchronises, and asks dropin_gauge_changed whether to redo it
// with a remade copy, which saves a synchronisation in the common case.
static GaugeField& dropin_gauge(void* gauge, const QcuParam* param, bool reload, bool* deferred = nullptr) {
  Context& c = ctx();
  if (deferred) *deferred = false;
  QREQ(gauge != nullptr, "gauge pointer is NULL");
  QREQ(param != nullptr, "QcuParam pointer is NULL");
  require_aligned(gauge, "gauge");
  validate_lattice(param->lattice_size);
  Context::Dropin& d = c.dropin;
  const bool same_geom = d.g && memcmp(d.g->L, param->lattice_size, sizeof(int) * 4) == 0;
  const bool same = same_geom && d.ptr == gauge;
  if (same && !reload && dropin_trust_pointer()) return *d.g;
  const long vh = (long)param->lattice_size[0] * param->lattice_size[1] * param->lattice_size[2] *
                  param->lattice_size[3] / 2;
  if (deferred && same && !reload) {
    DevBuf* sbuf = ctx().buf("fingerprint", sizeof(uint64_t) * (kFingerprintBlocks + 1));
    launch_fingerprint(gauge, (size_t)8 * vh * 9 * 16, static_cast<uint64_t*>(sbuf->p), stream());
    QHIP(hipGetLastError());
    QHIP(hipMemcpyAsync(fingerprint_host(), sbuf->p, sizeof(uint64_t), hipMemcpyDeviceToHost, stream()));
    *deferred = true;
    return *d.g;
  }
  const uint64_t fp = gauge_fingerprint(gauge, vh);
  if (same && !reload && fp == d.fp) return *d.g;
  dropin_rebuild(gauge, param, same_geom, fp);
  return *d.g;
}

// after a deferred dropin_gauge and a synchronisation of the library stream: true
// when the caller's gauge had changed, in which case the copy has been remade and the
// caller redoes its work
static bool dropin_gauge_changed(void* gauge, const QcuParam* param) {
  const uint64_t fp = *fingerprint_host();
  if (fp == ctx().dropin.fp) return false;
  dropin_rebuild(gauge, param, true, fp);
  return true;
}

extern "C" void initGridSize(QcuGrid_t* grid, QcuParam* p_param, void* gauge, void* fermion_in, void* fermion_out) {
  (void)fermion_in;
  (void)fermion_out;
  dropin("initGridSize", [&] {
    ensure_init();
    QREQ(grid && p_param, "initGridSize: NULL argument");
    for (int i = 0; i < 4; ++i)
      QREQ(grid->grid_size[i] == 1, "initGridSize: this build keeps the lattice on one GPU (grid_size must be 1,1,1,1)");
    validate_lattice(p_param->lattice_size);
    Context& c = ctx();
    memcpy(c.L, p_param->lattice_size, sizeof c.L);
    c.have_geom = true;
    g_param = *p_param;
    if (gauge) dropin_gauge(gauge, p_param, true);
  });
}

extern "C" void loadQcuGauge(void* gauge, QcuParam* param) {
  dropin("loadQcuGauge", [&] {
    ensure_init();
    GaugeField& g = dropin_gauge(gauge, param, true);
    g.get(F64, 12);
    QHIP(hipStreamSynchronize(stream()));
  });
}

// One kernel on the caller's buffers: the gauge is read on every call, as the
// reference's returned value (mpiDslashNaive) reads it, so no copy or fingerprint is
// needed here.
extern "C" void dslashQcu(void* fermion_out, void* fermion_in, void* gauge, QcuParam* param, int parity) {
  dropin("dslashQcu", [&] {
    ensure_init();
    QREQ(parity == 0 || parity == 1, "dslashQcu: parity must be 0 or 1");
    QREQ(fermion_out && fermion_in && gauge && param, "dslashQcu: NULL argument");
    require_aligned(fermion_in, "fermion_in");
    require_aligned(fermion_out, "fermion_out");
    require_aligned(gauge, "gauge");
    validate_lattice(param->lattice_size);
    const Geom g = make_geom(param->lattice_size);
    const char* i = static_cast<const char*>(fermion_in);
    char* o = static_cast<char*>(fermion_out);
    QREQ(o + g.vh * 192 <= i || i + g.vh * 192 <= o, "dslashQcu: fermion_out and fermion_in overlap");
    launch_hop_iface(fermion_in, fermion_out, gauge, g, parity, 0, 1.0, 1.0, stream());
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(stream()));
  });
}

extern "C" void fullDslashQcu(void* fermion_out, void* fermion_in, void* gauge, QcuParam* param, int dagger_flag) {
  dropin("fullDslashQcu", [&] {
    ensure_init();
    QREQ(fermion_out && fermion_in, "fullDslashQcu: NULL spinor");
    require_aligned(fermion_in, "fermion_in");
    require_aligned(fermion_out, "fermion_out");
    bool deferred = false;
    GaugeField& g = dropin_gauge(gauge, param, false, &deferred);
    qcuOperator op = {1.0, 1.0, 1, 12};
    apply_m_iface(g, op, fermion_out, fermion_in, dagger_flag ? 1 : 0);
    QHIP(hipStreamSynchronize(stream()));
    if (deferred && dropin_gauge_changed(gauge, param)) {  // the links changed in place: redo
      apply_m_iface(*ctx().dropin.g, op, fermion_out, fermion_in, dagger_flag ? 1 : 0);
      QHIP(hipStreamSynchronize(stream()));
    }
  });
}

extern "C" void cg_inverter(void* b_vector, void* x_vector, void* gauge, QcuParam* param) {
  qcuSolveStats st;
  memset(&st, 0, sizeof st);
  // A solve that stops at max_iter (or breaks down) is not an error of the call: like
  // the reference (src/cg_inverter.cu:331-345) it leaves the last iterate in x and
  // prints cg_failed!!!; qcuGetLastError() then returns QCU_ERROR_NOT_CONVERGED (or
  // QCU_ERROR_BREAKDOWN).  The solve has synchronised by then.
  int rc = guarded([&] {
    ensure_init();
    QREQ(b_vector && x_vector, "cg_inverter: NULL spinor");
    require_aligned(b_vector, "b_vector");
    require_aligned(x_vector, "x_vector");
    GaugeField& g = dropin_gauge(gauge, param, false);
    qcuOperator op = {1.0, 1.0, 1, 12};
    void* xs[1] = {x_vector};
    const void* bs[1] = {b_vector};
    solve_batch(g, op, ctx().default_solve, 1, xs, bs, &st);
    QHIP(hipStreamSynchronize(stream()));
  });
  if (rc != QCU_SUCCESS && rc != QCU_ERROR_NOT_CONVERGED && rc != QCU_ERROR_BREAKDOWN) {
    const std::string msg = g_last_msg;
    dropin("cg_inverter", [&] { throw Error(rc, msg); });
  }
  if (!quiet()) {
    // the reference's progress lines (src/cg_inverter.cu:330,342-349)
    if (st.converged) printf("number of iteration %d\ncg success!!!\n", st.iterations);
    else printf("cg_failed!!!\n");
    printf("true residual ||b - M x|| / ||b|| = %.3e\n", st.true_residual);
  }
}

// ==================================================== qcu_shift_storage.h ===
// [a, a + na) and [b, b + nb) must not overlap
static void require_disjoint(const void* a, size_t na, const void* b, size_t nb, const char* who) {
  const char* x = static_cast<const char*>(a);
  const char* y = static_cast<const char*>(b);
  QREQ(x + na <= y || y + nb <= x, std::string(who) + ": dst and src overlap");
}

// kind 0: one half spinor, kind 1: a gauge (src/qcu_shift_storage.cu:166-194)
static void shift_storage(const char* who, int kind, void* dst, void* src, int dir, int Lx, int Ly, int Lz, int Lt) {
  ensure_init();
  QREQ(dst && src, std::string(who) + ": NULL pointer");
  QREQ(dir == TO_COALESCE || dir == TO_NON_COALESCE, std::string(who) + ": shift_direction must be 0 or 1");
  const int L[4] = {Lx, Ly, Lz, Lt};
  validate_lattice(L);
  long vh = (long)Lx * Ly * Lz * Lt / 2;
  size_t aos = (size_t)vh * (kind == 0 ? 24 : 8 * 18) * 8;
  size_t soa = (size_t)vh * (kind == 0 ? 24 : 8 * 12) * 8;
  const bool to_soa = dir == TO_COALESCE;
  require_disjoint(dst, to_soa ? soa : aos, src, to_soa ? aos : soa, who);
  launch_shift_storage(kind, to_soa, src, dst, vh, stream());
  QHIP(hipGetLastError());
  QHIP(hipStreamSynchronize(stream()));
}

extern "C" void shiftVectorStorage(void* dst_vec, void* src_vec, int shift_direction, int Lx, int Ly, int Lz, int Lt) {
  dropin("shiftVectorStorage",
         [&] { shift_storage("shiftVectorStorage", 0, dst_vec, src_vec, shift_direction, Lx, Ly, Lz, Lt); });
}

extern "C" void shiftGaugeStorage(void* dst_vec, void* src_vec, int shift_direction, int Lx, int Ly, int Lz, int Lt) {
  dropin("shiftGaugeStorage",
         [&] { shift_storage("shiftGaugeStorage", 1, dst_vec, src_vec, shift_direction, Lx, Ly, Lz, Lt); });
}

// ================================================================ qcu_ex.h ===
extern "C" {

static int lex_eo(const char* who, int to_eo, void* dst, const void* src, const int L[4], int field) {
  return guarded([&] {
    ensure_init();
    QREQ(dst && src && L, std::string(who) + ": NULL argument");
    QREQ(field == QCU_FIELD_SPINOR || field == QCU_FIELD_GAUGE, std::string(who) + ": field must be spinor or gauge");
    validate_lattice(L);
    long vol = (long)L[0] * L[1] * L[2] * L[3];
    QREQ(vol < 2147483648L, std::string(who) + ": volume must be below 2^31 sites");
    int nc = field == QCU_FIELD_SPINOR ? 12 : 9, nf = field == QCU_FIELD_SPINOR ? 1 : 4;
    const size_t n = (size_t)vol * nc * nf * 16;
    require_disjoint(dst, n, src, n, who);
    launch_lex_eo(to_eo, src, dst, L, nc, nf, stream());
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(stream()));
  });
}

int qcuLexToEvenOdd(void* eo, const void* lex, const int L[4], int field) {
  return lex_eo("qcuLexToEvenOdd", 1, eo, lex, L, field);
}
int qcuEvenOddToLex(void* lex, const void* eo, const int L[4], int field) {
  return lex_eo("qcuEvenOddToLex", 0, lex, eo, L, field);
}

int qcuInit(int device) {
  return guarded([&] {
    Context& c = ctx();
    if (c.inited && c.device != device) throw Error(QCU_ERROR_STATE, "qcuInit: already initialised on another device");
    c.device = device;
    ensure_init();
  });
}

int qcuFinalize(void) {
  return guarded([&] {
    Context& c = ctx();
    if (!c.inited) return;
    QHIP(hipStreamSynchronize(c.stream));
    drop_cached_batches(nullptr);
    c.dropin.g.reset();
    c.dropin.ptr = nullptr;
    c.scratch.clear();
    QHIP(hipStreamDestroy(c.stream));
    c.stream = nullptr;
    c.inited = false;
  });
}

int qcuGetLastError(void) { return g_last_code; }
const char* qcuGetLastErrorString(void) { return g_last_msg.c_str(); }
void qcuClearLastError(void) { record(QCU_SUCCESS, ""); }
const char* qcuVersion(void) { return "qcu-mi355x 0.1 (gfx950)"; }

int qcuMalloc(void** ptr, size_t bytes) {
  return guarded([&] {
    ensure_init();
    QREQ(ptr, "qcuMalloc: NULL");
    QHIP(hipMalloc(ptr, bytes));
  });
}
int qcuFree(void* ptr) {
  return guarded([&] {
    ensure_init();
    QHIP(hipStreamSynchronize(stream()));
    QHIP(hipFree(ptr));
  });
}
int qcuMemcpyHtoD(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    ensure_init();
    QHIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    QHIP(hipStreamSynchronize(stream()));
  });
}
int qcuMemcpyDtoH(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    ensure_init();
    QHIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    QHIP(hipStreamSynchronize(stream()));
  });
}
int qcuMemcpyDtoD(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    ensure_init();
    QHIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
    QHIP(hipStreamSynchronize(stream()));
  });
}
int qcuMemset(void* dst, int value, size_t bytes) {
  return guarded([&] {
    ensure_init();
    QHIP(hipMemsetAsync(dst, value, bytes, stream()));
    QHIP(hipStreamSynchronize(stream()));
  });
}
int qcuSynchronize(void) {
  return guarded([&] {
    ensure_init();
    QHIP(hipStreamSynchronize(stream()));
  });
}
int qcuMemInfo(size_t* free_bytes, size_t* total_bytes) {
  return guarded([&] {
    ensure_init();
    QHIP(hipMemGetInfo(free_bytes, total_bytes));
  });
}

int qcuGaugeCreate(qcuGauge* out, const void* gauge, const int lattice_size[4]) {
  return guarded([&] {
    ensure_init();
    QREQ(out && gauge && lattice_size, "qcuGaugeCreate: NULL argument");
    auto g = make_gauge_from_aos(gauge, lattice_size);
    QHIP(hipStreamSynchronize(stream()));
    *out = reinterpret_cast<qcuGauge>(g.release());
  });
}
int qcuGaugeCreateRandom(qcuGauge* out, const int lattice_size[4], uint64_t seed, double eps) {
  return guarded([&] {
    ensure_init();
    QREQ(out && lattice_size, "qcuGaugeCreateRandom: NULL argument");
    QREQ(eps >= 0.0, "eps must be >= 0");
    auto g = make_gauge_random(lattice_size, seed, eps);
    QHIP(hipStreamSynchronize(stream()));
    *out = reinterpret_cast<qcuGauge>(g.release());
  });
}
int qcuGaugeDestroy(qcuGauge h) {
  return guarded([&] {
    ensure_init();
    QHIP(hipStreamSynchronize(stream()));
    drop_cached_batches(reinterpret_cast<GaugeField*>(h));
    delete reinterpret_cast<GaugeField*>(h);
  });
}
int qcuGaugeExport(qcuGauge h, void* gauge) {
  return guarded([&] {
    ensure_init();
    QREQ(h && gauge, "qcuGaugeExport: NULL argument");
    GaugeField& g = *reinterpret_cast<GaugeField*>(h);
    QHIP(hipMemcpyAsync(gauge, g.aos.p, g.aos.n, hipMemcpyDeviceToDevice, stream()));
    QHIP(hipStreamSynchronize(stream()));
  });
}
int qcuCloverExport(qcuGauge h, const qcuOperator* op, void* A, void* Ainv, int* nfallback) {
  return guarded([&] {
    ensure_init();
    QREQ(h && op, "qcuCloverExport: NULL argument");
    GaugeField& g = *reinterpret_cast<GaugeField*>(h);
    CloverField* cf = g.clover(op->kappa * op->csw, F64, true);
    size_t half = (size_t)g.geo.vh * 72 * 16;
    for (int p = 0; p < 2; ++p) {
      if (A) launch_clover_unpack(cf->a(F64, p), static_cast<char*>(A) + p * half, g.geo.vh, stream());
      if (Ainv) launch_clover_unpack(cf->ainv(F64, p), static_cast<char*>(Ainv) + p * half, g.geo.vh, stream());
    }
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(stream()));
    if (nfallback) *nfallback = cf->nbad;
  });
}

int qcuSourceGaussian(void* v, const int L[4], uint64_t seed) {
  return guarded([&] {
    ensure_init();
    QREQ(v && L, "qcuSourceGaussian: NULL argument");
    validate_lattice(L);
    long n = (long)L[0] * L[1] * L[2] * L[3] * 12;
    launch_source_gauss(v, n, seed, stream());
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(stream()));
  });
}

int qcuSourcePoint(void* v, const int L[4], const int coords[4], int spin, int colour) {
  return guarded([&] {
    ensure_init();
    QREQ(v && L && coords, "qcuSourcePoint: NULL argument");
    validate_lattice(L);
    QREQ(spin >= 0 && spin < 4 && colour >= 0 && colour < 3, "spin must be 0..3 and colour 0..2");
    for (int i = 0; i < 4; ++i) QREQ(coords[i] >= 0 && coords[i] < L[i], "coordinate outside the lattice");
    long vh = (long)L[0] * L[1] * L[2] * L[3] / 2;
    int parity = (coords[0] + coords[1] + coords[2] + coords[3]) & 1;
    long idx = (((long)coords[3] * L[2] + coords[2]) * L[1] + coords[1]) * (L[0] / 2) + coords[0] / 2;
    long off = ((parity * vh + idx) * 12 + spin * 3 + colour) * 16;
    static const double one[2] = {1.0, 0.0};
    QHIP(hipMemsetAsync(v, 0, (size_t)2 * vh * 192, stream()));
    QHIP(hipMemcpyAsync(static_cast<char*>(v) + off, one, 16, hipMemcpyHostToDevice, stream()));
    QHIP(hipStreamSynchronize(stream()));
  });
}

int qcuHop(qcuGauge h, void* out, const void* in, int parity, int dagger, int t_boundary, int recon, int precision) {
  return guarded([&] {
    ensure_init();
    QREQ(h && out && in, "qcuHop: NULL argument");
    QREQ(parity == 0 || parity == 1, "parity must be 0 or 1");
    QREQ(recon == 12 || recon == 18, "recon must be 12 or 18");
    QREQ(precision >= 0 && precision <= 2, "bad precision");
    QREQ(t_boundary == 1 || t_boundary == -1, "t_boundary must be +1 or -1");
    GaugeField& g = *reinterpret_cast<GaugeField*>(h);
    long vh = g.geo.vh;
    DevBuf* a = ctx().buf("hop_in", spinor_bytes(F64, vh));
    DevBuf* b = ctx().buf("hop_out", spinor_bytes(F64, vh));
    launch_spinor_from_aos(precision, in, 0, vec(a->p), vh, 1, stream());
    hop(g, precision, recon, EPI_PLAIN, parity, dagger ? 1 : 0, (double)t_boundary, vec(a->p), vec(b->p),
        vec(nullptr), nullptr, 1.0, 1, nullptr, stream());
    launch_spinor_to_aos(precision, vec(b->p), out, 0, vh, 1, stream());
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(stream()));
  });
}

int qcuHopBatch(qcuGauge h, int nrhs, void* const* out, const void* const* in, int parity, int dagger,
                int t_boundary, int recon, int precision) {
  return guarded([&] {
    ensure_init();
    QREQ(h && out && in, "qcuHopBatch: NULL argument");
    QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range (1..16)");
    QREQ(parity == 0 || parity == 1, "parity must be 0 or 1");
    QREQ(recon == 12 || recon == 18, "recon must be 12 or 18");
    QREQ(precision >= 0 && precision <= 2, "bad precision");
    QREQ(t_boundary == 1 || t_boundary == -1, "t_boundary must be +1 or -1");
    for (int r = 0; r < nrhs; ++r) QREQ(out[r] && in[r], "qcuHopBatch: NULL spinor");
    GaugeField& g = *reinterpret_cast<GaugeField*>(h);
    long vh = g.geo.vh;
    long fb = (long)spinor_bytes(precision, vh);
    DevBuf* a = ctx().buf("hopb_in", (size_t)fb * nrhs);
    DevBuf* b = ctx().buf("hopb_out", (size_t)fb * nrhs);
    for (int r = 0; r < nrhs; ++r) launch_spinor_from_aos(precision, in[r], 0, vec(a->c() + r * fb), vh, 1, stream());
    hop(g, precision, recon, EPI_PLAIN, parity, dagger ? 1 : 0, (double)t_boundary, vec(a->p, fb), vec(b->p, fb),
        vec(nullptr), nullptr, 1.0, nrhs, nullptr, stream());
    for (int r = 0; r < nrhs; ++r) launch_spinor_to_aos(precision, vec(b->c() + r * fb), out[r], 0, vh, 1, stream());
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(stream()));
  });
}

int qcuApplyM(qcuGauge h, const qcuOperator* op, void* out, const void* in, int dagger) {
  return guarded([&] {
    ensure_init();
    QREQ(h && op && out && in, "qcuApplyM: NULL argument");
    apply_m_iface(*reinterpret_cast<GaugeField*>(h), *op, out, in, dagger ? 1 : 0);
    QHIP(hipStreamSynchronize(stream()));
  });
}

int qcuApplyComposite(qcuGauge h, const qcuOperator* op, int kind, void* out, const void* in, int dagger) {
  return guarded([&] {
    ensure_init();
    QREQ(h && op && out && in, "qcuApplyComposite: NULL argument");
    apply_composite_iface(*reinterpret_cast<GaugeField*>(h), *op, kind, out, in, dagger ? 1 : 0);
    QHIP(hipStreamSynchronize(stream()));
  });
}

int qcuApplyMhat(qcuGauge h, const qcuOperator* op, void* out, const void* in, int dagger, int precision) {
  return guarded([&] {
    ensure_init();
    QREQ(h && op && out && in, "qcuApplyMhat: NULL argument");
    QREQ(precision >= 0 && precision <= 2, "bad precision");
    apply_mhat_iface(*reinterpret_cast<GaugeField*>(h), *op, out, in, dagger ? 1 : 0, precision);
    QHIP(hipStreamSynchronize(stream()));
  });
}

void qcuDefaultSolveParams(qcuSolveParams* sp) {
  if (!sp) return;
  sp->tol = 1e-10;
  sp->max_iter = 0;
  sp->precision = QCU_PREC_SINGLE;
  sp->delta = 0.0;  // by inner precision: 1e-3 for FP32, 0.1 for FP16
  sp->graph = 1;
  sp->chunk = 8;
}

int qcuSetDefaultSolve(const qcuSolveParams* sp) {
  return guarded([&] {
    ensure_init();
    QREQ(sp, "qcuSetDefaultSolve: NULL");
    ctx().default_solve = *sp;
  });
}

int qcuSolve(qcuGauge h, const qcuOperator* op, const qcuSolveParams* sp, void* x, const void* b,
             qcuSolveStats* stats) {
  return guarded([&] {
    ensure_init();
    QREQ(h && op && sp && x && b, "qcuSolve: NULL argument");
    void* xs[1] = {x};
    const void* bs[1] = {b};
    solve_batch(*reinterpret_cast<GaugeField*>(h), *op, *sp, 1, xs, bs, stats);
  });
}

int qcuSolveBatch(qcuGauge h, const qcuOperator* op, const qcuSolveParams* sp, int nrhs, void* const* x,
                  const void* const* b, qcuSolveStats* stats) {
  return guarded([&] {
    ensure_init();
    QREQ(h && op && sp && x && b, "qcuSolveBatch: NULL argument");
    QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range (1..16)");
    solve_batch(*reinterpret_cast<GaugeField*>(h), *op, *sp, nrhs, x, b, stats);
  });
}

int qcuBenchHop(qcuGauge h, int precision, int recon, int nrhs, int iters, double* ms) {
  return guarded([&] {
    ensure_init();
    QREQ(h && ms && iters > 0, "qcuBenchHop: bad argument");
    QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range");
    GaugeField& g = *reinterpret_cast<GaugeField*>(h);
    long vh = g.geo.vh;
    size_t fb = spinor_bytes(precision, vh);
    DevBuf* src = ctx().buf("bench_src", (size_t)vh * 192 * nrhs);
    DevBuf* a = ctx().buf("bench_a", fb * nrhs);
    DevBuf* b = ctx().buf("bench_b", fb * nrhs);
    launch_source_gauss(src->p, vh * 12 * nrhs, 12345, stream());
    launch_spinor_from_aos(precision, src->p, vh * 192, vec(a->p, (long)fb), vh, nrhs, stream());
    launch_spinor_from_aos(precision, src->p, vh * 192, vec(b->p, (long)fb), vh, nrhs, stream());
    Vec va = vec(a->p, (long)fb), vb = vec(b->p, (long)fb);
    for (int w = 0; w < 3; ++w)
      hop(g, precision, recon, EPI_PLAIN, w & 1, 0, -1.0, (w & 1) ? vb : va, (w & 1) ? va : vb, vec(nullptr), nullptr,
          1.0, nrhs, nullptr, stream());
    hipEvent_t e0, e1;
    QHIP(hipEventCreate(&e0));
    QHIP(hipEventCreate(&e1));
    QHIP(hipEventRecord(e0, stream()));
    for (int i = 0; i < iters; ++i)
      hop(g, precision, recon, EPI_PLAIN, i & 1, 0, -1.0, (i & 1) ? vb : va, (i & 1) ? va : vb, vec(nullptr), nullptr,
          1.0, nrhs, nullptr, stream());
    QHIP(hipEventRecord(e1, stream()));
    QHIP(hipEventSynchronize(e1));
    float t = 0.f;
    QHIP(hipEventElapsedTime(&t, e0, e1));
    QHIP(hipEventDestroy(e0));
    QHIP(hipEventDestroy(e1));
    QHIP(hipGetLastError());
    *ms = (double)t / iters;
  });
}

int qcuBenchLinkProduct(int mfma, long site_legs, int iters, double* ms, double* checksum) {
  return guarded([&] {
    ensure_init();
    QREQ(ms && iters > 0, "qcuBenchLinkProduct: bad argument");
    QREQ(mfma == 0 || mfma == 1, "mfma must be 0 or 1");
    QREQ(site_legs > 0 && site_legs % 262144 == 0, "site_legs must be a positive multiple of 262144");
    DevBuf* o = ctx().buf("linkprod", sizeof(double) * 262144);
    launch_linkprod(mfma, site_legs, static_cast<double*>(o->p), stream());
    hipEvent_t e0, e1;
    QHIP(hipEventCreate(&e0));
    QHIP(hipEventCreate(&e1));
    QHIP(hipEventRecord(e0, stream()));
    for (int i = 0; i < iters; ++i) launch_linkprod(mfma, site_legs, static_cast<double*>(o->p), stream());
    QHIP(hipEventRecord(e1, stream()));
    QHIP(hipEventSynchronize(e1));
    QHIP(hipGetLastError());
    float t = 0.f;
    QHIP(hipEventElapsedTime(&t, e0, e1));
    QHIP(hipEventDestroy(e0));
    QHIP(hipEventDestroy(e1));
    *ms = (double)t / iters;
    if (checksum) {
      std::vector<double> h(262144);
      QHIP(hipMemcpy(h.data(), o->p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
      double sum = 0.0;
      for (double v : h) sum += v;
      *checksum = sum;
    }
  });
}

int qcuBenchCopy(size_t bytes, int iters, double* ms) {
  return guarded([&] {
    ensure_init();
    QREQ(ms && iters > 0 && bytes > 0, "qcuBenchCopy: bad argument");
    DevBuf* a = ctx().buf("copy_a", bytes);
    DevBuf* b = ctx().buf("copy_b", bytes);
    QREQ(bytes % 16 == 0, "qcuBenchCopy: bytes must be a multiple of 16");
    QHIP(hipMemsetAsync(a->p, 1, bytes, stream()));
    launch_copy(b->p, a->p, bytes, stream());
    hipEvent_t e0, e1;
    QHIP(hipEventCreate(&e0));
    QHIP(hipEventCreate(&e1));
    QHIP(hipEventRecord(e0, stream()));
    for (int i = 0; i < iters; ++i) launch_copy((i & 1) ? a->p : b->p, (i & 1) ? b->p : a->p, bytes, stream());
    QHIP(hipEventRecord(e1, stream()));
    QHIP(hipEventSynchronize(e1));
    float t = 0.f;
    QHIP(hipEventElapsedTime(&t, e0, e1));
    QHIP(hipEventDestroy(e0));
    QHIP(hipEventDestroy(e1));
    *ms = (double)t / iters;
  });
}

}  // extern "C"
