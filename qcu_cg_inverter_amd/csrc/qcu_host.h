// qcu_host.h -- host-side runtime of libqcu.so: context, resident fields, errors.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "qcu_ex.h"
#include "qcu_kernels.h"

namespace qcu {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what, const char* file, int line);
#define QHIP(x) ::qcu::check_hip((x), #x, __FILE__, __LINE__)
#define QREQ(cond, msg)                                                         \
  do {                                                                          \
    if (!(cond)) throw ::qcu::Error(QCU_ERROR_INVALID_VALUE, std::string(msg)); \
  } while (0)

// reliable-update threshold: the caller's delta if in (0, 1), else 1e-3 for an FP32
// inner loop and 0.1 for FP16 (whose rounding, about 1e-3, needs the frequent
// updates).  Measured on 32^3x64 clover with 12 right-hand sides: FP32 takes the same
// 107 iterations from delta 0.1 to 0.001 with 8 updates down to 4; FP16 needs 107
// iterations at 0.1, 176 at 0.01 and 9090 at 0.001.
inline double reliable_delta(int prec, double d) {
  if (d > 0.0 && d < 1.0) return d;
  return prec == F16 ? 0.1 : 1e-3;
}

// device allocation owned by the library (grow-only when reused as scratch)
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void alloc(size_t bytes);
  void ensure(size_t bytes);
  void release();
  char* c() const { return static_cast<char*>(p); }
};

// A = 1 + ck T and A^-1, packed, per precision and parity, built on demand.  The
// even-odd solve reads A on both parities and A^-1 on the odd one in FP64, and A_ee and
// A^-1_oo in its inner precision; A^-1_ee (export, full-lattice compositions) is built
// only when asked for.  Low precisions are converted from FP64.
struct CloverField {
  double ck = 0.0;
  int nbad = 0;  // blocks inverted by the pivoted fallback, over the parities inverted so far
  long vh = 0;
  // serving slots (of every server) solving with this field; a server frees the field
  // only when this reaches zero, so servers that share a gauge never free a field
  // another one still uses
  int server_refs = 0;
  DevBuf A[3][2], Ainv[3][2];
  const void* a(int prec, int parity) const { return need(A[prec][parity]); }
  const void* ainv(int prec, int parity) const { return need(Ainv[prec][parity]); }
  static const void* need(const DevBuf& b) {
    if (!b.p) throw Error(QCU_ERROR_STATE, "internal: clover part not built");
    return b.p;
  }
};
// device bytes of the clover parts an even-odd solve with inner precision prec uses
inline size_t solver_clover_bytes(int prec, long vh) {
  size_t s = 3 * (clover_bytes(F64, vh) / 2);
  if (prec != F64) s += 2 * (clover_bytes(prec, vh) / 2);
  return s;
}

// a gauge field resident in HBM: the interface copy plus packed variants made on
// first use, and clover fields keyed by kappa * c_sw
struct GaugeField {
  int L[4];
  Geom geo;
  DevBuf aos;
  DevBuf packed[3][2];  // [precision][recon == 12]
  std::map<double, std::unique_ptr<CloverField>> clovers;
  const void* get(int prec, int recon);
  // the solve's parts in F64 and prec; full_inverse adds A^-1_ee (FP64)
  CloverField* clover(double ck, int prec, bool full_inverse = false);
  size_t bytes() const;
};

// launch shape of the hopping kernel: `block` threads in `rb` wave groups that share
// sites, `rpt` right-hand sides per thread, XCD-contiguous block order on/off, the
// (y, z, t) tile of the site walk (tz = 0: storage order; ty = 0: whole xy-planes),
// whether the right-hand side group is the slowest block index (gm) or the fastest,
// `rpw` right-hand sides per wave (lanes of a wave sharing sites), `spt` site blocks
// each block walks (sites per thread), and whether links / clover / x use ordinary
// (cached) loads instead of streaming ones; `lds` selects the kernel that stages a
// chunk's links in LDS for all right-hand sides (block = 64 x waves; rb, rpw, rpt and
// gm do not apply)
struct Tune {
  int block = 256;
  int rb = 1;
  int rpt = 1;
  int xcd = 1;
  int tz = 0;  // -1: all of z
  int tt = 0;
  int gm = 0;
  int ty = 0;
  int rpw = 1;
  int spt = 1;
  int cached = 0;
  int lds = 0;  // 1: every right-hand side in one block per 64-site chunk, its links staged in LDS;
                // 2: a clover shared by every right-hand side staged too
};

struct Context {
  bool inited = false;
  int device = 0;
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;
  // state of the drop-in entry points: one resident copy, of the gauge last passed,
  // with the fingerprint of the caller's buffer it was made from
  bool have_geom = false;
  int L[4] = {0, 0, 0, 0};
  struct Dropin {
    const void* ptr = nullptr;
    std::unique_ptr<GaugeField> g;
    uint64_t fp = 0;
  } dropin;
  qcuSolveParams default_solve;
  // scratch shared by calls (grow-only)
  std::map<std::string, std::unique_ptr<DevBuf>> scratch;
  std::map<std::string, Tune> tune;
  DevBuf* buf(const std::string& name, size_t bytes);
};

Context& ctx();
void ensure_init();
inline hipStream_t stream() { return ctx().stream; }

void validate_lattice(const int L[4]);
std::unique_ptr<GaugeField> make_gauge_from_aos(const void* aos_device, const int L[4]);
std::unique_ptr<GaugeField> make_gauge_random(const int L[4], uint64_t seed, double eps);

Tune tuned_hop(int prec, int recon, long vh, int nrhs, int epi = EPI_PLAIN);
std::string tune_key(int prec, int recon, long vh, int nrhs, int epi);
void set_tuned_hop(int prec, int recon, long vh, int nrhs, const Tune& t);
// autotuner (tune.cpp)
void load_tune_cache();
Tune autotune_hop(GaugeField& g, int prec, int recon, int nrhs, double* best_ms);
void maybe_autotune(GaugeField& g, int prec, int recon, int nrhs);

void set_last_error(int code, const std::string& msg);
// looks up the device address of qcuHopNoAlloc's static scratch (noalloc.cpp) on
// `device`, which must be current
void noalloc_setup(int device);

// one hop launch on nrhs strided fields
void hop(GaugeField& g, int prec, int recon, int epi, int parity, int dagger, double tbc, Vec in, Vec out, Vec x,
         const void* clov, double coef, int nrhs, double* partials, hipStream_t st);
// the same with a clover field and coefficient per right-hand side
void hop_v(GaugeField& g, int prec, int recon, int epi, int parity, int dagger, double tbc, Vec in, Vec out, Vec x,
           const void* const* clovs, const double* coefs, int nrhs, double* partials, hipStream_t st);

// solver entry (solver.cpp)
void solve_batch(GaugeField& g, const qcuOperator& op, const qcuSolveParams& sp, int nrhs, void* const* x,
                 const void* const* b, qcuSolveStats* stats);
void apply_mhat_iface(GaugeField& g, const qcuOperator& op, void* out, const void* in, int dagger, int prec);
// forget solver state that refers to gauge g (all gauges if null)
void drop_cached_batches(GaugeField* g);
void apply_m_iface(GaugeField& g, const qcuOperator& op, void* out, const void* in, int dagger);
void apply_composite_iface(GaugeField& g, const qcuOperator& op, int kind, void* out, const void* in, int dagger);

}  // namespace qcu
