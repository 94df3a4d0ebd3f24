// cg_batch.h -- a set of right-hand-side slots solved together by even-odd CGLS
// (see solver.cpp for the algorithm), filled and emptied at chunk boundaries.
//
// Slots [0, active) are contiguous: a finished slot is replaced by the last active
// one (device copies), so every kernel runs over a dense range of right-hand sides.
// A slot starts at a chunk boundary and every decision about it uses only its own
// scalars, so its result does not depend on which other slots share the batch.
#pragma once

#include <chrono>
#include <vector>

#include "qcu_host.h"

namespace qcu {

struct SlotSpec {
  const void* b;        // interface full spinor (device), or null: Gaussian source from `seed`
  uint64_t seed;
  void* x;              // interface full spinor for the solution (device), or null
  double kappa, csw, tol;
  int max_iter;         // <= 0: 10 * volume
  uint64_t tag;         // caller's identifier
};

struct SlotResult {
  uint64_t tag;
  qcuSolveStats stats;
};

// Precision of the inner iterate x between reliable updates: FP64 for the plain
// solve, else FP32.  x gathers the corrections of one reliable-update cycle.  Kept in
// FP16 next to an FP16 inner loop it saves a quarter of the p / x update's traffic but
// its rounding (2^-11 of the site norm per update) cost iterations: 109-110 instead of
// 107 on the 12-source 32^3x64 clover batch (1 % faster), 107 instead of 101 on the
// 24^3x48 clover solve and 117 instead of 107 on the 16^3x32 Wilson solve (5-7 %
// slower), measured on MI355X.
inline int iterate_precision(int prec) { return prec == F64 ? F64 : F32; }

class CgBatch {
 public:
  // prec: inner precision; clover: slots carry a clover term (csw != 0)
  CgBatch(GaugeField& g, int prec, int recon, int t_boundary, bool clover, int capacity, const qcuSolveParams& sp,
          hipStream_t st);
  ~CgBatch();
  CgBatch(const CgBatch&) = delete;
  CgBatch& operator=(const CgBatch&) = delete;

  static size_t bytes_for(long vh, int prec, int capacity);
  size_t bytes() const { return bytes_for(g_.geo.vh, prec_, cap_); }
  int capacity() const { return cap_; }
  int active() const { return n_; }
  int free_slots() const { return cap_ - n_; }
  GaugeField& gauge() { return g_; }
  bool clover() const { return clover_; }
  int precision() const { return prec_; }

  // prepare a right-hand side in a free slot (synchronises the batch stream)
  void add(const SlotSpec& s);
  // prepare m right-hand sides in free slots, their set-up batched; a non-finite
  // source rejects all m and leaves the batch as it was
  void add_slots(const SlotSpec* specs, int m);
  // enqueue one chunk of iterations for the active slots (asynchronous)
  void launch_chunk();
  // after the chunk: reliable updates, convergence, finished slots (synchronises)
  void harvest();
  std::vector<SlotResult> take_finished();

 private:
  struct Slot {
    SlotSpec spec;
    double bnorm2;
    int updates;
    CloverField* cf;
    std::chrono::steady_clock::time_point t0;
  };
  void check_operator(const SlotSpec& s) const;
  void mhat(int prec, int dag, Vec in, Vec out, Vec tmp, int r0, int n, double* partials);
  void flush_x(int r);
  double reliable_update(int r);
  std::vector<double> reliable_updates(int lo, int hi, const std::vector<char>& upd);
  // finish the slots marked in `done` (1 converged or capped, 2 broke down), their FP64
  // work batched over contiguous runs of marked slots, then compact top down
  void finish_slots(const std::vector<char>& done);
  // x_o and |b - M x|^2 (into nr[r]) for slots [lo, hi], solutions written out
  void finish_run(int lo, int hi, std::vector<double>& nr);
  // |v|^2 of slots [lo, hi] of field v, in order (asynchronous: valid after a sync)
  void norm2_run(Vec v, int lo, int hi, double* out);
  const void* source_aos(const Slot& sl);
  void move_slot(int from, int to);
  // slot r of a batch field; v must be the field's base (slot 0), never a slot view
  Vec at(Vec v, int r) const;
  double norm2_slot(Vec v, int r);
  void clear_graphs();

  GaugeField& g_;
  int prec_, xprec_, recon_, tbc_;
  bool clover_;
  int cap_, n_ = 0;
  qcuSolveParams sp_;
  hipStream_t st_;
  long vh_, nw_, fb_, pb_, xb_;
  DevBuf f64_, slo_, xlo_, part_, scal_, tmp_, aos_;
  Vec bh_, xe_, shi_, t64_, S_, R_, P_, T_, TO_, X_;
  double *part_rr_, *part_tt_, *part_ss_, *sums_;
  CgScal* dsc_;
  CgScal* hsc_;
  std::vector<Slot> slots_;
  std::vector<SlotResult> finished_;
  std::vector<std::pair<hipGraph_t, hipGraphExec_t>> graphs_;  // per active count, built lazily
  bool debug_;
};

}  // namespace qcu
