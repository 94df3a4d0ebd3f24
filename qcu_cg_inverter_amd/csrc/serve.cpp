// serve.cpp -- request loop with continuous batching over resident gauge fields.
//
// The reference has no serving path; this is the caller side of its solver.  One
// worker thread owns the device work and keeps every batch (one CgBatch per gauge and
// operator type, each on its own stream) going on its own clock.  Each pass it
//   1. harvests every batch whose chunk of CG iterations has finished: reliable
//      updates, converged right-hand sides leave their slots, and requests whose
//      right-hand sides are all done complete;
//   2. admits queued right-hand sides, in arrival order, into idle batches (or new
//      ones) as slots and the HBM budget allow;
//   3. launches the next chunk on every idle batch that has work.
// So a small lattice's batch is not held back by a large one's chunks, and batches
// on different gauges overlap on the device.  Empty batches are released when queued
// work needs their memory.  A right-hand side enters at a chunk boundary and every
// decision about it uses only its own numbers, so its result is the same whichever
// requests share its batch.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <thread>

#include "cg_batch.h"
#include "qcu_host.h"

namespace qcu {

using Clock = std::chrono::steady_clock;

struct Request {
  qcuRequest spec;
  std::vector<void*> x;
  std::vector<const void*> b;
  int admitted = 0, finished = 0;
  bool all_converged = true;
  int max_iter_seen = 0;
  double worst_res = 0.0;
  Clock::time_point submit, first_admit, done_at;
  Clock::time_point stalled_since;  // first pass that found no memory for it
  bool stalled = false;
  bool started = false, done = false;
};

struct BatchEntry {
  std::unique_ptr<CgBatch> batch;
  hipStream_t stream = nullptr;
  GaugeField* gauge = nullptr;
  bool clover = false;
  int recon = 0, tbc = 0;
  size_t bytes = 0;
  bool in_flight = false;  // a chunk is queued on the stream and not yet harvested
  std::map<uint64_t, std::pair<uint64_t, int>> slot_req;  // slot tag -> (request id, rhs)
};

struct Server {
  qcuServerConfig cfg;
  size_t budget = 0, used = 0, peak = 0;
  std::thread worker;
  std::mutex mu;  // guards queue, requests, stop
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::deque<uint64_t> queue;  // requests with right-hand sides not yet admitted
  std::map<uint64_t, Request> reqs;
  uint64_t next_id = 1, next_tag = 1;
  std::vector<std::unique_ptr<BatchEntry>> batches;
  // (gauge, kappa*csw) -> this server's slots using that clover field; the field's own
  // CloverField::server_refs counts the slots of every server
  std::map<std::pair<GaugeField*, double>, int> clover_refs;
  std::string error;

  void run();
  bool busy();
  enum Admit { ADMIT_ALL, ADMIT_WAIT, ADMIT_NO_MEMORY };
  Admit admit_one(uint64_t id, Request& r, bool allow_alloc);
  void complete_rhs(Request& r, bool converged, int iterations, double true_residual);
  BatchEntry* find_batch(GaugeField* g, bool clover, int recon, int tbc);
  BatchEntry* new_batch(GaugeField* g, bool clover, int recon, int tbc, size_t extra);
  void release_idle_batches(const BatchEntry* keep);
  size_t clover_cost(GaugeField* g, double ck);
  void release_clover(GaugeField* g, double ck, bool counted);
  bool round();
};

size_t Server::clover_cost(GaugeField* g, double ck) {
  if (clover_refs.count({g, ck})) return 0;
  return solver_clover_bytes(cfg.precision, g->geo.vh);
}

// one slot of this server stops using the clover field (g, ck); counted: the slot had
// been added to the field's server_refs.  The field is freed only when no slot of any
// server uses it any more.
void Server::release_clover(GaugeField* g, double ck, bool counted) {
  auto it = clover_refs.find({g, ck});
  if (it == clover_refs.end()) return;
  auto c = g->clovers.find(ck);
  if (counted && c != g->clovers.end()) c->second->server_refs--;
  if (--it->second > 0) return;
  clover_refs.erase(it);
  used -= solver_clover_bytes(cfg.precision, g->geo.vh);
  if (c != g->clovers.end() && c->second->server_refs <= 0) {
    QHIP(hipStreamSynchronize(ctx().stream));
    g->clovers.erase(c);
  }
}

// give back the memory of batches with no active slot and no chunk in flight
void Server::release_idle_batches(const BatchEntry* keep) {
  for (size_t i = 0; i < batches.size();) {
    BatchEntry* b = batches[i].get();
    if (b != keep && b->batch->active() == 0 && !b->in_flight) {
      used -= b->bytes;
      QHIP(hipStreamSynchronize(b->stream));
      QHIP(hipStreamDestroy(b->stream));
      batches.erase(batches.begin() + i);
    } else {
      ++i;
    }
  }
}

BatchEntry* Server::find_batch(GaugeField* g, bool clover, int recon, int tbc) {
  for (auto& b : batches)
    if (b->gauge == g && b->clover == clover && b->recon == recon && b->tbc == tbc && b->batch->free_slots() > 0)
      return b.get();
  return nullptr;
}

// a new batch for (g, clover, recon, tbc) if it and `extra` bytes (a clover field its
// first slot needs) fit the budget, after giving back idle batches if they do not
BatchEntry* Server::new_batch(GaugeField* g, bool clover, int recon, int tbc, size_t extra) {
  size_t need = CgBatch::bytes_for(g->geo.vh, cfg.precision, cfg.max_batch);
  if (used + need + extra > budget) {
    release_idle_batches(nullptr);
    if (used + need + extra > budget) return nullptr;
  }
  std::unique_ptr<BatchEntry> e(new BatchEntry);
  QHIP(hipStreamCreate(&e->stream));  // default priority
  qcuSolveParams sp;
  qcuDefaultSolveParams(&sp);
  sp.precision = cfg.precision;
  // Chunks of about the same device time whatever the volume: streams share hardware
  // queues, so a small lattice's chunk can wait behind a large one's already queued
  // chunk.  cfg.chunk iterations up to a 65536-site half lattice, proportionally fewer
  // (at least one) above.
  sp.chunk = (int)std::max<long>(1, std::min<long>(cfg.chunk, (long)cfg.chunk * 65536 / g->geo.vh));
  sp.graph = cfg.graph;
  e->batch.reset(new CgBatch(*g, cfg.precision, recon, tbc, clover, cfg.max_batch, sp, e->stream));
  e->gauge = g;
  e->clover = clover;
  e->recon = recon;
  e->tbc = tbc;
  e->bytes = need;
  used += need;
  if (used > peak) peak = used;
  batches.push_back(std::move(e));
  return batches.back().get();
}

// move as many right-hand sides of request id into batches as fit; true when all are in
// a right-hand side of r is done (solved, or rejected at admission); the request
// completes with its last one.  Caller holds mu.
void Server::complete_rhs(Request& r, bool converged, int iterations, double true_residual) {
  r.finished++;
  r.all_converged &= converged;
  r.max_iter_seen = std::max(r.max_iter_seen, iterations);
  r.worst_res = std::max(r.worst_res, true_residual);
  if (r.finished == r.spec.nrhs) {
    r.done = true;
    r.done_at = Clock::now();
    cv_done.notify_all();
  }
}

// allow_alloc = false: only free slots of existing batches and existing clover fields
Server::Admit Server::admit_one(uint64_t id, Request& r, bool allow_alloc) {
  GaugeField* g = reinterpret_cast<GaugeField*>(r.spec.gauge);
  const bool clover = r.spec.csw != 0.0;
  const double ck = r.spec.kappa * r.spec.csw;
  while (r.admitted < r.spec.nrhs) {
    BatchEntry* be = find_batch(g, clover, r.spec.recon, r.spec.t_boundary);
    if (be && be->in_flight) return ADMIT_WAIT;  // slots open at its next chunk boundary
    const size_t cc = clover ? clover_cost(g, ck) : 0;  // a clover field not yet in use here
    if (cc && !allow_alloc) return ADMIT_NO_MEMORY;
    if (!be) {
      if (!allow_alloc) return ADMIT_NO_MEMORY;
      be = new_batch(g, clover, r.spec.recon, r.spec.t_boundary, cc);
      if (!be) return ADMIT_NO_MEMORY;
    } else if (cc && used + cc > budget) {
      release_idle_batches(be);  // idle batches of other gauges may hold the memory
      if (used + cc > budget) return ADMIT_NO_MEMORY;
    }
    // as many of the request's right-hand sides as the batch has free slots, set up
    // together
    const int m = std::min(r.spec.nrhs - r.admitted, be->batch->free_slots());
    if (clover) {
      used += cc;
      if (used > peak) peak = used;
      clover_refs[{g, ck}] += m;
    }
    std::vector<SlotSpec> specs(m);
    for (int j = 0; j < m; ++j) {
      const int k = r.admitted + j;
      SlotSpec& s = specs[j];
      s.b = r.b.empty() ? nullptr : r.b[k];
      s.seed = r.spec.seed + (uint64_t)k;
      s.x = r.x.empty() ? nullptr : r.x[k];
      s.kappa = r.spec.kappa;
      s.csw = r.spec.csw;
      s.tol = r.spec.tol;
      s.max_iter = r.spec.max_iter;
      s.tag = next_tag++;
    }
    std::vector<char> added(m, 1);
    try {
      be->batch->add_slots(specs.data(), m);  // device work: the request lock is not held here
    } catch (const Error& e) {
      if (e.code == QCU_ERROR_HIP || e.code == QCU_ERROR_OUT_OF_MEMORY) throw;
      // a right-hand side that cannot be solved (e.g. a non-finite source) fails
      // alone, the server and the other requests go on: add the group one by one
      for (int j = 0; j < m; ++j) {
        try {
          be->batch->add(specs[j]);
        } catch (const Error& e1) {
          if (e1.code == QCU_ERROR_HIP || e1.code == QCU_ERROR_OUT_OF_MEMORY) throw;
          added[j] = 0;
          if (clover) release_clover(g, ck, false);
        }
      }
    }
    int nadded = 0;
    for (int j = 0; j < m; ++j)
      if (added[j]) {
        be->slot_req[specs[j].tag] = {id, r.admitted + j};
        nadded++;
      }
    if (clover && nadded > 0) {
      auto c = g->clovers.find(ck);  // built by add_slots()
      if (c != g->clovers.end()) c->second->server_refs += nadded;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!r.started) {
      r.started = true;
      r.first_admit = Clock::now();
    }
    r.admitted += m;
    for (int j = 0; j < m; ++j)
      if (!added[j]) complete_rhs(r, false, 0, std::numeric_limits<double>::infinity());
  }
  return ADMIT_ALL;
}

bool Server::round() {
  bool progress = false;
  // 1. harvest every batch whose chunk has finished
  for (auto& b : batches) {
    if (!b->in_flight) continue;
    const hipError_t q = hipStreamQuery(b->stream);
    if (q == hipErrorNotReady) continue;
    QHIP(q);
    b->in_flight = false;
    progress = true;
    b->batch->harvest();
    for (const SlotResult& res : b->batch->take_finished()) {
      auto it = b->slot_req.find(res.tag);
      if (it == b->slot_req.end()) continue;
      uint64_t id = it->second.first;
      b->slot_req.erase(it);
      std::lock_guard<std::mutex> lk(mu);
      Request& r = reqs[id];
      if (r.spec.csw != 0.0) release_clover(b->gauge, r.spec.kappa * r.spec.csw, true);
      complete_rhs(r, res.stats.converged != 0, res.stats.iterations, res.stats.true_residual);
    }
  }
  // 2. admission in arrival order.  A request whose batch is mid-chunk keeps its place
  //    while later ones go ahead.  Once one does not fit the memory budget, nothing new
  //    is allocated until it is in, and the ones behind it may still fill free slots of
  //    existing batches -- but only for kStallGrace: after that nothing behind it is
  //    admitted, so the batches it waits on drain and give their memory back (with a
  //    steady stream they would otherwise stay busy and starve it)
  constexpr double kStallGrace = 0.5;  // seconds
  std::vector<uint64_t> ids;
  {
    std::lock_guard<std::mutex> lk(mu);
    ids.assign(queue.begin(), queue.end());
  }
  std::vector<uint64_t> admitted_all;
  bool allow_alloc = true;
  const Clock::time_point now = Clock::now();
  for (uint64_t id : ids) {
    Request* r;
    {
      std::lock_guard<std::mutex> lk(mu);
      r = &reqs[id];
    }
    const int before = r->admitted;
    const Admit a = admit_one(id, *r, allow_alloc);
    progress |= r->admitted != before;
    if (a == ADMIT_ALL) {
      admitted_all.push_back(id);
    } else if (a == ADMIT_NO_MEMORY) {
      allow_alloc = false;
      if (!r->stalled) {
        r->stalled = true;
        r->stalled_since = now;
      }
      if (std::chrono::duration<double>(now - r->stalled_since).count() > kStallGrace) break;
    }
  }
  if (!admitted_all.empty()) {
    std::lock_guard<std::mutex> lk(mu);
    for (uint64_t id : admitted_all)
      for (auto it = queue.begin(); it != queue.end(); ++it)
        if (*it == id) {
          queue.erase(it);
          break;
        }
  }
  // 3. the next chunk on every idle batch with work
  for (auto& b : batches)
    if (!b->in_flight && b->batch->active() > 0) {
      b->batch->launch_chunk();
      b->in_flight = true;
      progress = true;
    }
  return progress;
}

bool Server::busy() {
  for (auto& b : batches)
    if (b->in_flight || b->batch->active() > 0) return true;
  return false;
}

void Server::run() {
  bool progress = true;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      if (!busy()) cv_work.wait_for(lk, std::chrono::milliseconds(2), [&] { return stop || !queue.empty(); });
      if (stop) break;
    }
    // nothing finished and nothing admitted last pass: give the device a moment
    if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(50));
    try {
      std::lock_guard<std::recursive_mutex> g(ctx().mu);
      progress = round();
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> lk(mu);
      error = e.what();
      // fail every pending request rather than hang their waiters
      for (auto& kv : reqs)
        if (!kv.second.done) {
          kv.second.done = true;
          kv.second.all_converged = false;
          kv.second.done_at = Clock::now();
        }
      queue.clear();
      cv_done.notify_all();
      stop = true;
      break;
    }
  }
}

static void fill_status(const Request& r, qcuRequestStatus* st) {
  if (!st) return;
  st->done = r.done;
  st->converged = r.done && r.all_converged;
  st->iterations = r.max_iter_seen;
  st->true_residual = r.worst_res;
  st->queue_seconds = r.started ? std::chrono::duration<double>(r.first_admit - r.submit).count() : -1.0;
  st->latency_seconds = r.done ? std::chrono::duration<double>(r.done_at - r.submit).count() : -1.0;
}

}  // namespace qcu

using namespace qcu;

extern "C" {

void qcuDefaultServerConfig(qcuServerConfig* cfg) {
  if (!cfg) return;
  cfg->hbm_budget = 0;
  cfg->max_batch = 12;
  cfg->precision = QCU_PREC_SINGLE;
  cfg->chunk = 8;
  cfg->graph = 1;
}

int qcuServerCreate(qcuServer* out, const qcuServerConfig* cfg) {
  try {
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    ensure_init();
    QREQ(out, "qcuServerCreate: NULL");
    std::unique_ptr<Server> s(new Server);
    if (cfg) s->cfg = *cfg;
    else qcuDefaultServerConfig(&s->cfg);
    QREQ(s->cfg.max_batch >= 1 && s->cfg.max_batch <= kMaxRhs, "max_batch must be 1..16");
    QREQ(s->cfg.precision >= 0 && s->cfg.precision <= 2, "bad precision");
    if (s->cfg.chunk < 1) s->cfg.chunk = 1;
    s->budget = s->cfg.hbm_budget;
    if (s->budget == 0) {
      size_t fr = 0, tot = 0;
      QHIP(hipMemGetInfo(&fr, &tot));
      s->budget = fr / 4 * 3;
    }
    Server* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    *out = reinterpret_cast<qcuServer>(s.release());
    return QCU_SUCCESS;
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    return e.code;
  }
}

int qcuServerSubmit(qcuServer h, const qcuRequest* req, uint64_t* id) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s || !req || !id || !req->gauge) {
    set_last_error(QCU_ERROR_INVALID_VALUE, "qcuServerSubmit: NULL argument");
    return QCU_ERROR_INVALID_VALUE;
  }
  if (req->nrhs < 1 || req->nrhs > kMaxRhs || !(req->tol > 0.0) || (req->recon != 12 && req->recon != 18) ||
      (req->t_boundary != 1 && req->t_boundary != -1)) {
    set_last_error(QCU_ERROR_INVALID_VALUE, "qcuServerSubmit: bad request parameters");
    return QCU_ERROR_INVALID_VALUE;
  }
  GaugeField* g = reinterpret_cast<GaugeField*>(req->gauge);
  // a batch of the request's gauge, plus the clover field its first slot builds
  size_t need = CgBatch::bytes_for(g->geo.vh, s->cfg.precision, s->cfg.max_batch);
  if (req->csw != 0.0) need += solver_clover_bytes(s->cfg.precision, g->geo.vh);
  if (need > s->budget) {
    set_last_error(QCU_ERROR_ADMISSION, "request needs more device memory than the server's budget");
    return QCU_ERROR_ADMISSION;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) {
    set_last_error(QCU_ERROR_STATE, "server stopped: " + s->error);
    return QCU_ERROR_STATE;
  }
  uint64_t i = s->next_id++;
  Request& r = s->reqs[i];
  r.spec = *req;
  if (req->x) r.x.assign(req->x, req->x + req->nrhs);
  if (req->b) r.b.assign(req->b, req->b + req->nrhs);
  r.spec.x = nullptr;
  r.spec.b = nullptr;
  r.submit = Clock::now();
  s->queue.push_back(i);
  s->cv_work.notify_one();
  *id = i;
  return QCU_SUCCESS;
}

int qcuServerPoll(qcuServer h, uint64_t id, qcuRequestStatus* st) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return QCU_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) {
    set_last_error(QCU_ERROR_INVALID_VALUE, "unknown request id");
    return QCU_ERROR_INVALID_VALUE;
  }
  fill_status(it->second, st);
  return QCU_SUCCESS;
}

int qcuServerWait(qcuServer h, uint64_t id, qcuRequestStatus* st) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return QCU_ERROR_INVALID_VALUE;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) {
    set_last_error(QCU_ERROR_INVALID_VALUE, "unknown request id");
    return QCU_ERROR_INVALID_VALUE;
  }
  s->cv_done.wait(lk, [&] { return it->second.done; });
  fill_status(it->second, st);
  if (!s->error.empty()) {
    set_last_error(QCU_ERROR_STATE, "server failed: " + s->error);
    return QCU_ERROR_STATE;
  }
  return it->second.all_converged ? QCU_SUCCESS : QCU_ERROR_NOT_CONVERGED;
}

int qcuServerDrain(qcuServer h) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return QCU_ERROR_INVALID_VALUE;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] {
    for (auto& kv : s->reqs)
      if (!kv.second.done) return false;
    return true;
  });
  if (!s->error.empty()) {
    set_last_error(QCU_ERROR_STATE, "server failed: " + s->error);
    return QCU_ERROR_STATE;
  }
  return QCU_SUCCESS;
}

int qcuServerUsage(qcuServer h, size_t* bytes, size_t* peak, int* nb) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return QCU_ERROR_INVALID_VALUE;
  std::lock_guard<std::recursive_mutex> g(ctx().mu);
  if (bytes) *bytes = s->used;
  if (peak) *peak = s->peak;
  if (nb) *nb = (int)s->batches.size();
  return QCU_SUCCESS;
}

int qcuServerDestroy(qcuServer h) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return QCU_ERROR_INVALID_VALUE;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  try {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    for (auto& b : s->batches) {
      QHIP(hipStreamSynchronize(b->stream));
      b->batch.reset();
      QHIP(hipStreamDestroy(b->stream));
    }
    // this server's slots stop using their clover fields; a field another server
    // still uses stays
    for (auto& kv : s->clover_refs) {
      auto c = kv.first.first->clovers.find(kv.first.second);
      if (c == kv.first.first->clovers.end()) continue;
      c->second->server_refs -= kv.second;
      if (c->second->server_refs <= 0) kv.first.first->clovers.erase(c);
    }
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    delete s;
    return e.code;
  }
  delete s;
  return QCU_SUCCESS;
}

}  // extern "C"
