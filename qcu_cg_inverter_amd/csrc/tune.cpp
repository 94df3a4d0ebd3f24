// tune.cpp -- autotuner of the hopping kernel's launch shape.
//
// For one (precision, link storage, half volume, right-hand sides) it times every
// candidate shape -- block size, wave groups per block sharing sites, right-hand
// sides per wave, group order, XCD-contiguous block order, walk tile, and the kernel
// that stages a site chunk's links in LDS for every right-hand side -- on scratch
// fields with hipEvents and the plain epilogue, keeps the fastest, and re-times the
// eight fastest, and the two fastest that take every right-hand side in one block per
// site range, with each epilogue that reads x or a clover field (those add loads that
// can favour another shape).  Results live in the context's table and, when
// QCU_TUNE_CACHE names a file, are appended there and reloaded by later processes.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "qcu_host.h"

namespace qcu {

static const char* cache_file() {
  const char* p = getenv("QCU_TUNE_CACHE");
  return (p && *p) ? p : nullptr;
}

// The checks a launch shape must pass (qcuSetHopShape and the cache file): null when the
// shape is valid for `precision`, else the reason.  Shapes outside them would divide by
// zero in the launcher or launch blocks the kernels do not expect.
static const char* shape_error(const Tune& t, int precision) {
  if (!(t.lds >= 0 && t.lds <= 2)) return "lds must be 0, 1 or 2";
  if (t.lds && !(t.block == 64 || t.block == 128 || t.block == 256))
    return "the LDS kernel takes 1, 2 or 4 waves per block";
  if (!(t.block == 64 || t.block == 128 || t.block == 192 || t.block == 256)) return "block must be 64, 128, 192 or 256";
  if (!(t.rpw == 1 || t.rpw == 2 || t.rpw == 4)) return "right-hand sides per wave must be 1, 2 or 4";
  if (!(t.spt >= 1 && t.spt <= 64)) return "site blocks per block must be 1..64";
  if (!(t.cached == 0 || t.cached == 1)) return "load policy must be 0 (streaming) or 1 (cached)";
  if (!(t.rb >= 1 && t.block % (64 * t.rb) == 0)) return "rb must divide block into 64-thread groups";
  if (!(t.rpt >= 1 && t.rpt <= 4 && (precision != F64 || t.rpt == 1))) return "rpt must be 1..4 (1 for FP64)";
  if (!((t.xcd == 0 || t.xcd == 1 || (t.lds && (t.xcd == 2 || t.xcd == 4))) && (t.gm == 0 || t.gm == 1)))
    return "xcd is 0 or 1 (2 or 4: XCD column groups, LDS kernel only); group order is 0 or 1";
  if (!(t.tz >= 0 && t.tt >= 0 && t.ty >= 0 && (t.tz == 0) == (t.tt == 0)))
    return "tile extents must be non-negative, tz and tt both 0 or both positive";
  return nullptr;
}

static bool cache_loaded = false;

void load_tune_cache() {
  if (cache_loaded) return;
  cache_loaded = true;
  const char* path = cache_file();
  if (!path) return;
  FILE* f = fopen(path, "r");
  if (!f) return;
  char line[256], key[128];
  while (fgets(line, sizeof line, f)) {
    Tune t;
    double ms;
    int prec = -1;
    // lines that do not parse, or whose shape qcuSetHopShape would refuse, are skipped
    if (sscanf(line, "%127s %d %d %d %d %d %d %d %d %d %d %d %d %lf", key, &t.block, &t.rb, &t.rpt, &t.xcd, &t.tz,
               &t.tt, &t.gm, &t.ty, &t.rpw, &t.spt, &t.cached, &t.lds, &ms) == 14 &&
        sscanf(key, "hop:%d:", &prec) == 1 && prec >= 0 && prec <= 2 && !shape_error(t, prec))
      ctx().tune[key] = t;
  }
  fclose(f);
}

static void append_cache(const std::string& key, const Tune& t, double ms) {
  const char* path = cache_file();
  if (!path) return;
  FILE* f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d %d %d %.6f\n", key.c_str(), t.block, t.rb, t.rpt, t.xcd, t.tz, t.tt,
          t.gm, t.ty, t.rpw, t.spt, t.cached, t.lds, ms);
  fclose(f);
}

// Several right-hand sides per thread (qcuSetHopShape) spill in FP64 and drop FP32 /
// FP16 to one wave per SIMD, so the search keeps one per thread and varies how many
// share a wave (rpw) and a block (rb).
static std::vector<Tune> candidates(const Geom& geo, int prec, int nrhs) {
  (void)prec;
  std::vector<Tune> out;
  const int blocks[] = {64, 128, 192, 256};
  // (ty, tz, tt) walk tiles; ty = 0 keeps whole xy-planes
  const int tiles[][3] = {{0, 0, 0}, {0, 4, 4}, {0, 8, 8}, {8, 4, 4}, {4, 2, 4}};
  for (auto& tl : tiles)
    for (int xcd = 0; xcd <= 1; ++xcd)
      for (int block : blocks)
        for (int rb = 1; rb <= 4; ++rb) {
          if (block % (64 * rb) != 0) continue;
          for (int rpw = 1; rpw <= 4; rpw *= 2) {
            const int slots = rb * rpw;  // right-hand sides one block serves
            if (rpw > nrhs || slots > nrhs + rpw - 1) continue;  // every wave of a full block has work
            const bool groups = slots < nrhs;  // more than one block per site range
            for (int gm = 0; gm <= (groups ? 1 : 0); ++gm) {
              Tune t;
              t.block = block;
              t.rb = rb;
              t.rpt = 1;
              t.rpw = rpw;
              t.xcd = xcd;
              t.ty = tl[0];
              t.tz = tl[1];
              t.tt = tl[2];
              t.gm = gm;
              out.push_back(t);
            }
          }
        }
  // the LDS-staged kernel: every right-hand side in one block of 1, 2 or 4 waves, also
  // with tiles of all of z by 4 or 8 t-planes (on 32^3x64 with 12 FP32 right-hand
  // sides, XCD round-robin with tiles of 32 z by 8 t was the fastest of ten shapes
  // tried: 784 us against 847 us for 4 x 4 tiles)
  std::vector<std::array<int, 3>> ltiles;
  for (auto& tl : tiles) ltiles.push_back({tl[0], tl[1], tl[2]});
  ltiles.push_back({0, geo.L[2], 16});
  ltiles.push_back({0, geo.L[2], 8});
  ltiles.push_back({0, geo.L[2], 4});
  // block orders: none (round-robin over the XCDs), contiguous ranges, and XCD column
  // groups 2 and 4 chunks wide
  const int xorders[] = {0, 1, 2, 4};
  if (nrhs > 1)
    for (auto& tl : ltiles)
      for (int xcd : xorders)
        for (int nw = 1; nw <= 4; nw *= 2) {
          if (nw > nrhs) continue;
          Tune t;
          t.block = 64 * nw;
          t.rb = 1;
          t.rpt = 1;
          t.rpw = 1;
          t.xcd = xcd;
          t.ty = tl[0];
          t.tz = tl[1];
          t.tt = tl[2];
          t.lds = 1;
          out.push_back(t);
        }
  return out;
}


Tune autotune_hop(GaugeField& g, int prec, int recon, int nrhs, double* best_ms) {
  load_tune_cache();
  const long vh = g.geo.vh;
  const size_t fb = spinor_bytes(prec, vh);
  DevBuf* src = ctx().buf("tune_src", (size_t)vh * 192 * nrhs);
  DevBuf* a = ctx().buf("tune_a", fb * nrhs);
  DevBuf* b = ctx().buf("tune_b", fb * nrhs);
  launch_source_gauss(src->p, vh * 12 * nrhs, 777, stream());
  launch_spinor_from_aos(prec, src->p, vh * 192, vec(a->p, (long)fb), vh, nrhs, stream());
  // x of the xpay epilogues: a field of its own, as in a solve (not the hop's input,
  // whose loads would hit the caches)
  DevBuf* xb = ctx().buf("tune_x", fb * nrhs);
  launch_spinor_from_aos(prec, src->p, vh * 192, vec(xb->p, (long)fb), vh, nrhs, stream());
  // a clover-sized scratch field for the clover epilogues (its values do not matter)
  DevBuf* cl = ctx().buf("tune_clov", clover_bytes(prec, vh) / 2);
  QHIP(hipMemsetAsync(cl->p, 0, clover_bytes(prec, vh) / 2, stream()));
  const void* gauge = g.get(prec, recon);
  HopArgs args;
  memset(&args, 0, sizeof args);
  args.in = a->p;
  args.out = b->p;
  args.x = xb->p;
  args.in_stride = args.out_stride = args.x_stride = (long)fb;
  args.gauge = gauge;
  for (int r = 0; r < nrhs; ++r) {
    args.coef[r] = 1.0;
    args.clov[r] = cl->p;
  }
  args.clov_shared = 1;
  args.geo = g.geo;
  args.tbc = -1.0;
  args.nrhs = nrhs;
  hipEvent_t e0, e1;
  QHIP(hipEventCreate(&e0));
  QHIP(hipEventCreate(&e1));
  const int iters = vh * nrhs > (1 << 22) ? 10 : 30;
  auto time_shape = [&](const Tune& t, int epi) -> double {
    args.xcd_remap = t.xcd;
    args.gmajor = t.gm;
    args.rpw = t.rpw;
    args.spt = t.spt;
    args.cached = t.cached;
    args.lds = t.lds;
    args.walk = make_walk(g.geo, t.ty, t.tz, t.tt);
    for (int w = 0; w < 2; ++w) launch_hop(prec, recon, epi, args, t.block, t.rb, t.rpt, stream());
    QHIP(hipEventRecord(e0, stream()));
    for (int i = 0; i < iters; ++i) {
      args.parity = i & 1;
      launch_hop(prec, recon, epi, args, t.block, t.rb, t.rpt, stream());
    }
    QHIP(hipEventRecord(e1, stream()));
    QHIP(hipEventSynchronize(e1));
    QHIP(hipGetLastError());
    float ms = 0.f;
    QHIP(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms / iters;
  };
  // every candidate with the plain epilogue ...
  std::vector<std::pair<double, Tune>> timed;
  for (const Tune& t : candidates(g.geo, prec, nrhs)) {
    if (t.tz && !make_walk(g.geo, t.ty, t.tz, t.tt).tile_z) continue;  // tile does not fit this lattice
    timed.push_back({time_shape(t, EPI_PLAIN), t});
  }
  auto by_time = [](const std::pair<double, Tune>& x, const std::pair<double, Tune>& y) { return x.first < y.first; };
  std::stable_sort(timed.begin(), timed.end(), by_time);
  // ... then the fastest few walking 2 or 4 site blocks per block (sites per thread),
  // and with ordinary instead of streaming loads
  const size_t top = timed.size() < 8 ? timed.size() : 8;
  for (size_t k = 0; k < top; ++k)
    for (int v = 0; v < 3; ++v) {
      Tune t = timed[k].second;
      if (v < 2) t.spt = 2 << v;
      else t.cached = 1;
      timed.push_back({time_shape(t, EPI_PLAIN), t});
    }
  std::stable_sort(timed.begin(), timed.end(), by_time);
  const Tune best = timed.front().second;
  const double bms = timed.front().first;
  std::string key = tune_key(prec, recon, vh, nrhs, EPI_PLAIN);
  ctx().tune[key] = best;
  append_cache(key, best, bms);
  // ... and the fastest few again with each epilogue that reads x or the clover, plus
  // the two fastest shapes that serve every right-hand side from one block per site
  // range (links and clover read once per launch), streaming and cached: the epilogue
  // loads can favour them although they lose with the plain one
  std::vector<Tune> epi_shapes;
  auto same = [](const Tune& x, const Tune& y) {
    return x.block == y.block && x.rb == y.rb && x.rpt == y.rpt && x.xcd == y.xcd && x.tz == y.tz &&
           x.tt == y.tt && x.gm == y.gm && x.ty == y.ty && x.rpw == y.rpw && x.spt == y.spt && x.cached == y.cached &&
           x.lds == y.lds;
  };
  auto add_shape = [&](const Tune& t) {
    for (const Tune& u : epi_shapes)
      if (same(t, u)) return;
    epi_shapes.push_back(t);
  };
  for (size_t k = 0; k < top; ++k) add_shape(timed[k].second);
  // the LDS kernel's shapes also with the clover staged in LDS (lds = 2; the clover
  // epilogues use it when every right-hand side shares one clover field)
  for (size_t k = 0; k < top; ++k)
    if (timed[k].second.lds == 1) {
      Tune t = timed[k].second;
      t.lds = 2;
      add_shape(t);
    }
  int one_group = 0;
  for (size_t k = 0; k < timed.size() && one_group < 2; ++k) {
    Tune t = timed[k].second;
    if ((!t.lds && t.rb * t.rpw * t.rpt < nrhs) || t.spt != 1) continue;
    t.cached = 0;
    add_shape(t);
    t.cached = 1;
    add_shape(t);
    ++one_group;
  }
  for (int epi = EPI_XPAY; epi <= EPI_CLOV_XPAY; ++epi) {
    Tune be = best;
    double bt = 1e30;
    for (const Tune& t : epi_shapes) {
      double ms = time_shape(t, epi);
      if (ms < bt) {
        bt = ms;
        be = t;
      }
    }
    std::string ke = tune_key(prec, recon, vh, nrhs, epi);
    ctx().tune[ke] = be;
    append_cache(ke, be, bt);
  }
  QHIP(hipEventDestroy(e0));
  QHIP(hipEventDestroy(e1));
  if (best_ms) *best_ms = bms;
  return best;
}

// tune once per key when QCU_TUNE=1 (used by the solver before capturing graphs)
void maybe_autotune(GaugeField& g, int prec, int recon, int nrhs) {
  load_tune_cache();
  const char* e = getenv("QCU_TUNE");
  if (!(e && strcmp(e, "0") != 0)) return;
  if (ctx().tune.count(tune_key(prec, recon, g.geo.vh, nrhs, EPI_PLAIN))) return;
  autotune_hop(g, prec, recon, nrhs, nullptr);
}

}  // namespace qcu

extern "C" int qcuTuneHop(qcuGauge h, int precision, int recon, int nrhs, int* shape, double* ms) {
  using namespace qcu;
  try {
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    ensure_init();
    QREQ(h, "qcuTuneHop: NULL gauge");
    QREQ(precision >= 0 && precision <= 2, "bad precision");
    QREQ(recon == 12 || recon == 18, "recon must be 12 or 18");
    QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range");
    Tune t = autotune_hop(*reinterpret_cast<GaugeField*>(h), precision, recon, nrhs, ms);
    if (shape) {
      shape[0] = t.block;
      shape[1] = t.rb;
      shape[2] = t.rpt;
      shape[3] = t.xcd;
      shape[4] = t.tz;
      shape[5] = t.tt;
      shape[6] = t.gm;
      shape[7] = t.ty;
      shape[8] = t.rpw;
      shape[9] = t.spt;
      shape[10] = t.cached;
      shape[11] = t.lds;
    }
    return QCU_SUCCESS;
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    return e.code;
  }
}

extern "C" int qcuGetHopShape(qcuGauge h, int precision, int recon, int nrhs, int* shape) {
  using namespace qcu;
  try {
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    ensure_init();
    QREQ(h && shape, "qcuGetHopShape: NULL argument");
    QREQ(precision >= 0 && precision <= 2, "bad precision");
    QREQ(recon == 12 || recon == 18, "recon must be 12 or 18");
    QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range");
    const Tune t = tuned_hop(precision, recon, reinterpret_cast<GaugeField*>(h)->geo.vh, nrhs);
    const int v[12] = {t.block, t.rb, t.rpt, t.xcd, t.tz, t.tt, t.gm, t.ty, t.rpw, t.spt, t.cached, t.lds};
    for (int k = 0; k < 12; ++k) shape[k] = v[k];
    return QCU_SUCCESS;
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    return e.code;
  }
}

extern "C" int qcuSetHopShape(qcuGauge h, int precision, int recon, int nrhs, const int* shape) {
  using namespace qcu;
  try {
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    ensure_init();
    QREQ(h && shape, "qcuSetHopShape: NULL argument");
    QREQ(precision >= 0 && precision <= 2, "bad precision");
    QREQ(recon == 12 || recon == 18, "recon must be 12 or 18");
    QREQ(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs out of range");
    Tune t;
    t.block = shape[0];
    t.rb = shape[1];
    t.rpt = shape[2];
    t.xcd = shape[3];
    t.tz = shape[4];
    t.tt = shape[5];
    t.gm = shape[6];
    t.ty = shape[7];
    t.rpw = shape[8];
    t.spt = shape[9];
    t.cached = shape[10];
    t.lds = shape[11];
    if (const char* why = shape_error(t, precision)) throw Error(QCU_ERROR_INVALID_VALUE, why);
    GaugeField& g = *reinterpret_cast<GaugeField*>(h);
    set_tuned_hop(precision, recon, g.geo.vh, nrhs, t);
    return QCU_SUCCESS;
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    return e.code;
  }
}
