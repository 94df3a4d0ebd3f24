#!/usr/bin/env python3
"""Benchmark of the qcu hot path on one MI355X.

Workload (step): one even-odd hopping-term application D_eo / D_oe (FP64, 18-real
links, antiperiodic t) on the 16^3x32 lattice of the Wilson CG configuration
(seeded near-unit SU(3) gauge, eps = 0.2).  `value` is the algorithmic Dslash
bandwidth: 2880 B per output half-site (8 links x 144 B + 8 neighbour spinors x
192 B + 192 B written) x Vh / time per application.

The same JSON line also carries:
  roofline      the hop kernel against the 8 TB/s HBM3E peak, HIP-event timed
  cpu_baseline  the FP64 oracle hop timed on the host cores (bounded sample)
  cg            FP64 vs mixed-precision (FP32 inner, FP64 reliable updates) CG to a
                true residual of 1e-10 on the same lattice, kappa = 0.124, one
                Gaussian source
  dslash_large  the same kernel on 24^3x48 and 32^3x64 (working set above the
                256 MB Infinity Cache), FP64 / FP32 / FP16 and reconstruct-12
  copy          device-to-device copy bandwidth for scale
  dslash_16x32_median  the headline hop as the median of 100 timed windows
  dropin_16x32  host-synchronous dslashQcu / fullDslashQcu calls through the drop-in
  config1_..., dslash_12rhs, config3_..., config4_..., config5_serving
                the other configurations: solves, the 12-RHS hop, serving

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import hashlib
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Dslash GB/s & %HBM peak; CG time-to-1e-10 mixed vs FP64; p50/p99 solve ms"
HBM_PEAK_GBS = 8000.0
BYTES_PER_SITE = {("f64", 18): 2880, ("f64", 12): 2496, ("f32", 18): 1440, ("f32", 12): 1248,
                  ("f16", 18): 756, ("f16", 12): 660}
PREC = {"f64": 0, "f32": 1, "f16": 2}
L_MAIN = (16, 16, 16, 32)
LINK_BYTES = {("f64", 18): 144, ("f64", 12): 96, ("f32", 18): 72, ("f32", 12): 48, ("f16", 18): 36, ("f16", 12): 24}
SPINOR_BYTES = {"f64": 192, "f32": 96, "f16": 52}


def hop_bytes(prec, recon, nrhs=1):
    """Algorithmic bytes per output half-site of one launch over nrhs right-hand
    sides: 8 links shared by all, then per right-hand side 8 neighbour spinors read
    and one written."""
    return 8 * LINK_BYTES[(prec, recon)] + nrhs * 9 * SPINOR_BYTES[prec]


def compulsory_bytes(prec, recon, nrhs=1):
    """HBM bytes per output half-site that one launch cannot avoid: each of the 8
    links once, and per right-hand side one read of the input spinor and one write of
    the output (the 8 neighbour reads of a spinor can come from cache after the
    first).  FP64 recon-18: 8 x 144 + 2 x 192 = 1536 B."""
    return 8 * LINK_BYTES[(prec, recon)] + nrhs * 2 * SPINOR_BYTES[prec]


# sources that fix the hop kernel's code: a PMC traffic figure measured on other
# sources is stale and is not reported
HOP_SOURCES = ("k_hop.inc", "k_hop.hip", "qcu_fields.h", "qcu_common.h", "qcu_kernels.h")


def hop_source_hash():
    h = hashlib.sha256()
    for n in HOP_SOURCES:
        with open(os.path.join(ROOT, "qcu_cg_inverter_amd", "csrc", n), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def roofline(kernel, prec, recon, vh, ms, traffic_name, nrhs=1):
    """The hop against the 8 TB/s HBM peak.  `achieved` counts the compulsory bytes of
    one launch (compulsory_bytes x Vh), so a correct kernel cannot exceed the peak;
    `achieved_algorithmic` counts the 8 neighbour reads of every spinor as well
    (hop_bytes, the per-site figure the metric is quoted in), which the caches serve
    after the first.  `traffic` is the HBM bytes per launch read off rocprofv3 PMC
    counters (FETCH_SIZE and WRITE_SIZE, scaled by what they read for a copy of known
    size) for these kernel sources, or None."""
    comp = compulsory_bytes(prec, recon, nrhs) * vh
    alg = hop_bytes(prec, recon, nrhs) * vh
    ach = comp / (ms * 1e-3) / 1e9
    t = measured_traffic(traffic_name)
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": t, "kernel": kernel,
            "avg_launch_us": round(ms * 1e3, 3), "bytes_per_launch": comp,
            "bytes_basis": f"compulsory: {compulsory_bytes(prec, recon, nrhs)} B per half site",
            "traffic_GBps": round(t / (ms * 1e-3) / 1e9, 1) if t else None,
            "achieved_algorithmic": round(alg / (ms * 1e-3) / 1e9, 1), "algorithmic_bytes_per_launch": alg,
            "frac_algorithmic": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_sources_sha16": hop_source_hash()}


def timed(fn):
    t0 = time.perf_counter()
    r = fn()
    return r, time.perf_counter() - t0


def multi_rhs_section(q):
    """12-RHS hop (links loaded once per site for all) vs 12 single launches, 32^3x64."""
    L = (32, 32, 32, 64)
    g = q.Gauge.random(L, seed=3, eps=0.25)
    vh = q.half_volume(L)
    res = {}
    for prec, recon in (("f64", 18), ("f64", 12), ("f32", 12), ("f16", 12)):
        shape, _ = q.tune_hop(g, PREC[prec], recon, 12)
        q.tune_hop(g, PREC[prec], recon, 1)
        ms12 = q.bench_hop(g, precision=PREC[prec], recon=recon, nrhs=12, iters=20)
        ms1 = q.bench_hop(g, precision=PREC[prec], recon=recon, nrhs=1, iters=50)
        gbs = hop_bytes(prec, recon, 12) * vh / (ms12 * 1e-3) / 1e9
        res[f"{prec}_r{recon}"] = {"us_12rhs": round(ms12 * 1e3, 1), "us_1rhs": round(ms1 * 1e3, 1),
                                   "speedup_vs_12_single": round(12 * ms1 / ms12, 3),
                                   "GBps_algorithmic": round(gbs, 1), "bytes_per_site": hop_bytes(prec, recon, 12),
                                   "shape": shape}
        log(f"12-rhs hop {prec} r{recon}: {ms12*1e3:.0f} us (12 x single {12*ms1*1e3:.0f} us), {gbs:.0f} GB/s")
    g.destroy()
    return res


def config2_section(q, g=None):
    """16^3x32 Wilson (kappa = 0.124, eps = 0.2, antiperiodic t), one Gaussian source,
    true residual 1e-10: plain FP64 CG against FP32- and FP16-inner CG with FP64
    reliable updates."""
    own = g is None
    if own:
        g = q.Gauge.random(L_MAIN, seed=1, eps=0.2)
    b = q.source_gaussian(L_MAIN, seed=2)
    op = q.Operator(0.124, 0.0, -1, 12)
    cg = {}
    for name, prec in (("fp64", 0), ("mixed_fp32", 1), ("mixed_fp16", 2)):
        sp = q.default_solve_params(tol=1e-10, precision=prec)
        if prec == 0:
            op64 = q.Operator(0.124, 0.0, -1, 18)
            q.solve(g, op64, b, sp)  # warm (graph capture paths, buffers)
            x, st = q.solve(g, op64, b, sp)
        else:
            q.solve(g, op, b, sp)
            x, st = q.solve(g, op, b, sp)
        cg[name] = {"seconds": round(st.seconds, 5), "iterations": st.iterations,
                    "reliable_updates": st.reliable_updates, "true_residual": st.true_residual,
                    "converged": bool(st.converged)}
        log(f"cg {name}: {st.seconds*1e3:.2f} ms, {st.iterations} it, true res {st.true_residual:.2e}")
    cg["speedup_mixed_fp32_vs_fp64"] = round(cg["fp64"]["seconds"] / cg["mixed_fp32"]["seconds"], 3)
    cg["speedup_mixed_fp16_vs_fp64"] = round(cg["fp64"]["seconds"] / cg["mixed_fp16"]["seconds"], 3)
    cg["problem"] = "16^3x32 Wilson, kappa=0.124, eps=0.2, antiperiodic t, 1 Gaussian source, tol 1e-10"
    if own:
        g.destroy()
    return cg


def config3_section(q):
    """24^3x48 Wilson-clover c_sw=1, kappa=0.122, eps=0.25, 1 Gaussian source, tol 1e-10:
    FP16 (+ per-site norm) inner with FP64 reliable updates vs plain FP64."""
    L = (24, 24, 24, 48)
    g = q.Gauge.random(L, seed=5, eps=0.25)
    op = q.Operator(0.122, 1.0, -1, 12)
    b = q.source_gaussian_device(L, seed=6)
    x = q.DeviceBuffer(2 * q.half_volume(L) * 192)
    out = {}
    for name, prec, graph in (("fp64_plain", 0, 0), ("mixed_fp16", 2, 1), ("mixed_fp32", 1, 1)):
        sp = q.default_solve_params(tol=1e-10, precision=prec, graph=graph)
        q.solve_batch_device(g, op, [x], [b], sp)  # warm
        st = q.solve_batch_device(g, op, [x], [b], sp)[0]
        out[name] = {"seconds": round(st.seconds, 4), "iterations": st.iterations,
                     "reliable_updates": st.reliable_updates, "true_residual": st.true_residual}
        log(f"config3 {name}: {st.seconds*1e3:.1f} ms, {st.iterations} it, true res {st.true_residual:.2e}")
    out["speedup_fp16_vs_fp64"] = round(out["fp64_plain"]["seconds"] / out["mixed_fp16"]["seconds"], 3)
    g.destroy()
    return out


def serving_section(q, n_requests=2000, seed=2024, precision=1):
    """Poisson stream of solve requests over 8 resident gauges (2 per volume in
    {8^4, 16^3x32, 24^3x48, 32^3x64}); each request draws kappa ~ U[0.11, 0.124],
    Wilson or clover (c_sw = 1.0), 1-12 Gaussian right-hand sides and tol in
    {1e-8, 1e-10}.  Capacity is measured with a closed burst, arrivals then run at
    0.8x capacity through the continuous-batching server (FP16 + per-site norm inner
    loop by default, or FP32; FP64 reliable updates and true residuals).  The baseline is the plain FP64 solve, one right-hand side at a time:
    one solve per request class is timed and the whole stream is replayed through a
    FIFO queue with the same arrival times; that model is checked against a measured
    one-at-a-time FP64 run of a seeded subsample of 40 of the stream's requests, each
    at its own kappa, and its error is reported."""
    import numpy as np

    rng = np.random.default_rng(seed)
    vols = [(8, 8, 8, 8), (16, 16, 16, 32), (24, 24, 24, 48), (32, 32, 32, 64)]
    gauges = [q.Gauge.random(vols[i // 2], seed=900 + i, eps=0.2 + 0.05 * (i % 2)) for i in range(8)]

    def draw(n):
        out = []
        for _ in range(n):
            gi = int(rng.integers(8))
            out.append({"g": gi, "kappa": float(rng.uniform(0.11, 0.124)), "csw": 1.0 if rng.random() < 0.5 else 0.0,
                        "nrhs": int(rng.integers(1, 13)), "tol": 1e-8 if rng.random() < 0.5 else 1e-10,
                        "seed": int(rng.integers(1 << 40))})
        return out

    def submit(server, r):
        return server.submit(gauges[r["g"]], r["kappa"], r["csw"], r["tol"], r["nrhs"], seed=r["seed"])

    # warm-up: build packed gauges, tune the common shapes (tuned shapes are kept per
    # volume, so one gauge of each volume is enough)
    for gi in range(0, 8, 2):
        t1 = time.perf_counter()
        q.tune_hop(gauges[gi], precision, 12, 12)
        q.tune_hop(gauges[gi], precision, 12, 1)
        log(f"serving warm-up: tuned {'x'.join(map(str, vols[gi // 2]))} in {time.perf_counter() - t1:.1f} s")
    server = q.Server(precision=precision, max_batch=12, chunk=8)
    for r in draw(8):
        submit(server, r)
    server.drain()
    # capacity: closed burst
    burst = draw(48)
    t0 = time.perf_counter()
    for r in burst:
        submit(server, r)
    server.drain()
    tb = time.perf_counter() - t0
    cap = len(burst) / tb
    rate = 0.8 * cap
    # open Poisson stream at 0.8x capacity
    reqs = draw(n_requests)
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    arrivals = np.cumsum(gaps)
    ids = []
    t0 = time.perf_counter()
    log(f"serving: capacity {cap:.1f} requests/s from a burst of {len(burst)}; "
        f"{n_requests} Poisson arrivals at {rate:.1f}/s")
    for k, (r, ta) in enumerate(zip(reqs, arrivals)):
        d = ta - (time.perf_counter() - t0)
        if d > 0:
            time.sleep(d)
        ids.append(submit(server, r))
        if (k + 1) % 60 == 0:
            log(f"serving: {k + 1} submitted at {time.perf_counter() - t0:.1f} s")
    last, pending = time.perf_counter(), list(ids)
    while pending:  # drain, with a progress line every 10 s
        pending = [i for i in pending if not server.poll(i).done]
        if pending and time.perf_counter() - last > 10.0:
            log(f"serving: {len(ids) - len(pending)} of {len(ids)} done at {time.perf_counter() - t0:.1f} s")
            last = time.perf_counter()
        if pending:
            time.sleep(0.2)
    server.drain()
    tend = time.perf_counter() - t0
    sts = [server.poll(i) for i in ids]
    usage = server.usage()
    server.destroy()
    lat = np.array([s.latency_seconds for s in sts]) * 1e3
    nrhs_total = sum(r["nrhs"] for r in reqs)
    out = {"requests": n_requests, "rhs": nrhs_total, "inner_precision": ("fp64", "fp32", "fp16")[precision],
           "arrival_rate_per_s": round(rate, 2),
           "capacity_req_per_s": round(cap, 2), "wall_s": round(tend, 2),
           "solves_per_s": round(nrhs_total / tend, 2), "requests_per_s": round(n_requests / tend, 2),
           "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
           "all_converged": all(s.converged for s in sts),
           "max_true_residual_over_tol": max(s.true_residual / r["tol"] for s, r in zip(sts, reqs)),
           "peak_device_bytes": usage["peak_bytes"]}
    by_vol = {}
    for vi, vol in enumerate(vols):
        lv = np.array([l for l, r in zip(lat, reqs) if r["g"] // 2 == vi])
        if lv.size:
            by_vol["x".join(map(str, vol))] = {"requests": int(lv.size), "p50_ms": round(float(np.percentile(lv, 50)), 1),
                                               "p99_ms": round(float(np.percentile(lv, 99)), 1)}
    out["latency_by_volume"] = by_vol
    # baseline: plain FP64, one right-hand side at a time.  One solve is timed per
    # class (gauge, Wilson or clover, tol) at kappa = 0.117, then the whole stream is
    # replayed through a FIFO queue with the same arrival times, each request taking
    # nrhs times its class's solve time
    log(f"serving: stream drained at {tend:.1f} s; timing the plain FP64 baseline per class")
    sp = q.default_solve_params(tol=1e-10, precision=0, graph=0)
    cls_time = {}
    for gi in range(8):
        L = vols[gi // 2]
        x = q.DeviceBuffer(2 * q.half_volume(L) * 192)
        b = q.source_gaussian_device(L, 4242 + gi)
        for csw in (0.0, 1.0):
            for tol in (1e-8, 1e-10):
                sp.tol = tol
                op = q.Operator(0.117, csw, -1, 12)
                q.solve_batch_device(gauges[gi], op, [x], [b], sp)  # clover and buffers built
                t1 = time.perf_counter()
                q.solve_batch_device(gauges[gi], op, [x], [b], sp)
                cls_time[(gi, csw, tol)] = time.perf_counter() - t1
        b.free()
        x.free()
    service = [r["nrhs"] * cls_time[(r["g"], r["csw"], r["tol"])] for r in reqs]
    busy_until, blat = 0.0, []
    for ta, s in zip(arrivals, service):
        start = max(ta, busy_until)
        busy_until = start + s
        blat.append(busy_until - ta)
    blat = np.array(blat) * 1e3
    out["baseline_fp64_one_at_a_time"] = {
        "solves_per_s": round(nrhs_total / sum(service), 2),
        "p50_ms": round(float(np.percentile(blat, 50)), 2), "p99_ms": round(float(np.percentile(blat, 99)), 2),
        "note": "FIFO replay of the same arrivals; per-request service = nrhs x the measured plain FP64 solve "
                "time of its class (gauge, Wilson/clover, tol) at kappa 0.117"}
    # the model checked against a measured run: a seeded subsample of the stream's
    # requests solved in plain FP64, one right-hand side at a time, each at its own
    # kappa, c_sw, tol and sources
    sub_n = min(40, n_requests)
    sub = sorted(rng.choice(n_requests, size=sub_n, replace=False).tolist())
    meas, model = [], []
    log(f"serving: measuring the plain FP64 one-at-a-time baseline on {sub_n} of the stream's requests")
    for k in sub:
        r = reqs[k]
        L = vols[r["g"] // 2]
        sp.tol = r["tol"]
        op = q.Operator(r["kappa"], r["csw"], -1, 12)
        x = q.DeviceBuffer(2 * q.half_volume(L) * 192)
        bsrc = [q.source_gaussian_device(L, r["seed"] + j) for j in range(r["nrhs"])]
        t1 = time.perf_counter()
        for b in bsrc:
            st = q.solve_batch_device(gauges[r["g"]], op, [x], [b], sp)[0]
            assert st.converged, ("baseline solve did not converge", r)
        meas.append(time.perf_counter() - t1)
        model.append(service[k])
        for b in bsrc:
            b.free()
        x.free()
    meas, model = np.array(meas), np.array(model)
    relerr = (model - meas) / meas
    out["baseline_fp64_one_at_a_time"]["measured_subsample"] = {
        "requests": sub_n, "rhs": int(sum(reqs[k]["nrhs"] for k in sub)),
        "measured_s": round(float(meas.sum()), 3), "modelled_s": round(float(model.sum()), 3),
        "model_error_total": round(float(model.sum() / meas.sum() - 1.0), 4),
        "model_error_per_request_median": round(float(np.median(relerr)), 4),
        "model_error_per_request_max_abs": round(float(np.abs(relerr).max()), 4),
        "solves_per_s_measured": round(sum(reqs[k]["nrhs"] for k in sub) / float(meas.sum()), 2),
        "note": "each sampled request's right-hand sides solved one at a time in plain FP64 at the request's "
                "own kappa, c_sw, tol and seeded sources; model = nrhs x its class time"}
    # the baseline rate with the model's service times rescaled by its measured error
    out["baseline_fp64_one_at_a_time"]["solves_per_s_corrected"] = round(
        nrhs_total / (sum(service) * float(meas.sum() / model.sum())), 2)
    log(f"serving: baseline model error over the subsample {100 * (model.sum() / meas.sum() - 1):.1f} %")
    out["speedup_solves_per_s"] = round(out["solves_per_s"] / out["baseline_fp64_one_at_a_time"]["solves_per_s"], 3)
    out["speedup_solves_per_s_vs_corrected"] = round(
        out["solves_per_s"] / out["baseline_fp64_one_at_a_time"]["solves_per_s_corrected"], 3)
    log(f"serving: {out['solves_per_s']} solves/s p50 {out['p50_ms']} ms p99 {out['p99_ms']} ms; "
        f"fp64 one-at-a-time {out['baseline_fp64_one_at_a_time']['solves_per_s']} solves/s")
    for g in gauges:
        g.destroy()
    return out


def config4_section(q):
    """32^3x64 clover (c_sw=1, kappa=0.122, eps=0.25), 12 point sources (all spin x
    colour at the origin), tol 1e-10 per RHS: one batched FP32-inner solve with
    reliable updates and graph-replayed iterations against 12 plain FP64 solves run
    one at a time (no graph)."""
    L = (32, 32, 32, 64)
    g = q.Gauge.random(L, seed=4, eps=0.25)
    op = q.Operator(0.122, 1.0, -1, 12)
    vh = q.half_volume(L)
    srcs = [q.source_point_device(L, (0, 0, 0, 0), s, c) for s in range(4) for c in range(3)]
    xs = [q.DeviceBuffer(2 * vh * 192) for _ in range(12)]
    for prec, nrhs in ((1, 12), (0, 1), (1, 1)):
        q.tune_hop(g, prec, 12, nrhs)
    sp64 = q.default_solve_params(tol=1e-10, precision=0, graph=0)
    q.solve_batch_device(g, op, xs[:1], srcs[:1], sp64)  # warm
    sts64, t64 = timed(lambda: [q.solve_batch_device(g, op, [xs[i]], [srcs[i]], sp64)[0] for i in range(12)])
    spm = q.default_solve_params(tol=1e-10, precision=1, graph=1)
    q.solve_batch_device(g, op, xs, srcs, spm)  # warm (tuning, graph capture, buffers)
    stsm, tm = timed(lambda: q.solve_batch_device(g, op, xs, srcs, spm))
    out = {"plain_fp64_sequential": {"seconds": round(t64, 3), "solves_per_s": round(12 / t64, 3),
                                     "iterations": [s.iterations for s in sts64],
                                     "max_true_residual": max(s.true_residual for s in sts64)},
           "batched_mixed_fp32": {"seconds": round(tm, 3), "solves_per_s": round(12 / tm, 3),
                                  "iterations": [s.iterations for s in stsm],
                                  "reliable_updates": [s.reliable_updates for s in stsm],
                                  "max_true_residual": max(s.true_residual for s in stsm)}}
    out["speedup_solves_per_s"] = round(t64 / tm, 3)
    # the same plain FP64 solves with graph-replayed iterations, for comparison
    sp64g = q.default_solve_params(tol=1e-10, precision=0, graph=1)
    q.solve_batch_device(g, op, xs[:1], srcs[:1], sp64g)  # warm (graph capture)
    _, t64g = timed(lambda: [q.solve_batch_device(g, op, [xs[i]], [srcs[i]], sp64g)[0] for i in range(12)])
    out["plain_fp64_sequential_graphs"] = {"seconds": round(t64g, 3), "solves_per_s": round(12 / t64g, 3),
                                           "speedup_batched_mixed_fp32": round(t64g / tm, 3)}
    # beyond the configuration (which names FP32): the same batch with the FP16 +
    # per-site-norm inner solve
    q.tune_hop(g, 2, 12, 12)
    sph = q.default_solve_params(tol=1e-10, precision=2, graph=1)
    q.solve_batch_device(g, op, xs, srcs, sph)  # warm
    stsh, th = timed(lambda: q.solve_batch_device(g, op, xs, srcs, sph))
    out["extra_batched_mixed_fp16"] = {"seconds": round(th, 3), "solves_per_s": round(12 / th, 3),
                                       "iterations": [s.iterations for s in stsh],
                                       "max_true_residual": max(s.true_residual for s in stsh),
                                       "speedup_solves_per_s": round(t64 / th, 3)}
    log(f"config4: plain fp64 12 sequential {t64:.2f} s ({t64g:.2f} s with graphs), batched mixed {tm:.2f} s, "
        f"x{t64/tm:.2f} (x{t64g/tm:.2f}); fp16 inner {th:.2f} s, x{t64/th:.2f}")
    g.destroy()
    return out


def dropin_section(q, g=None):
    """The reference's entry points as a caller sees them, on the 16^3x32 gauge of the
    headline: host-synchronous dslashQcu (one parity, s = +1, periodic, kappa = 1) and
    fullDslashQcu (A - D with kappa = c_sw = 1, even-first full spinors), each call
    including the interface <-> internal layout conversions."""
    import ctypes

    import numpy as np

    if g is None:
        g = q.Gauge.random(L_MAIN, seed=1, eps=0.2)
    L = g.L
    vh = q.half_volume(L)
    gd = q.DeviceBuffer.from_numpy(g.to_numpy())
    prm = q.param(L)
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    q.lib().loadQcuGauge(gd.ptr, ctypes.byref(prm))
    full = q.source_gaussian_device(L, seed=7)
    out = q.DeviceBuffer(2 * vh * 192)
    res = {}
    for name, n in (("dslashQcu", 200), ("fullDslashQcu", 100)):
        def call():
            if name == "dslashQcu":
                q.lib().dslashQcu(out.ptr, full.ptr + vh * 192, gd.ptr, ctypes.byref(prm), 0)
            else:
                q.lib().fullDslashQcu(out.ptr, full.ptr, gd.ptr, ctypes.byref(prm), 0)
        for _ in range(5):
            call()  # warm (clover build on the first fullDslashQcu)
        t0 = time.perf_counter()
        for _ in range(n):
            call()
        us = 1e6 * (time.perf_counter() - t0) / n
        res[name] = {"us_per_call": round(us, 2), "calls": n}
        log(f"drop-in {name} 16^3x32: {us:.1f} us per call")
    q.check(q.lib().qcuGetLastError())
    res["note"] = ("host-synchronous calls through the drop-in symbols, including the layout conversions; "
                   "the hop kernel alone is roofline.avg_launch_us")
    del np
    return res


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def measured_traffic(name):
    """HBM bytes per launch measured with rocprofv3 PMC counters (profiles/traffic.json,
    each entry stamped with the hop sources' hash), or None when there is no
    measurement of this kernel and workload on the current sources."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            e = json.load(f).get(name)
    except (OSError, ValueError):
        return None
    if not isinstance(e, dict) or e.get("kernel_sources_sha16") != hop_source_hash():
        return None
    return e.get("bytes")


def cpu_baseline(seconds_target=15.0):
    """Oracle FP64 hop on the host cores, on a bounded sample of the workload:
    repeated applications on the 16^3x32 lattice for about seconds_target."""
    import numpy as np

    import oracle as orc

    threads = min(os.cpu_count() or 1, 16)
    orc.set_threads(threads)
    L = L_MAIN
    g = orc.gauge_random(L, 1, 0.2)
    v = orc.source_gauss(L, 2)
    orc.hop(v[1], g, L, 0, 1, -1.0, 18)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        orc.hop(v[1 - (n & 1)], g, L, n & 1, 1, -1.0, 18)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds_target:
            break
    vh = orc.half_volume(L)
    gbs = BYTES_PER_SITE[("f64", 18)] * vh * n / el / 1e9
    # the GPU's 16^3x32 solve on the same inputs: Wilson kappa 0.124, antiperiodic t,
    # FP64 even-odd CG to 1e-10, all threads
    t0 = time.perf_counter()
    _, st2 = orc.cg_eo(v, g, L, 0.124, tbc=-1.0, recon=18, tol=1e-10)
    tcg2 = time.perf_counter() - t0
    # the same on one core, and on the 24^3x48 lattice with all of them
    orc.set_threads(1)
    n1, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 3.0:
        orc.hop(v[1 - (n1 & 1)], g, L, n1 & 1, 1, -1.0, 18)
        n1 += 1
    el1 = time.perf_counter() - t0
    orc.set_threads(threads)
    L3 = (24, 24, 24, 48)
    g3 = orc.gauge_random(L3, 1, 0.25)
    v3 = orc.source_gauss(L3, 2)
    orc.hop(v3[1], g3, L3, 0, 1, -1.0, 18)
    n3, t0 = 0, time.perf_counter()
    while n3 < 2 or time.perf_counter() - t0 < 3.0:
        orc.hop(v3[1 - (n3 & 1)], g3, L3, n3 & 1, 1, -1.0, 18)
        n3 += 1
    el3 = time.perf_counter() - t0
    # configuration 3's FP64 clover solve on the same inputs (generators seeded as on
    # the GPU), as a bounded sample: the clover build, then the cost of one CG
    # iteration from the difference of a 5- and a 10-iteration run
    g3 = orc.gauge_random(L3, 5, 0.25)
    b3 = orc.source_gauss(L3, 6)
    t0 = time.perf_counter()
    A3, Ainv3 = orc.clover_fields(g3, L3, 0.122, 1.0, recon=12)
    tcl3 = time.perf_counter() - t0
    tit = []
    for m in (5, 10):
        t0 = time.perf_counter()
        orc.cg_eo(b3, g3, L3, 0.122, A=A3, Ainv=Ainv3, tbc=-1.0, recon=12, tol=1e-10, max_iter=m)
        tit.append(time.perf_counter() - t0)
    it3 = max(tit[1] - tit[0], 1e-9) / 5
    del g3, v3, b3, A3, Ainv3
    # the 32^3x64 hop (configuration 4's lattice) on all threads
    L4 = (32, 32, 32, 64)
    g4 = orc.gauge_random(L4, 3, 0.25)
    v4 = orc.source_gauss(L4, 2)
    orc.hop(v4[1], g4, L4, 0, 1, -1.0, 18)
    n4, t0 = 0, time.perf_counter()
    while n4 < 2 or time.perf_counter() - t0 < 3.0:
        orc.hop(v4[1 - (n4 & 1)], g4, L4, n4 & 1, 1, -1.0, 18)
        n4 += 1
    el4 = time.perf_counter() - t0
    del g4, v4
    # the CPU-runnable configuration: 8^4 Wilson, kappa 0.12, eps 0.1, antiperiodic t,
    # one Gaussian source, FP64 even-odd CG to 1e-10
    L1 = (8, 8, 8, 8)
    g1 = orc.gauge_random(L1, 1, 0.1)
    b1 = orc.source_gauss(L1, 2)
    t0 = time.perf_counter()
    _, st = orc.cg_eo(b1, g1, L1, 0.12, tbc=-1.0, recon=18, tol=1e-10)
    tcg = time.perf_counter() - t0
    del np
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{n} FP64 hop applications on 16^3x32 ({el:.1f} s, oracle, OpenMP)",
            "ms_per_apply": round(1e3 * el / n, 3),
            "one_core": {"GBps": round(BYTES_PER_SITE[("f64", 18)] * vh * n1 / el1 / 1e9, 3),
                         "ms_per_apply": round(1e3 * el1 / n1, 3), "applications": n1},
            "hop_24x24x24x48": {"GBps": round(BYTES_PER_SITE[("f64", 18)] * orc.half_volume(L3) * n3 / el3 / 1e9, 3),
                                "ms_per_apply": round(1e3 * el3 / n3, 3), "applications": n3, "cores": threads},
            "hop_32x32x32x64": {"GBps": round(BYTES_PER_SITE[("f64", 18)] * orc.half_volume(L4) * n4 / el4 / 1e9, 3),
                                "ms_per_apply": round(1e3 * el4 / n4, 3), "applications": n4, "cores": threads},
            "cg_config3_24x24x24x48_clover_sample": {
                "clover_build_seconds": round(tcl3, 3), "ms_per_iteration": round(1e3 * it3, 2),
                "projected_seconds_100_iterations": round(tcl3 + tit[1] + 90 * it3, 2), "cores": threads,
                "sample": "clover build, then FP64 CG runs of 5 and 10 iterations (per-iteration cost from their "
                          "difference; the projection is the build, the 10-iteration run and 90 more iterations); "
                          "the GPU's plain FP64 solve of the same system takes 100 iterations"},
            "host_cpu": platform.processor() or platform.machine(),
            "cg_config1_8x8x8x8": {"seconds": round(tcg, 4), "iterations": st.iterations,
                                   "true_residual": st.true_residual},
            "cg_16x16x16x32_wilson": {"seconds": round(tcg2, 3), "iterations": st2.iterations,
                                      "true_residual": st2.true_residual, "cores": threads}}


def config1_gpu(q):
    """8^4 Wilson, kappa 0.12, eps 0.1, antiperiodic t, 1 Gaussian source: plain FP64 GPU
    solve and the mixed solve, true residual 1e-10."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=1, eps=0.1)
    b = q.source_gaussian_device(L, seed=2)
    x = q.DeviceBuffer(2 * q.half_volume(L) * 192)
    out = {}
    for name, prec, graph in (("fp64_plain", 0, 0), ("mixed_fp32", 1, 1)):
        sp = q.default_solve_params(tol=1e-10, precision=prec, graph=graph)
        q.solve_batch_device(g, q.Operator(0.12, 0.0, -1, 18), [x], [b], sp)
        st = q.solve_batch_device(g, q.Operator(0.12, 0.0, -1, 18), [x], [b], sp)[0]
        out[name] = {"seconds": round(st.seconds, 5), "iterations": st.iterations, "true_residual": st.true_residual}
    g.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--quick", action="store_true", help="skip the large lattices and the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="requests in the Poisson serving stream (2000 for the full configuration; 0 skips)")
    ap.add_argument("--serving-precision", type=int, default=2, choices=[1, 2],
                    help="inner precision of the serving solves: 1 FP32, 2 FP16 + per-site norm (default; "
                         "67 solves/s against 57 with FP32 on the 2,000-request stream, every request converged)")
    ap.add_argument("--kernels-12rhs", action="store_true",
                    help="profiler driver: copy kernel, then the 32^3x64 FP32 and FP16 hops with 1 and 12 "
                         "right-hand sides")
    ap.add_argument("--section", choices=["config1", "config2", "config3", "config4", "serving", "dslash_12rhs",
                                          "dropin"],
                    help="profiler driver: run one section only and print its JSON")
    ap.add_argument("--kernels-large", action="store_true",
                    help="profiler driver: 1 GiB copy, then the FP64 recon-18 and recon-12 hops on 32^3x64")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profiler driver: 1 GiB copy kernel x5 then the 16^3x32 FP64 hop x steps, nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build keeps the lattice on one GPU; --gpus {args.gpus} is not supported "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import qcu_cg_inverter_amd as q

    q.init(0)
    if args.kernels_only:
        q.bench_copy(1 << 30, iters=5)
        gk = q.Gauge.random(L_MAIN, seed=1, eps=0.2)
        ms = q.bench_hop(gk, precision=0, recon=18, nrhs=1, iters=args.steps)
        log(f"kernels-only: hop f64 r18 16^3x32 {ms*1e3:.2f} us/launch over {args.steps}")
        return 0
    if args.kernels_large:
        # profiler driver: with QCU_TUNE_CACHE naming an existing file the shapes come
        # from it, so profiled runs launch no tuning
        q.bench_copy(1 << 30, iters=5)
        gk = q.Gauge.random((32, 32, 32, 64), seed=3, eps=0.25)
        cache = os.environ.get("QCU_TUNE_CACHE", "")
        have = bool(cache) and os.path.exists(cache)
        for recon in (18, 12):
            if not have:
                q.tune_hop(gk, 0, recon, 1)
            ms = q.bench_hop(gk, precision=0, recon=recon, nrhs=1, iters=args.steps)
            log(f"kernels-large: f64 r{recon} 32^3x64 {ms*1e3:.1f} us/launch")
        return 0
    if args.section:
        fn = {"config1": config1_gpu, "config2": config2_section, "config3": config3_section,
              "config4": config4_section, "dropin": dropin_section,
              "serving": lambda q_: serving_section(q_, args.serving_requests, precision=args.serving_precision),
              "dslash_12rhs": multi_rhs_section}[args.section]
        print(json.dumps({args.section: fn(q)}))
        return 0
    if args.kernels_12rhs:
        # profiler driver: 1 GiB copy x5, then on 32^3x64 the FP32 and the FP16 recon-12
        # hops with 1 right-hand side and with 12 in the autotuned shape.  With
        # QCU_TUNE_CACHE naming an existing file the shapes come from it, so profiled
        # runs launch no tuning.
        q.bench_copy(1 << 30, iters=5)
        gk = q.Gauge.random((32, 32, 32, 64), seed=3, eps=0.25)
        cache = os.environ.get("QCU_TUNE_CACHE", "")
        have = bool(cache) and os.path.exists(cache)
        for prec in (1, 2):
            for n in (1, 12):
                if not have:
                    q.tune_hop(gk, prec, 12, n)
                ms = q.bench_hop(gk, precision=prec, recon=12, nrhs=n, iters=args.steps)
                log(f"kernels-12rhs: {('f64', 'f32', 'f16')[prec]} r12 32^3x64 nrhs={n} {ms*1e3:.1f} us/launch")
        return 0
    out = {"metric": METRIC, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded near-unit SU(3) gauge eps=0.2, Gaussian spinors)"}
    vh = q.half_volume(L_MAIN)
    g = q.Gauge.random(L_MAIN, seed=1, eps=0.2)

    # ---- headline: FP64 recon-18 hop on 16^3x32 --------------------------------
    q.bench_hop(g, precision=0, recon=18, nrhs=1, iters=max(1, args.warmup))
    q.check(q.lib().qcuSynchronize())
    t0 = time.perf_counter()
    ms_event = q.bench_hop(g, precision=0, recon=18, nrhs=1, iters=args.steps)
    q.check(q.lib().qcuSynchronize())
    wall = time.perf_counter() - t0
    bps = BYTES_PER_SITE[("f64", 18)] * vh
    achieved = bps / (ms_event * 1e-3) / 1e9
    out["value"] = round(bps / (wall / args.steps) / 1e9, 2)
    out["ms_per_step"] = round(1e3 * wall / args.steps, 5)
    out["config"] = {"workload": "even-odd Wilson hop D_eo/D_oe, FP64, 18-real links, 16^3x32, antiperiodic t",
                     "lattice": list(L_MAIN), "half_volume": vh, "bytes_per_site": 2880,
                     "flop_per_site": 1320}
    out["roofline"] = roofline("hop_kernel<F64,18,EPI_PLAIN> 16^3x32", "f64", 18, vh, ms_event, "hop_f64_r18_16x32")
    log(f"hop f64 r18 16^3x32: {ms_event*1e3:.2f} us/launch, {achieved:.0f} GB/s algorithmic, "
        f"{out['roofline']['achieved']:.0f} GB/s compulsory")
    # the same hop as the median of 100 windows of 20 back-to-back launches, each
    # window timed by its own pair of HIP events
    win = sorted(q.bench_hop(g, precision=0, recon=18, nrhs=1, iters=20) for _ in range(100))
    med = 0.5 * (win[49] + win[50])
    out["dslash_16x32_median"] = {"windows": 100, "launches_per_window": 20,
                                  "median_us": round(med * 1e3, 3), "min_us": round(win[0] * 1e3, 3),
                                  "max_us": round(win[-1] * 1e3, 3),
                                  "GBps": round(bps / (med * 1e-3) / 1e9, 1),
                                  "pct_hbm_peak": round(100 * bps / (med * 1e-3) / 1e9 / HBM_PEAK_GBS, 1)}
    log(f"hop f64 r18 16^3x32 median of 100 windows: {med*1e3:.2f} us/launch")

    # ---- copy bandwidth -----------------------------------------------------------
    nb = 1 << 30
    ms = q.bench_copy(nb, iters=20)
    out["copy"] = {"bytes": nb, "GBps_read_plus_write": round(2 * nb / (ms * 1e-3) / 1e9, 1)}
    out["dropin_16x32"] = dropin_section(q, g)

    # ---- CG: FP64 vs mixed on 16^3x32 Wilson ----------------------------------------
    out["cg"] = config2_section(q, g)

    # ---- larger lattices (beyond the Infinity Cache) ---------------------------------
    if not args.quick:
        large = {}
        for L in ((24, 24, 24, 48), (32, 32, 32, 64)):
            gl = q.Gauge.random(L, seed=3, eps=0.25)
            vhl = q.half_volume(L)
            key = "x".join(map(str, L))
            large[key] = {}
            for prec in ("f64", "f32", "f16"):
                for recon in (18, 12):
                    ms = q.bench_hop(gl, precision=PREC[prec], recon=recon, nrhs=1, iters=100)
                    gbs = BYTES_PER_SITE[(prec, recon)] * vhl / (ms * 1e-3) / 1e9
                    row = {"us": round(ms * 1e3, 2), "GBps": round(gbs, 1),
                           "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 1)}
                    # real traffic: PMC-measured bytes per launch (profiles/hop_f64_32x64.md)
                    tb = measured_traffic(f"hop_{prec}_r{recon}_{key}")
                    if tb:
                        real = tb / (ms * 1e-3) / 1e9
                        row.update({"traffic_bytes": tb, "real_GBps": round(real, 1),
                                    "real_pct_hbm_peak": round(100 * real / HBM_PEAK_GBS, 1)})
                    large[key][f"{prec}_r{recon}"] = row
                    log(f"hop {prec} r{recon} {key}: {ms*1e3:.1f} us, {gbs:.0f} GB/s")
                    if key == "32x32x32x64" and prec == "f64" and recon == 18:
                        out["roofline_32x64"] = roofline("hop_kernel<F64,18,EPI_PLAIN> 32^3x64", prec, recon, vhl,
                                                         ms, f"hop_{prec}_r{recon}_{key}")
            gl.destroy()
        out["dslash_large"] = large
        out["config1_wilson_8x8x8x8"] = config1_gpu(q)
        out["dslash_12rhs"] = multi_rhs_section(q)
        out["config3_clover_24x48"] = config3_section(q)
        out["config4_clover_32x64_12rhs"] = config4_section(q)
        if args.serving_requests > 0:
            out["config5_serving"] = serving_section(q, args.serving_requests, precision=args.serving_precision)

    if not (args.quick or args.no_cpu):
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baseline"]["host"] = cpu_info()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
