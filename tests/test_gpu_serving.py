"""GPU tests of the continuous-batching server (qcuServer*): every request converges
to its tolerance (true residual recomputed in FP64 and checked by the oracle), a
request gives the same bits alone and in a mix of other requests that arrive while
it runs, and admission rejects what cannot fit the HBM budget."""
import time

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def q():
    import qcu_cg_inverter_amd as q

    q.init(0)
    return q


def run_request(q, server, g, L, kappa, csw, tol, bs):
    vh = q.half_volume(L)
    xs = [q.DeviceBuffer(2 * vh * 192) for _ in bs]
    rid = server.submit(g, kappa, csw, tol, len(bs), xs=xs, bs=bs)
    return rid, xs


def test_server_alone_equals_mix(q):
    La, Lb = (8, 8, 8, 8), (8, 8, 4, 8)
    ga = q.Gauge.random(La, seed=41, eps=0.25)
    gb = q.Gauge.random(Lb, seed=42, eps=0.25)
    vh = q.half_volume(La)
    src = [q.DeviceBuffer.from_numpy(orc.source_gauss(La, 70 + i)) for i in range(3)]
    # alone
    s1 = q.Server(precision=1, max_batch=8, chunk=4)
    rid, xs = run_request(q, s1, ga, La, 0.121, 1.0, 1e-10, src)
    st = s1.wait(rid)
    alone = [x.to_numpy((2, vh, 4, 3)) for x in xs]
    it_alone = st.iterations
    s1.destroy()
    assert st.converged and st.true_residual <= 1e-10
    # in a mix: others before, during and after, on both gauges, Wilson and clover,
    # different kappa, some arriving while the request runs
    s2 = q.Server(precision=1, max_batch=8, chunk=4)
    others = []
    for k in range(3):
        others.append(s2.submit(ga, 0.115 + 0.002 * k, 1.0, 1e-8, 2, seed=100 + k))
        others.append(s2.submit(gb, 0.12, 0.0, 1e-10, 1 + k, seed=200 + k))
    time.sleep(0.002)
    rid, xs = run_request(q, s2, ga, La, 0.121, 1.0, 1e-10, src)
    for k in range(4):
        time.sleep(0.001)
        others.append(s2.submit(ga, 0.118, 0.0, 1e-10, 3, seed=300 + k))
        others.append(s2.submit(ga, 0.121, 1.0, 1e-10, 2, seed=400 + k))
    st2 = s2.wait(rid)
    mix = [x.to_numpy((2, vh, 4, 3)) for x in xs]
    for o in others:
        so = s2.wait(o)
        assert so.converged
    s2.destroy()
    assert st2.iterations == it_alone
    for a, m in zip(alone, mix):
        assert np.array_equal(a, m)
    # and the answer is right (oracle residual in FP64)
    gref = ga.to_numpy()
    A = orc.clover_fields(gref, La, 0.121, 1.0, recon=12)[0]
    for i in range(3):
        b = src[i].to_numpy((2, vh, 4, 3))
        r = b - orc.full_op(mix[i], gref, La, 0.121, A=A, s=1, tbc=-1.0, recon=12)
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-10


def test_server_burst_all_converge(q):
    rng = np.random.default_rng(5)
    Ls = [(8, 8, 8, 8), (16, 8, 8, 8)]
    gs = [q.Gauge.random(L, seed=50 + i, eps=0.25) for i, L in enumerate(Ls)]
    s = q.Server(precision=1, max_batch=12, chunk=8)
    ids = []
    for k in range(24):
        gi = int(rng.integers(2))
        kappa = float(rng.uniform(0.11, 0.124))
        csw = 1.0 if rng.random() < 0.5 else 0.0
        tol = 1e-8 if rng.random() < 0.5 else 1e-10
        nrhs = int(rng.integers(1, 13))
        ids.append((s.submit(gs[gi], kappa, csw, tol, nrhs, seed=1000 + 13 * k), tol))
    s.drain()
    for rid, tol in ids:
        st = s.poll(rid)
        assert st.done and st.converged and st.true_residual <= tol, st.as_dict()
        assert st.latency_seconds >= st.queue_seconds >= 0
    use = s.usage()
    assert use["peak_bytes"] > 0
    s.destroy()


def test_server_rejects_over_budget(q):
    L = (16, 16, 16, 16)
    g = q.Gauge.random(L, seed=3, eps=0.2)
    s = q.Server(precision=1, max_batch=12, hbm_budget=1 << 20)
    with pytest.raises(q.QcuError) as e:
        s.submit(g, 0.12, 0.0, 1e-8, 1, seed=1)
    assert e.value.code == 8
    s.destroy()


def test_server_bad_request_fails_alone(q):
    """A request with a non-finite source fails on its own (not converged, infinite
    residual); requests before and after it on the same gauge still converge."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=44, eps=0.25)
    vh = q.half_volume(L)
    bad = orc.source_gauss(L, 9)
    bad[1, 5, 2, 1] = np.nan
    bbad = q.DeviceBuffer.from_numpy(bad)
    xbad = q.DeviceBuffer(2 * vh * 192)
    s = q.Server(precision=1, max_batch=4, chunk=4)
    before = s.submit(g, 0.12, 1.0, 1e-10, 2, seed=11)
    rid = s.submit(g, 0.12, 1.0, 1e-10, 1, xs=[xbad], bs=[bbad])
    after = s.submit(g, 0.12, 1.0, 1e-10, 3, seed=12)
    s.drain()
    st = s.poll(rid)
    assert st.done and not st.converged and not np.isfinite(st.true_residual)
    for other in (before, after):
        so = s.poll(other)
        assert so.done and so.converged and so.true_residual <= 1e-10
    s.destroy()


def test_server_bad_rhs_of_a_group_fails_alone(q):
    """A request's right-hand sides are set up together; when one of them has a
    non-finite source it fails alone: the request ends not converged with an infinite
    residual, and its other right-hand sides are solved, with the bits they get in a
    request of their own and their true residual by the oracle."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=45, eps=0.25)
    vh = q.half_volume(L)
    good = [orc.source_gauss(L, 80), orc.source_gauss(L, 81)]
    bad = orc.source_gauss(L, 82)
    bad[0, 3, 1, 2] = np.inf
    s = q.Server(precision=1, max_batch=8, chunk=4)
    bs = [q.DeviceBuffer.from_numpy(good[0]), q.DeviceBuffer.from_numpy(bad), q.DeviceBuffer.from_numpy(good[1])]
    rid, xs = run_request(q, s, g, L, 0.121, 1.0, 1e-10, bs)
    st = s.wait(rid, raise_on_error=False)
    assert st.done and not st.converged and not np.isfinite(st.true_residual)
    ref_bs = [q.DeviceBuffer.from_numpy(b) for b in good]
    rid2, xs2 = run_request(q, s, g, L, 0.121, 1.0, 1e-10, ref_bs)
    st2 = s.wait(rid2)
    assert st2.converged and st2.true_residual <= 1e-10
    s.destroy()
    gref = g.to_numpy()
    A = orc.clover_fields(gref, L, 0.121, 1.0, recon=12)[0]
    for i, j in ((0, 0), (2, 1)):
        x = xs[i].to_numpy((2, vh, 4, 3))
        assert np.array_equal(x, xs2[j].to_numpy((2, vh, 4, 3))), i
        r = good[j] - orc.full_op(x, gref, L, 0.121, A=A, s=1, tbc=-1.0, recon=12)
        assert np.linalg.norm(r) / np.linalg.norm(good[j]) <= 1.01e-10


def test_server_stalled_request_not_starved(q):
    """A request that needs memory held by a busy batch of another gauge gets in while
    a backlog for that batch is still queued: after a short grace nothing behind it is
    admitted, the busy batch drains and gives its memory back.  (Without the grace the
    backlog would keep the batch busy and the request would wait for all of it.)"""
    L = (8, 8, 8, 8)
    g1 = q.Gauge.random(L, seed=51, eps=0.2)
    g2 = q.Gauge.random(L, seed=52, eps=0.2)
    # the budget of exactly one batch (measured with a server that makes one)
    probe = q.Server(precision=1, max_batch=12, chunk=2)
    probe.wait(probe.submit(g1, 0.12, 0.0, 1e-10, 1, seed=1))
    one_batch = probe.usage()["bytes"]
    probe.destroy()
    assert one_batch > 0
    s = q.Server(precision=1, max_batch=12, chunk=2, hbm_budget=one_batch + 4096)
    first = s.submit(g1, 0.12, 0.0, 1e-12, 12, seed=100)
    stalled = s.submit(g2, 0.12, 0.0, 1e-10, 1, seed=200)  # needs a second batch: no memory
    backlog = [s.submit(g1, 0.12, 0.0, 1e-12, 12, seed=300 + k) for k in range(3000)]
    st = s.wait(stalled)
    pending = sum(1 for i in backlog if not s.poll(i).done)
    assert st.converged and st.true_residual <= 1e-10
    assert pending >= 1000, f"the stalled request waited for the backlog ({pending} of 3000 left)"
    s.drain()
    assert s.poll(first).converged and all(s.poll(i).converged for i in backlog)
    assert s.usage()["peak_bytes"] <= one_batch + 4096
    s.destroy()
    g1.destroy()
    g2.destroy()


def test_server_config5_mix_all_volumes(q):
    """Configuration 5's request mix under test: 8 resident gauges, two per volume over
    8^4, 16^3x32, 24^3x48 and 32^3x64; Wilson and clover (c_sw = 1), kappa ~ U[0.11,
    0.124], 1-12 right-hand sides, tol in {1e-8, 1e-10}; 40 requests submitted in
    waves while earlier ones run.  Every request converges to its tolerance; for one
    request per volume the oracle recomputes the true residual of its solutions; and a
    32^3x64 clover request gives the same bits as when it is served alone."""
    rng = np.random.default_rng(2025)
    vols = [(8, 8, 8, 8), (16, 16, 16, 32), (24, 24, 24, 48), (32, 32, 32, 64)]
    gauges = [q.Gauge.random(vols[i // 2], seed=900 + i, eps=0.2 + 0.05 * (i % 2)) for i in range(8)]

    def draw(gi):
        return {"g": gi, "kappa": float(rng.uniform(0.11, 0.124)), "csw": 1.0 if rng.random() < 0.5 else 0.0,
                "nrhs": int(rng.integers(1, 13)), "tol": 1e-8 if rng.random() < 0.5 else 1e-10,
                "seed": int(rng.integers(1 << 40))}

    order = [int(i) for i in rng.permutation(8)] + [int(rng.integers(8)) for _ in range(32)]
    reqs = [draw(gi) for gi in order]
    # checked requests (explicit sources, solutions kept): the first of each volume,
    # and a 32^3x64 clover one for the alone == mix comparison
    checked = {}
    for k, r in enumerate(reqs):
        vi = r["g"] // 2
        if vi not in checked:
            r["nrhs"] = min(r["nrhs"], 2)
            checked[vi] = k
    big = {"g": 6, "kappa": 0.12, "csw": 1.0, "nrhs": 1, "tol": 1e-10, "seed": 77}
    reqs.insert(12, big)
    k_big = 12
    checked = {vi: (k if k < k_big else k + 1) for vi, k in checked.items()}
    bufs = {}

    def buffers(k, r):
        L = vols[r["g"] // 2]
        vh = q.half_volume(L)
        bs_h = [orc.source_gauss(L, r["seed"] + j) for j in range(r["nrhs"])]
        return bs_h, [q.DeviceBuffer.from_numpy(b) for b in bs_h], [q.DeviceBuffer(2 * vh * 192) for _ in bs_h]

    s = q.Server(precision=1, max_batch=12, chunk=8)
    ids = []
    for k, r in enumerate(reqs):
        if k in checked.values() or k == k_big:
            bufs[k] = buffers(k, r)
            ids.append(s.submit(gauges[r["g"]], r["kappa"], r["csw"], r["tol"], r["nrhs"], xs=bufs[k][2],
                                bs=bufs[k][1]))
        else:
            ids.append(s.submit(gauges[r["g"]], r["kappa"], r["csw"], r["tol"], r["nrhs"], seed=r["seed"]))
        if k % 8 == 7:
            time.sleep(0.05)  # the next wave arrives while this one runs
    s.drain()
    for rid, r in zip(ids, reqs):
        st = s.poll(rid)
        assert st.done and st.converged and st.true_residual <= r["tol"], (r, st.as_dict())
    peak = s.usage()["peak_bytes"]
    s.destroy()
    assert peak > 0
    # the oracle's residuals, one request per volume
    for vi, k in sorted(checked.items()):
        r = reqs[k]
        L = vols[vi]
        vh = q.half_volume(L)
        gref = gauges[r["g"]].to_numpy()
        A = orc.clover_fields(gref, L, r["kappa"], r["csw"], recon=12)[0] if r["csw"] else None
        bs_h, _, xs = bufs[k]
        for b, xd in zip(bs_h, xs):
            x = xd.to_numpy((2, vh, 4, 3))
            res = b - orc.full_op(x, gref, L, r["kappa"], A=A, s=1, tbc=-1.0, recon=12)
            assert np.linalg.norm(res) / np.linalg.norm(b) <= 1.01 * r["tol"], (L, r)
        del A, gref
    # the 32^3x64 clover request alone gives the bits it got in the mix
    L = vols[3]
    vh = q.half_volume(L)
    mix = bufs[k_big][2][0].to_numpy((2, vh, 4, 3))
    s1 = q.Server(precision=1, max_batch=12, chunk=8)
    xa = q.DeviceBuffer(2 * vh * 192)
    st = s1.wait(s1.submit(gauges[6], 0.12, 1.0, 1e-10, 1, xs=[xa], bs=[bufs[k_big][1][0]]))
    s1.destroy()
    assert st.converged
    assert np.array_equal(xa.to_numpy((2, vh, 4, 3)), mix)
    for g in gauges:
        g.destroy()


def test_server_rejects_clover_beyond_budget(q):
    """A budget that holds one batch but not the clover field a clover request also
    needs: the request is rejected at submit (code 8) instead of waiting forever."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=61, eps=0.2)
    probe = q.Server(precision=1, max_batch=4, chunk=2)
    probe.wait(probe.submit(g, 0.12, 0.0, 1e-10, 1, seed=1))
    one_batch = probe.usage()["bytes"]
    probe.destroy()
    s = q.Server(precision=1, max_batch=4, chunk=2, hbm_budget=one_batch + 4096)
    with pytest.raises(q.QcuError) as e:
        s.submit(g, 0.12, 1.0, 1e-10, 1, seed=2)
    assert e.value.code == 8
    rid = s.submit(g, 0.12, 0.0, 1e-10, 1, seed=3)  # a Wilson request still fits
    assert s.wait(rid).converged
    s.destroy()
    g.destroy()


def test_server_clover_after_idle_batch_of_other_gauge(q):
    """A clover request finds an idle batch of its own kind with free slots, but its
    new clover field fits only once an idle batch of another gauge gives its memory
    back: that batch is released and the request admitted, not left waiting."""
    L = (8, 8, 8, 8)
    g1 = q.Gauge.random(L, seed=62, eps=0.2)
    g2 = q.Gauge.random(L, seed=63, eps=0.2)
    probe = q.Server(precision=1, max_batch=4, chunk=2)
    probe.wait(probe.submit(g1, 0.12, 0.0, 1e-10, 1, seed=1))
    one_batch = probe.usage()["bytes"]
    probe.destroy()
    s = q.Server(precision=1, max_batch=4, chunk=2, hbm_budget=2 * one_batch + 4096)
    assert s.wait(s.submit(g1, 0.115, 1.0, 1e-10, 1, seed=2)).converged  # idle clover batch of g1
    assert s.wait(s.submit(g2, 0.12, 0.0, 1e-10, 1, seed=3)).converged  # idle Wilson batch of g2
    assert s.usage()["batches"] == 2
    st = s.wait(s.submit(g1, 0.121, 1.0, 1e-10, 2, seed=4))  # a second kappa: a new clover field
    assert st.converged and st.true_residual <= 1e-10
    assert s.usage()["peak_bytes"] <= 2 * one_batch + 4096
    s.destroy()
    g1.destroy()
    g2.destroy()


def test_two_servers_share_a_clover_field(q):
    """Two servers (FP32 and FP16 inner) solve clover requests on the same gauge and
    kappa * c_sw, so their slots use one clover field.  One server finishes (and
    releases its use) while the other still solves; the other's results stay right."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=64, eps=0.25)
    vh = q.half_volume(L)
    sa = q.Server(precision=1, max_batch=12, chunk=2)
    sb = q.Server(precision=2, max_batch=12, chunk=2)
    bs_h = [orc.source_gauss(L, 500 + j) for j in range(6)]
    bs = [q.DeviceBuffer.from_numpy(b) for b in bs_h]
    xs = [q.DeviceBuffer(2 * vh * 192) for _ in bs]
    rb = sb.submit(g, 0.121, 1.0, 1e-10, 6, xs=xs, bs=bs)
    ra = sa.submit(g, 0.121, 1.0, 1e-8, 1, seed=9)
    assert sa.wait(ra).converged
    sa.destroy()
    st = sb.wait(rb)
    sb.destroy()
    assert st.converged and st.true_residual <= 1e-10
    gref = g.to_numpy()
    A = orc.clover_fields(gref, L, 0.121, 1.0, recon=12)[0]
    for b, xd in zip(bs_h, xs):
        r = b - orc.full_op(xd.to_numpy((2, vh, 4, 3)), gref, L, 0.121, A=A, s=1, tbc=-1.0, recon=12)
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-10
    g.destroy()
