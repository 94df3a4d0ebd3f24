"""GPU parity: libqcu.so (through its C ABI) against the FP64 oracle on the same
seeded inputs.  FP64 tolerance 1e-14 relative (norm-wise) for the hopping term and
operators; FP32 1e-6, FP16 (+ per-site norm) 3e-3."""
import ctypes

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-14


@pytest.fixture(scope="module")
def q():
    import qcu_cg_inverter_amd as q

    q.init(0)
    return q


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (8, 8, 8, 8), (6, 4, 2, 8), (16, 16, 16, 32)])
def test_generators_match_oracle(q, L):
    g = q.Gauge.random(L, seed=7, eps=0.2)
    ref = orc.gauge_random(L, 7, 0.2)
    assert np.abs(g.to_numpy() - ref).max() < 1e-14
    v = q.source_gaussian(L, seed=3)
    assert np.abs(v - orc.source_gauss(L, 3)).max() < 1e-13


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (8, 8, 8, 8), (6, 4, 2, 8), (2, 2, 2, 2)])
@pytest.mark.parametrize("recon", [12, 18])
def test_hop_f64_matches_oracle(q, L, recon):
    g = q.Gauge.random(L, seed=11, eps=0.5)
    gref = g.to_numpy()
    v = orc.source_gauss(L, 5)
    for parity in (0, 1):
        for dag in (0, 1):
            for tb in (1, -1):
                out = q.hop(g, v[1 - parity], parity, dagger=dag, t_boundary=tb, recon=recon)
                ref = orc.hop(v[1 - parity], gref, L, parity, s=-1 if dag else 1, tbc=float(tb), recon=recon)
                assert rel(out, ref) < FP64_TOL, (parity, dag, tb)


def test_hop_f64_16x32(q):
    L = (16, 16, 16, 32)
    g = q.Gauge.random(L, seed=1, eps=0.2)
    gref = g.to_numpy()
    v = orc.source_gauss(L, 2)
    out = q.hop(g, v[1], 0, t_boundary=-1, recon=18)
    ref = orc.hop(v[1], gref, L, 0, s=1, tbc=-1.0, recon=18)
    assert rel(out, ref) < FP64_TOL


@pytest.mark.parametrize("prec,tol", [(1, 1e-6), (2, 3e-3)])
@pytest.mark.parametrize("recon", [12, 18])
def test_hop_low_precision(q, prec, tol, recon):
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=4, eps=0.3)
    gref = g.to_numpy()
    v = orc.source_gauss(L, 6)
    for parity in (0, 1):
        out = q.hop(g, v[1 - parity], parity, dagger=1, t_boundary=-1, recon=recon, precision=prec)
        ref = orc.hop(v[1 - parity], gref, L, parity, s=-1, tbc=-1.0, recon=recon)
        assert rel(out, ref) < tol


@pytest.mark.parametrize("ck", [0.122, 1.0])
def test_clover_matches_oracle(q, ck):
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=2, eps=0.25)
    gref = g.to_numpy()
    A, Ainv, nf = g.clover(ck, 1.0)
    T = orc.clover_T(gref, L, recon=18)
    Aref = orc.clover_A(T, ck)
    assert np.abs(A - Aref).max() < 1e-14
    inv_ref, nbad = orc.clover_invert(Aref, chol=True)
    assert nf == nbad
    cond = np.linalg.cond(Aref.reshape(-1, 6, 6)).max()
    assert np.abs(Ainv - inv_ref).max() < 1e-14 * cond * max(1.0, np.abs(inv_ref).max())
    assert np.abs(np.einsum("...ij,...jk->...ik", A, Ainv) - np.eye(6)).max() < 1e-14 * cond


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_full_operator_matches_oracle(q, csw):
    L = (8, 4, 4, 8)
    kappa = 0.124
    g = q.Gauge.random(L, seed=9, eps=0.25)
    gref = g.to_numpy()
    A = orc.clover_fields(gref, L, kappa, csw, recon=18)[0] if csw else None
    v = orc.source_gauss(L, 1)
    for dag in (0, 1):
        for tb in (1, -1):
            op = q.Operator(kappa, csw, tb, 18)
            out = q.apply_m(g, op, v, dagger=dag)
            ref = orc.full_op(v, gref, L, kappa, A=A, s=-1 if dag else 1, tbc=float(tb), recon=18)
            assert rel(out, ref) < FP64_TOL


@pytest.mark.parametrize("csw", [0.0, 1.0])
@pytest.mark.parametrize("prec,tol", [(0, FP64_TOL), (1, 1e-6), (2, 3e-3)])
def test_schur_operator_matches_oracle(q, csw, prec, tol):
    L = (8, 8, 4, 4)
    kappa = 0.122
    g = q.Gauge.random(L, seed=12, eps=0.25)
    gref = g.to_numpy()
    A = Ainv = None
    if csw:
        A, Ainv = orc.clover_fields(gref, L, kappa, csw, recon=12)
    v = orc.source_gauss(L, 8)[0]
    for dag in (0, 1):
        op = q.Operator(kappa, csw, -1, 12)
        out = q.apply_mhat(g, op, v, dagger=dag, precision=prec)
        ref = orc.schur_op(v, gref, L, kappa, Aee=None if A is None else A[0],
                           Ainv_oo=None if Ainv is None else Ainv[1], s=-1 if dag else 1, tbc=-1.0, recon=12)
        assert rel(out, ref) < (tol if prec else 10 * tol)


def test_fp16_site_norms_down_to_denormals(q):
    """FP16 + per-site norm on a field whose site magnitudes fall from 1 through the
    float denormals to below their end (1e-46): every stored site stays finite (a norm
    below 2^-128 has no finite float reciprocal), and the hop keeps the FP16 tolerance
    against the oracle over the whole field and over the slice of sites between 1e-20
    and 1e-40."""
    L = (8, 8, 4, 8)
    g = q.Gauge.random(L, seed=13, eps=0.25)
    gref = g.to_numpy()
    v = orc.source_gauss(L, 9)[1]
    vh = v.shape[0]
    v = v * (10.0 ** (-46.0 * np.arange(vh) / (vh - 1)))[:, None, None]
    lo, hi = int(vh * 20 / 46), int(vh * 40 / 46)
    for dag in (0, 1):
        out = q.hop(g, v, 0, dagger=dag, t_boundary=-1, recon=12, precision=2)
        ref = orc.hop(v, gref, L, 0, s=-1 if dag else 1, tbc=-1.0, recon=12)
        assert np.isfinite(out).all(), dag
        assert rel(out, ref) < 3e-3, dag
        assert rel(out[lo:hi], ref[lo:hi]) < 3e-3, dag


def test_fp16_batch_point_sources_converge(q):
    """Twelve point sources (every spin and colour at the origin) solved together with
    the FP16 + per-site-norm inner loop: the fields start as one nonzero site, so the
    hops see sites of exact zeros and exactly cancelling spin projections, and
    magnitudes that fall off fast with the distance from the source.  Every right-hand
    side converges to 1e-10 with the batch's true residual."""
    L = (8, 8, 8, 16)
    g = q.Gauge.random(L, seed=4, eps=0.25)
    op = q.Operator(0.122, 1.0, -1, 12)
    vh = q.half_volume(L)
    srcs = [q.source_point_device(L, (0, 0, 0, 0), s, c) for s in range(4) for c in range(3)]
    xs = [q.DeviceBuffer(2 * vh * 192) for _ in range(12)]
    sts = q.solve_batch_device(g, op, xs, srcs, q.default_solve_params(tol=1e-10, precision=2))
    for r, st in enumerate(sts):
        assert st.converged and st.true_residual <= 1e-10, (r, st.true_residual)
    g.destroy()


# ---------------------------------------------------------------- drop-in ---
def dev(a):
    import qcu_cg_inverter_amd as q

    return q.DeviceBuffer.from_numpy(np.ascontiguousarray(a, np.complex128))


def test_dropin_dslash_and_full(q):
    L = (8, 8, 8, 16)
    gref = orc.gauge_random(L, 21, 0.25)
    prm = q.param(L)
    gd = dev(gref)
    v = orc.source_gauss(L, 4)
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    q.lib().loadQcuGauge(gd.ptr, ctypes.byref(prm))
    vh = orc.half_volume(L)
    for parity in (0, 1):
        inp = dev(v[1 - parity])
        out = q.DeviceBuffer(vh * 192)
        q.lib().dslashQcu(out.ptr, inp.ptr, gd.ptr, ctypes.byref(prm), parity)
        res = out.to_numpy((vh, 4, 3))
        ref = orc.hop(v[1 - parity], gref, L, parity, s=1, tbc=1.0, recon=12)
        assert rel(res, ref) < FP64_TOL
    # full operator (A - D_s) with A = 1 + T, kappa = c_sw = 1
    A = orc.clover_fields(gref, L, 1.0, 1.0, recon=12)[0]
    inp = dev(v)
    for dag in (0, 1):
        out = q.DeviceBuffer(2 * vh * 192)
        q.lib().fullDslashQcu(out.ptr, inp.ptr, gd.ptr, ctypes.byref(prm), dag)
        res = out.to_numpy((2, vh, 4, 3))
        ref = orc.full_op(v, gref, L, 1.0, A=A, s=-1 if dag else 1, tbc=1.0, recon=12)
        assert rel(res, ref) < FP64_TOL


def test_dropin_dslash_16x32(q):
    """The headline lattice through the drop-in symbols: dslashQcu on both parities and
    fullDslashQcu with s = -1, against the oracle (periodic, kappa = c_sw = 1)."""
    L = (16, 16, 16, 32)
    gref = orc.gauge_random(L, 22, 0.2)
    prm = q.param(L)
    gd = dev(gref)
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    q.lib().loadQcuGauge(gd.ptr, ctypes.byref(prm))
    vh = orc.half_volume(L)
    v = orc.source_gauss(L, 5)
    for parity in (0, 1):
        inp = dev(v[1 - parity])
        out = q.DeviceBuffer(vh * 192)
        q.lib().dslashQcu(out.ptr, inp.ptr, gd.ptr, ctypes.byref(prm), parity)
        ref = orc.hop(v[1 - parity], gref, L, parity, s=1, tbc=1.0, recon=12)
        assert rel(out.to_numpy((vh, 4, 3)), ref) < FP64_TOL
    A = orc.clover_fields(gref, L, 1.0, 1.0, recon=12)[0]
    inp = dev(v)
    out = q.DeviceBuffer(2 * vh * 192)
    q.lib().fullDslashQcu(out.ptr, inp.ptr, gd.ptr, ctypes.byref(prm), 1)
    ref = orc.full_op(v, gref, L, 1.0, A=A, s=-1, tbc=1.0, recon=12)
    assert rel(out.to_numpy((2, vh, 4, 3)), ref) < FP64_TOL


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (6, 2, 2, 2), (8, 8, 8, 16)])
def test_dropin_dslash_bits_equal_soa_hop(q, L):
    """dslashQcu reads the interface layouts directly (one kernel); its bits equal the
    SoA hop with reconstruct-12 on the same links, including half volumes that are not
    a multiple of the 64-site wave (24 and 8 sites)."""
    gref = orc.gauge_random(L, 17, 0.3)
    prm = q.param(L)
    gd = dev(gref)
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    g = q.Gauge.from_numpy(gref, L)
    vh = orc.half_volume(L)
    v = orc.source_gauss(L, 3)
    for parity in (0, 1):
        inp = dev(v[1 - parity])
        out = q.DeviceBuffer(vh * 192)
        q.lib().dslashQcu(out.ptr, inp.ptr, gd.ptr, ctypes.byref(prm), parity)
        assert q.lib().qcuGetLastError() == 0
        soa = q.hop(g, v[1 - parity], parity, dagger=0, t_boundary=1, recon=12)
        np.testing.assert_array_equal(out.to_numpy((vh, 4, 3)), soa)
        ref = orc.hop(v[1 - parity], gref, L, parity, s=1, tbc=1.0, recon=12)
        assert rel(out.to_numpy((vh, 4, 3)), ref) < FP64_TOL
    g.destroy()


def test_dropin_gauge_changed_in_place(q):
    """The caller overwrites its gauge buffer between calls without loadQcuGauge (an
    HMC update): dslashQcu, fullDslashQcu and cg_inverter must all apply the new
    links, as the reference reads the gauge argument on every call."""
    L = (8, 8, 8, 16)
    g1 = orc.gauge_random(L, 31, 0.25)
    g2 = orc.gauge_random(L, 32, 0.25)
    prm = q.param(L)
    gd = dev(g1)
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    q.lib().loadQcuGauge(gd.ptr, ctypes.byref(prm))
    vh = orc.half_volume(L)
    v = orc.source_gauss(L, 6)
    inp_h, inp_f = dev(v[1]), dev(v)
    out_h, out_f = q.DeviceBuffer(vh * 192), q.DeviceBuffer(2 * vh * 192)
    for gref in (g1, g2, g1):
        gd.upload(gref)  # in place: same pointer, new links
        q.lib().dslashQcu(out_h.ptr, inp_h.ptr, gd.ptr, ctypes.byref(prm), 0)
        ref = orc.hop(v[1], gref, L, 0, s=1, tbc=1.0, recon=12)
        assert rel(out_h.to_numpy((vh, 4, 3)), ref) < FP64_TOL
        q.lib().fullDslashQcu(out_f.ptr, inp_f.ptr, gd.ptr, ctypes.byref(prm), 0)
        A = orc.clover_fields(gref, L, 1.0, 1.0, recon=12)[0]
        ref = orc.full_op(v, gref, L, 1.0, A=A, s=1, tbc=1.0, recon=12)
        assert rel(out_f.to_numpy((2, vh, 4, 3)), ref) < FP64_TOL
    assert q.lib().qcuGetLastError() == 0
    # the solver's clover and packed links follow the change too
    Ls = (4, 4, 4, 8)
    h1, h2 = orc.gauge_random(Ls, 33, 0.25), orc.gauge_random(Ls, 34, 0.25)
    prm = q.param(Ls)
    gd = dev(h1)
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    b = orc.source_gauss(Ls, 9)
    bd = dev(b)
    vh = orc.half_volume(Ls)
    xd = q.DeviceBuffer(2 * vh * 192)
    for gref in (h1, h2):
        gd.upload(gref)
        q.lib().cg_inverter(bd.ptr, xd.ptr, gd.ptr, ctypes.byref(prm))
        x = xd.to_numpy((2, vh, 4, 3))
        A = orc.clover_fields(gref, Ls, 1.0, 1.0, recon=12)[0]
        r = b - orc.full_op(x, gref, Ls, 1.0, A=A, s=1, tbc=1.0, recon=12)
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-11


def test_dropin_cg_inverter(q):
    L = (4, 4, 4, 8)
    gref = orc.gauge_random(L, 5, 0.25)
    prm = q.param(L)
    gd = dev(gref)
    b = orc.source_gauss(L, 9)
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    bd = dev(b)
    vh = orc.half_volume(L)
    xd = q.DeviceBuffer(2 * vh * 192)
    q.lib().cg_inverter(bd.ptr, xd.ptr, gd.ptr, ctypes.byref(prm))
    x = xd.to_numpy((2, vh, 4, 3))
    assert np.array_equal(bd.to_numpy((2, vh, 4, 3)), b)  # b untouched
    A, Ainv = orc.clover_fields(gref, L, 1.0, 1.0, recon=12)
    r = b - orc.full_op(x, gref, L, 1.0, A=A, s=1, tbc=1.0, recon=12)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-11
    xr, st = orc.cg_eo(b, gref, L, 1.0, A=A, Ainv=Ainv, tbc=1.0, recon=12, tol=1e-12, max_iter=50000)
    assert rel(x, xr) < 1e-8


# ----------------------------------------------------------------- solver ---
def test_solve_config1_wilson_fp64(q):
    """8^4 Wilson, kappa 0.12, eps 0.1, antiperiodic t, 1 Gaussian source, 1e-10."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=1, eps=0.1)
    gref = g.to_numpy()
    b = orc.source_gauss(L, 2)
    op = q.Operator(0.12, 0.0, -1, 18)
    sp = q.default_solve_params(tol=1e-10, precision=q.PREC_DOUBLE)
    x, st = q.solve(g, op, b, sp)
    assert st.converged and st.true_residual <= 1e-10
    r = b - orc.full_op(x, gref, L, 0.12, s=1, tbc=-1.0, recon=18)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-10
    xr, sr = orc.cg_eo(b, gref, L, 0.12, tbc=-1.0, recon=18, tol=1e-10)
    assert abs(st.iterations - sr.iterations) <= 2
    assert rel(x, xr) < 1e-8


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_solve_mixed_precision(q, prec, csw):
    L = (8, 8, 8, 8)
    kappa = 0.122 if csw else 0.124
    g = q.Gauge.random(L, seed=3, eps=0.25)
    gref = g.to_numpy()
    b = orc.source_gauss(L, 4)
    op = q.Operator(kappa, csw, -1, 12)
    sp = q.default_solve_params(tol=1e-10, precision=prec)
    x, st = q.solve(g, op, b, sp)
    assert st.converged and st.true_residual <= 1e-10 and st.reliable_updates >= 1
    A = orc.clover_fields(gref, L, kappa, csw, recon=12)[0] if csw else None
    r = b - orc.full_op(x, gref, L, kappa, A=A, s=1, tbc=-1.0, recon=12)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-10


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_equals_alone(q, prec):
    """Each right-hand side gives the same bits solved alone or in a batch."""
    L = (8, 8, 4, 8)
    g = q.Gauge.random(L, seed=8, eps=0.25)
    op = q.Operator(0.122, 1.0, -1, 12)
    sp = q.default_solve_params(tol=1e-10, precision=prec)
    bs = [orc.source_gauss(L, 100 + i) for i in range(3)]
    xs, sts = q.solve_batch(g, op, bs, sp)
    for i in range(3):
        xa, sa = q.solve(g, op, bs[i], sp)
        assert np.array_equal(xa, xs[i]), i
        assert sa.iterations == sts[i].iterations
        assert sts[i].true_residual <= 1e-10


@pytest.mark.parametrize("prec", [1, 2])
def test_batch_slots_finish_apart(q, prec):
    """Point and Gaussian sources interleaved in one batch, checked after every
    iteration (chunk 1), so slots converge at different harvests and the ones that
    finish together are not neighbours: each right-hand side still equals its own solve
    alone bit for bit, with the same iteration count and its true residual by the
    oracle."""
    L = (8, 8, 4, 8)
    g = q.Gauge.random(L, seed=8, eps=0.25)
    kappa = 0.122
    op = q.Operator(kappa, 1.0, -1, 12)
    sp = q.default_solve_params(tol=1e-10, precision=prec, chunk=1)
    bs = [orc.source_point(L, (0, 0, 0, 0), 0, 0), orc.source_gauss(L, 101),
          orc.source_point(L, (3, 2, 1, 5), 2, 1), orc.source_gauss(L, 102),
          orc.source_point(L, (7, 0, 3, 2), 3, 2)]
    xs, sts = q.solve_batch(g, op, bs, sp)
    its = [s.iterations for s in sts]
    assert len(set(its)) >= 2, its
    gref = g.to_numpy()
    A = orc.clover_fields(gref, L, kappa, 1.0, recon=12)[0]
    for i, b in enumerate(bs):
        xa, sa = q.solve(g, op, b, sp)
        assert np.array_equal(xa, xs[i]), (i, its)
        assert sa.iterations == sts[i].iterations and sts[i].converged
        r = b - orc.full_op(xs[i], gref, L, kappa, A=A, s=1, tbc=-1.0, recon=12)
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-10


def test_zero_source(q):
    """b = 0 gives x = 0 after no iterations with true residual 0, alone and in a batch
    beside nonzero sources, whose bits it leaves unchanged."""
    L = (8, 8, 4, 8)
    g = q.Gauge.random(L, seed=8, eps=0.25)
    op = q.Operator(0.122, 1.0, -1, 12)
    sp = q.default_solve_params(tol=1e-10, precision=1)
    b = orc.source_gauss(L, 100)
    z = np.zeros_like(b)
    x, st = q.solve(g, op, z, sp)
    assert st.converged and st.iterations == 0 and st.true_residual == 0.0
    assert not np.any(x)
    xs, sts = q.solve_batch(g, op, [b, z, b], sp)
    assert not np.any(xs[1]) and sts[1].iterations == 0 and sts[1].converged
    xa, sa = q.solve(g, op, b, sp)
    assert np.array_equal(xs[0], xa) and np.array_equal(xs[2], xa)
    assert sts[0].iterations == sa.iterations == sts[2].iterations


def test_graph_equals_eager(q):
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=6, eps=0.2)
    op = q.Operator(0.124, 0.0, -1, 18)
    b = orc.source_gauss(L, 1)
    xg, sg = q.solve(g, op, b, q.default_solve_params(tol=1e-10, precision=1, graph=1))
    xe, se = q.solve(g, op, b, q.default_solve_params(tol=1e-10, precision=1, graph=0))
    assert np.array_equal(xg, xe) and sg.iterations == se.iterations


def test_error_paths(q):
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=6, eps=0.2)
    op = q.Operator(0.124, 0.0, -1, 18)
    b = orc.source_gauss(L, 1)
    b[0, 3, 1, 2] = np.nan
    with pytest.raises(q.QcuError) as e:
        q.solve(g, op, b)
    assert e.value.code == 6
    with pytest.raises(q.QcuError) as e:
        q.Gauge.random((5, 4, 4, 4), seed=1, eps=0.1)
    assert e.value.code == 1
    b = orc.source_gauss(L, 1)
    x, st = q.solve(g, op, b, q.default_solve_params(tol=1e-12, max_iter=3), raise_on_error=False)
    assert not st.converged and st.iterations <= 8
    assert q.lib().qcuGetLastError() == 4


def test_batch_with_bad_source_leaves_no_stale_slots(q):
    """A batched solve whose second source holds a NaN fails with code 6; the slot
    already added for the first source must not stay behind, so the next clean solve
    of the same shape (same gauge, operator, precision and batch size) converges."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=8, eps=0.2)
    op = q.Operator(0.12, 1.0, -1, 12)
    sp = q.default_solve_params(tol=1e-10, precision=1)
    good = [orc.source_gauss(L, 21), orc.source_gauss(L, 22)]
    bad = [good[0], good[1].copy()]
    bad[1][1, 7, 0, 2] = np.nan
    with pytest.raises(q.QcuError) as e:
        q.solve_batch(g, op, bad, sp)
    assert e.value.code == 6
    xs, sts = q.solve_batch(g, op, good, sp)
    assert all(s.converged and s.true_residual <= 1e-10 for s in sts)
    xa, sa = q.solve(g, op, good[1], sp)
    assert np.array_equal(xs[1], xa)
    g.destroy()


# --------------------------------------------------------------- batching ---
@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("nrhs", [2, 5, 12, 16])
@pytest.mark.parametrize("L", [(8, 4, 8, 8), (16, 8, 8, 8)])
def test_hop_batch_bitwise_equals_single(q, prec, nrhs, L):
    """One launch over nrhs right-hand sides (shared link loads, several per thread,
    tiled site walk where the lattice allows it) gives the same bits as single
    launches, so results never depend on the batch."""
    g = q.Gauge.random(L, seed=31, eps=0.3)
    vs = [orc.source_gauss(L, 50 + r)[1] for r in range(nrhs)]
    outs = q.hop_batch(g, vs, 0, dagger=1, t_boundary=-1, recon=12, precision=prec)
    for r in range(nrhs):
        single = q.hop(g, vs[r], 0, dagger=1, t_boundary=-1, recon=12, precision=prec)
        assert np.array_equal(outs[r], single), r


@pytest.mark.parametrize("prec", [0, 1])
def test_tuned_shape_keeps_bits(q, prec):
    L = (8, 8, 8, 16)
    g = q.Gauge.random(L, seed=32, eps=0.3)
    vs = [orc.source_gauss(L, 60 + r)[0] for r in range(12)]
    before = q.hop_batch(g, vs, 1, recon=18, precision=prec)
    shape, ms = q.tune_hop(g, prec, 18, 12)
    assert ms > 0 and shape["block"] in (64, 128, 192, 256) and shape["rpw"] in (1, 2, 4)
    after = q.hop_batch(g, vs, 1, recon=18, precision=prec)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [None, (256, 4, 3, 1, 0, 0, 0), (128, 2, 2, 1, 0, 0, 1),
                                   (192, 3, 1, 1, 0, 0, 0, 0, 4, 2, 1), (64, 1, 1, 0, 0, 0, 1, 0, 2, 4, 0),
                                   (256, 1, 1, 1, 4, 4, 0, 0, 1, 1, 1, 1), (128, 1, 1, 0, 0, 0, 0, 0, 1, 2, 0, 1),
                                   (256, 1, 1, 0, 8, 8, 0, 0, 1, 1, 1, 2), (64, 1, 1, 1, 0, 0, 0, 0, 1, 3, 0, 2)])
def test_batch12_clover_mixed(q, shape):
    """12 right-hand sides in one batched FP32-inner clover solve (default launch
    shape, several right-hand sides per thread in either group order, the LDS kernel
    with links or links and clover staged): every one converges to 1e-10 and equals
    its own solve alone bit for bit."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=33, eps=0.25)
    if shape is not None:
        q.set_hop_shape(g, 1, 12, 12, dict(zip(q.SHAPE_KEYS, shape)))
    op = q.Operator(0.122, 1.0, -1, 12)
    sp = q.default_solve_params(tol=1e-10, precision=1)
    bs = [orc.source_point(L, (0, 0, 0, 0), s, c) for s in range(4) for c in range(3)]
    xs, sts = q.solve_batch(g, op, bs, sp)
    gref = g.to_numpy()
    A = orc.clover_fields(gref, L, 0.122, 1.0, recon=12)[0]
    for i in (0, 5, 11):
        r = bs[i] - orc.full_op(xs[i], gref, L, 0.122, A=A, s=1, tbc=-1.0, recon=12)
        assert np.linalg.norm(r) / np.linalg.norm(bs[i]) <= 1.01e-10
        xa, sa = q.solve(g, op, bs[i], sp)
        assert np.array_equal(xa, xs[i]) and sa.iterations == sts[i].iterations


@pytest.mark.parametrize("shape", [None, (256, 1, 1, 0, 8, 8, 0, 0, 1, 1, 1, 1), (64, 1, 1, 1, 0, 0, 0, 0, 1, 3, 0, 2),
                                   (256, 4, 3, 1, 0, 0, 0), (128, 1, 1, 0, 0, 0, 0, 0, 4, 1, 1)])
def test_batch12_clover_fp16(q, shape):
    """The same 12-point-source clover batch with the FP16 + per-site-norm inner loop,
    whose clover epilogues multiply in packed halves (the clover from LDS or from
    memory, several right-hand sides per thread or per wave): every right-hand side
    converges to 1e-10 by the oracle and equals its own solve alone bit for bit."""
    L = (8, 8, 8, 8)
    g = q.Gauge.random(L, seed=33, eps=0.25)
    if shape is not None:
        q.set_hop_shape(g, 2, 12, 12, dict(zip(q.SHAPE_KEYS, shape)))
    op = q.Operator(0.122, 1.0, -1, 12)
    sp = q.default_solve_params(tol=1e-10, precision=2)
    bs = [orc.source_point(L, (0, 0, 0, 0), s, c) for s in range(4) for c in range(3)]
    xs, sts = q.solve_batch(g, op, bs, sp)
    gref = g.to_numpy()
    A = orc.clover_fields(gref, L, 0.122, 1.0, recon=12)[0]
    for i in (0, 7, 11):
        r = bs[i] - orc.full_op(xs[i], gref, L, 0.122, A=A, s=1, tbc=-1.0, recon=12)
        assert np.linalg.norm(r) / np.linalg.norm(bs[i]) <= 1.01e-10
        xa, sa = q.solve(g, op, bs[i], sp)
        assert np.array_equal(xa, xs[i]) and sa.iterations == sts[i].iterations


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_hop_shapes_keep_bits(q, prec):
    """Every launch shape of a 12-RHS hop -- block size, wave groups sharing sites,
    right-hand sides per thread, XCD order, walk tile, right-hand side groups as the
    fastest or the slowest block index, the kernel with links staged in LDS -- gives
    the bits of single launches."""
    L = (16, 16, 8, 16)  # 128-site planes: whole-plane and 8-row (y, z, t) walk tiles apply
    g = q.Gauge.random(L, seed=34, eps=0.3)
    vs = [orc.source_gauss(L, 80 + r)[1] for r in range(12)]
    singles = [q.hop(g, v, 0, dagger=0, t_boundary=-1, recon=12, precision=prec) for v in vs]
    # (block, rb, rpt, xcd, tile_z, tile_t, group_major, tile_y, rpw, spt, cached[, lds])
    shapes = [(256, 4, 1, 1, 4, 4, 1, 0, 1, 1, 0), (128, 2, 3, 1, 4, 4, 1, 8, 1, 2, 1), (64, 1, 2, 0, 0, 0, 1, 0, 1, 1, 0),
              (256, 2, 1, 1, 8, 8, 0, 8, 1, 4, 1), (128, 1, 4, 0, 4, 4, 0, 0, 1, 1, 0), (256, 4, 1, 1, 2, 4, 0, 8, 1, 3, 0),
              (192, 3, 1, 1, 4, 4, 0, 0, 4, 1, 1), (64, 1, 1, 1, 0, 0, 1, 0, 4, 2, 0), (128, 2, 1, 0, 4, 4, 1, 8, 2, 4, 1),
              (256, 4, 1, 1, 0, 0, 0, 0, 4, 1, 0), (256, 2, 1, 1, 8, 8, 1, 0, 2, 2, 1),
              # the LDS-staged kernel (last entry 1): 4, 2 and 1 waves per block
              (256, 1, 1, 1, 4, 4, 0, 8, 1, 1, 0, 1), (128, 1, 1, 0, 0, 0, 0, 0, 1, 2, 1, 1),
              (64, 1, 1, 1, 8, 8, 0, 0, 1, 3, 0, 1), (256, 1, 1, 0, 2, 4, 0, 8, 1, 1, 1, 1),
              # XCD column groups 2 and 4 chunks wide (4 falls back to round-robin here)
              (256, 1, 1, 2, 8, 16, 0, 0, 1, 1, 1, 1), (128, 1, 1, 2, 0, 0, 0, 0, 1, 1, 0, 1),
              (256, 1, 1, 4, 8, 8, 0, 0, 1, 1, 1, 2)]
    for sh in shapes:
        shape = dict(zip(q.SHAPE_KEYS, sh))
        if prec == 0:
            shape["rpt"] = 1
        q.set_hop_shape(g, prec, 12, 12, shape)
        outs = q.hop_batch(g, vs, 0, dagger=0, t_boundary=-1, recon=12, precision=prec)
        for r in range(12):
            assert np.array_equal(outs[r], singles[r]), (sh, r)


def test_tune_cache_file_skips_invalid_shapes(tmp_path):
    """A QCU_TUNE_CACHE file is read by a fresh process: a valid line sets its shape, a
    line with a shape qcuSetHopShape would refuse (block 100; a load policy of 7 after a
    valid line of the same key) is skipped, and hops on the affected keys still run and
    match the oracle."""
    import os
    import subprocess
    import sys
    import textwrap

    vh = 8 * 8 * 8 * 8 // 2
    cache = tmp_path / "tune.cache"
    cache.write_text(f"hop:1:12:{vh}:1 64 1 1 0 0 0 0 0 1 1 1 0 0.1\n"
                     f"hop:1:18:{vh}:1 100 5 1 0 0 0 0 0 1 1 0 0 0.1\n"
                     f"hop:2:12:{vh}:1 128 1 1 1 0 0 0 0 1 1 0 0 0.1\n"
                     f"hop:2:12:{vh}:1 128 1 1 1 0 0 0 0 1 1 7 0 0.1\n"
                     f"not a shape line\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = textwrap.dedent("""
        import sys
        sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
        import numpy as np
        import oracle as orc
        import qcu_cg_inverter_amd as q
        q.init(0)
        L = (8, 8, 8, 8)
        g = q.Gauge.random(L, seed=4, eps=0.3)
        a = q.get_hop_shape(g, 1, 12, 1)
        b = q.get_hop_shape(g, 1, 18, 1)
        c = q.get_hop_shape(g, 2, 12, 1)
        assert (a["block"], a["cached"]) == (64, 1), a
        assert b["block"] != 100 and b["rb"] != 5, b
        assert (c["block"], c["xcd"], c["cached"]) == (128, 1, 0), c
        gref = g.to_numpy()
        v = orc.source_gauss(L, 6)
        for prec, recon, tol in ((1, 12, 1e-6), (1, 18, 1e-6), (2, 12, 3e-3)):
            out = q.hop(g, v[1], 0, dagger=0, t_boundary=-1, recon=recon, precision=prec)
            ref = orc.hop(v[1], gref, L, 0, s=1, tbc=-1.0, recon=recon)
            assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol, (prec, recon)
        print("cache ok")
    """).format(root=root, tests=os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, QCU_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "cache ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_composites_match_oracle(q, csw):
    """The reference's full-lattice compositions (cloverDslashOneRound, its
    1 - A^-1 D form, invertCloverDslash, preCloverDslash, MmV_one_round) against the
    oracle's hop and clover apply, at kappa = 0.13 and the reference normalisation."""
    L = (8, 4, 4, 8)
    g = q.Gauge.random(L, seed=35, eps=0.3)
    gref = g.to_numpy()
    v = orc.source_gauss(L, 7)
    for kappa in (0.13, 1.0):
        op = q.Operator(kappa, csw, -1, 12)
        if csw:
            A, Ainv = orc.clover_fields(gref, L, kappa, csw, recon=12)
        else:
            A = Ainv = None

        def clov(M, x, p):
            return x.copy() if M is None else orc.clover_apply(x, M[p])

        def ainv_d(x, s):
            return np.stack([kappa * clov(Ainv, orc.hop(x[1 - p], gref, L, p, s=s, tbc=-1.0, recon=12), p)
                             for p in range(2)])

        want = {q.COMP_AINV_D: (ainv_d(v, 1), ainv_d(v, -1)),
                q.COMP_ONE_MINUS_AINV_D: (v - ainv_d(v, 1), v - ainv_d(v, -1)),
                q.COMP_AINV: (np.stack([clov(Ainv, v[p], p) for p in range(2)]),) * 2,
                q.COMP_A: (np.stack([clov(A, v[p], p) for p in range(2)]),) * 2}
        mv = v - ainv_d(v, 1)
        want[q.COMP_NORMAL] = (mv - ainv_d(mv, -1),) * 2
        # A^-1 agrees to about 1e-16 times the blocks' condition number (3.2e3 at
        # kappa c_sw = 1 on this gauge, 1.4 at 0.13)
        tol = 1e-13 if not csw or kappa < 0.5 else 1e-11
        for kind, (w0, w1) in want.items():
            for dag, w in ((0, w0), (1, w1)):
                got = q.apply_composite(g, op, kind, v, dagger=dag)
                assert rel(got, w) < tol, (kappa, kind, dag)


def test_link_product_mfma_matches_valu(q):
    """The link-product measurement computes the same 12-RHS FP64 products on MFMA
    (v_mfma_f64_16x16x4f64, padded operands) as on VALU: the sums of every product
    entry agree to rounding, so its MFMA timing is of real work."""
    S = 262144 * 4
    ms_v, cs_v = q.bench_link_product(0, site_legs=S, iters=2)
    ms_m, cs_m = q.bench_link_product(1, site_legs=S, iters=2)
    assert ms_v > 0 and ms_m > 0
    # the closed form: sum over r, c, j of U[r][c] H[c][j] = sum_c (sum_r U[r][c]) (sum_j H[c][j])
    b = 1e-3 * (np.arange(S) & 1023)
    want = 0.0
    for c in range(6):
        su = sum(b + 0.01 * (6 * r + c) for r in range(6))
        sh = sum(0.5 * b + 0.001 * (j + 1) + 0.1 * c for j in range(24))
        want += float((su * sh).sum())
    assert abs(cs_v - want) <= 1e-12 * want, (cs_v, want)
    assert abs(cs_m - want) <= 1e-12 * want, (cs_m, want)
