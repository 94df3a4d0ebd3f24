"""Parity at the largest configurations' volumes (32^3x64, 2^20 sites per parity, and
24^3x48): the FP64 hop against the oracle, the 12-right-hand-side launch against 12
single launches bit for bit, the s = +1 / s = -1 adjointness of the hops, and the
true residuals of FP64 and mixed-precision solves recomputed by the oracle."""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

L = (32, 32, 32, 64)


@pytest.fixture(scope="module")
def q():
    import qcu_cg_inverter_amd as q

    q.init(0)
    return q


@pytest.fixture(scope="module")
def gauge(q):
    g = q.Gauge.random(L, seed=4, eps=0.25)
    yield g
    g.destroy()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_hop_f64_32x64_matches_oracle(q, gauge):
    gref = gauge.to_numpy()
    v = orc.source_gauss(L, 11)
    for parity, dag, recon in ((0, 0, 12), (1, 1, 18)):
        out = q.hop(gauge, v[1 - parity], parity, dagger=dag, t_boundary=-1, recon=recon)
        ref = orc.hop(v[1 - parity], gref, L, parity, s=-1 if dag else 1, tbc=-1.0, recon=recon)
        assert rel(out, ref) < 1e-14, (parity, dag, recon)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_hop_12rhs_32x64_equals_single_launches(q, gauge, prec):
    """12 right-hand sides in one launch give the bits of 12 single launches; at this
    volume the batch's fields are stored nontemporally and a single FP32 / FP16 field
    is not, so this also checks that the store policy leaves the values alone."""
    vh = q.half_volume(L)
    rng = np.random.default_rng(5)
    vs = [(rng.standard_normal((vh, 4, 3)) + 1j * rng.standard_normal((vh, 4, 3))) for _ in range(12)]
    outs = q.hop_batch(gauge, vs, 0, dagger=1, t_boundary=-1, recon=12, precision=prec)
    for r in (0, 5, 11):
        single = q.hop(gauge, vs[r], 0, dagger=1, t_boundary=-1, recon=12, precision=prec)
        np.testing.assert_array_equal(outs[r], single)


def test_hop_adjointness_32x64(q, gauge):
    """<phi_e, D_{s=+1} psi_o> = <D_{s=-1} phi_e, psi_o>: a property of the operator
    that holds at any volume."""
    vh = q.half_volume(L)
    rng = np.random.default_rng(6)
    phi = rng.standard_normal((vh, 4, 3)) + 1j * rng.standard_normal((vh, 4, 3))
    psi = rng.standard_normal((vh, 4, 3)) + 1j * rng.standard_normal((vh, 4, 3))
    lhs = np.vdot(phi, q.hop(gauge, psi, 0, dagger=0, t_boundary=-1, recon=12))
    rhs = np.vdot(q.hop(gauge, phi, 1, dagger=1, t_boundary=-1, recon=12), psi)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_mixed_solve_32x64_true_residual_by_oracle(q, gauge):
    """FP32-inner Wilson solve on the full volume; the oracle recomputes its residual."""
    gref = gauge.to_numpy()
    b = orc.source_gauss(L, 12)
    op = q.Operator(0.122, 0.0, -1, 12)
    x, st = q.solve(gauge, op, b, q.default_solve_params(tol=1e-10, precision=1))
    assert st.converged and st.true_residual <= 1e-10
    r = b - orc.full_op(x, gref, L, 0.122, s=1, tbc=-1.0, recon=12)
    res = np.linalg.norm(r) / np.linalg.norm(b)
    assert res <= 1e-10
    assert abs(res - st.true_residual) <= 1e-3 * st.true_residual


def test_fp64_solve_32x64_true_residual_by_oracle(q, gauge):
    """Plain FP64 clover solve on the full volume, whose vector kernels store their
    201 MB fields nontemporally; the oracle recomputes the residual with its own
    clover field."""
    gref = gauge.to_numpy()
    b = orc.source_gauss(L, 13)
    x, st = q.solve(gauge, q.Operator(0.122, 1.0, -1, 12), b, q.default_solve_params(tol=1e-10, precision=0))
    assert st.converged and st.true_residual <= 1e-10
    A = orc.clover_fields(gref, L, 0.122, 1.0, recon=12)[0]
    res = np.linalg.norm(b - orc.full_op(x, gref, L, 0.122, A=A, s=1, tbc=-1.0, recon=12)) / np.linalg.norm(b)
    assert res <= 1e-10
    assert abs(res - st.true_residual) <= 1e-3 * st.true_residual


def test_config4_batch_true_residual_by_oracle(q, gauge):
    """The largest configuration as run: 12 point sources on 32^3x64 clover (c_sw 1,
    kappa 0.122, antiperiodic), one batched FP32-inner solve; the oracle builds its own
    clover field and recomputes the residual of three of the solutions."""
    vh = q.half_volume(L)
    gref = gauge.to_numpy()
    kappa, csw = 0.122, 1.0
    srcs = [q.source_point_device(L, (0, 0, 0, 0), s, c) for s in range(4) for c in range(3)]
    xs = [q.DeviceBuffer(2 * vh * 192) for _ in range(12)]
    sts = q.solve_batch_device(gauge, q.Operator(kappa, csw, -1, 12), xs, srcs,
                               q.default_solve_params(tol=1e-10, precision=1))
    assert all(s.converged and s.true_residual <= 1e-10 for s in sts)
    A = orc.clover_fields(gref, L, kappa, csw, recon=12)[0]
    for r in (0, 5, 11):
        b = srcs[r].to_numpy((2, vh, 4, 3))
        x = xs[r].to_numpy((2, vh, 4, 3))
        res = np.linalg.norm(b - orc.full_op(x, gref, L, kappa, A=A, s=1, tbc=-1.0, recon=12)) / np.linalg.norm(b)
        assert res <= 1e-10, (r, res)
        assert abs(res - sts[r].true_residual) <= 1e-3 * sts[r].true_residual


def test_config3_fp16_solve_true_residual_by_oracle(q):
    """24^3x48 clover (c_sw 1, kappa 0.122, eps 0.25), one Gaussian source, FP16 +
    per-site-norm inner loop with FP64 reliable updates; the oracle recomputes the
    residual with its own clover field."""
    L3 = (24, 24, 24, 48)
    g = q.Gauge.random(L3, seed=3, eps=0.25)
    gref = g.to_numpy()
    b = orc.source_gauss(L3, 2)
    x, st = q.solve(g, q.Operator(0.122, 1.0, -1, 12), b, q.default_solve_params(tol=1e-10, precision=2))
    assert st.converged and st.true_residual <= 1e-10
    A = orc.clover_fields(gref, L3, 0.122, 1.0, recon=12)[0]
    res = np.linalg.norm(b - orc.full_op(x, gref, L3, 0.122, A=A, s=1, tbc=-1.0, recon=12)) / np.linalg.norm(b)
    assert res <= 1e-10
    assert abs(res - st.true_residual) <= 1e-3 * st.true_residual
    g.destroy()


def test_config2_fp64_and_mixed_solves_by_oracle(q):
    """Configuration 2: 16^3x32 Wilson, kappa 0.124, eps 0.2, antiperiodic t, one
    Gaussian source (the benchmark's seeds: gauge 1, source 2), tol 1e-10.  The plain
    FP64 solve (18-real links) and the FP32-inner mixed solve with FP64 reliable
    updates (reconstruct-12) both reach the tolerance, and the oracle's operator
    recomputes each true residual.  Both solve the same system, so their solutions
    agree to the tolerance times the condition number's share."""
    L2 = (16, 16, 16, 32)
    g = q.Gauge.random(L2, seed=1, eps=0.2)
    gref = g.to_numpy()
    b = q.source_gaussian(L2, seed=2)
    xs = {}
    for name, prec, recon in (("fp64", 0, 18), ("mixed_fp32", 1, 12)):
        op = q.Operator(0.124, 0.0, -1, recon)
        x, st = q.solve(g, op, b, q.default_solve_params(tol=1e-10, precision=prec))
        assert st.converged and st.true_residual <= 1e-10, (name, st.as_dict())
        r = b - orc.full_op(x, gref, L2, 0.124, s=1, tbc=-1.0, recon=recon)
        res = np.linalg.norm(r) / np.linalg.norm(b)
        assert res <= 1.001e-10, (name, res)
        assert abs(res - st.true_residual) <= 1e-3 * st.true_residual, (name, res, st.true_residual)
        xs[name] = x
    assert rel(xs["mixed_fp32"], xs["fp64"]) < 1e-8
    g.destroy()
