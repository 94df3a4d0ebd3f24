"""libqcu.so against the committed golden vectors (tests/golden/golden_4x4x4x4.npz):
the five drop-in symbols, the clover export and the extended solver, on the stored
gauge and source.  FP64 operators within 1e-14 relative; solutions within what the
solve tolerance allows."""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_4x4x4x4.npz")
FP64_TOL = 1e-14


@pytest.fixture(scope="module")
def q():
    import qcu_cg_inverter_amd as q

    q.init(0)
    return q


@pytest.fixture(scope="module")
def gold():
    z = np.load(PATH, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(str(d["meta"]))
    return d


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dev(q, a):
    return q.DeviceBuffer.from_numpy(np.ascontiguousarray(a, np.complex128))


def test_dropin_symbols_on_golden(q, gold):
    L = tuple(gold["meta"]["L"])
    vh = L[0] * L[1] * L[2] * L[3] // 2
    prm = q.param(L)
    gd, b = dev(q, gold["gauge"]), gold["b"]
    q.lib().qcuClearLastError()
    q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
    q.lib().loadQcuGauge(gd.ptr, ctypes.byref(prm))
    for parity, key in ((0, "dslash_p0"), (1, "dslash_p1")):
        inp, out = dev(q, b[1 - parity]), q.DeviceBuffer(vh * 192)
        q.lib().dslashQcu(out.ptr, inp.ptr, gd.ptr, ctypes.byref(prm), parity)
        assert rel(out.to_numpy((vh, 4, 3)), gold[key]) < FP64_TOL, key
    inp = dev(q, b)
    for dag, key in ((0, "full_ref_s1"), (1, "full_ref_sm1")):
        out = q.DeviceBuffer(2 * vh * 192)
        q.lib().fullDslashQcu(out.ptr, inp.ptr, gd.ptr, ctypes.byref(prm), dag)
        assert rel(out.to_numpy((2, vh, 4, 3)), gold[key]) < FP64_TOL, key
    assert q.lib().qcuGetLastError() == 0, q.lib().qcuGetLastErrorString()
    # kappa = 1 is far past the critical kappa, so this system is ill-conditioned: it
    # takes about 1e4 CG iterations, more than cg_inverter's default cap of 20 Vh
    xd = q.DeviceBuffer(2 * vh * 192)
    sp = q.default_solve_params(tol=1e-12, precision=q.PREC_DOUBLE, max_iter=40000)
    q.check(q.lib().qcuSetDefaultSolve(ctypes.byref(sp)))
    try:
        q.lib().cg_inverter(inp.ptr, xd.ptr, gd.ptr, ctypes.byref(prm))
        assert q.lib().qcuGetLastError() == 0, q.lib().qcuGetLastErrorString()
    finally:
        sp.max_iter = 0
        q.check(q.lib().qcuSetDefaultSolve(ctypes.byref(sp)))
    # the solution agrees to what 1e-12 residuals allow
    assert rel(xd.to_numpy((2, vh, 4, 3)), gold["x_ref"]) < 1e-8


def test_dropin_cg_inverter_cap_reports_failure(q, gold):
    """Stopping at the iteration cap leaves the last iterate and reports, as the
    reference's cg_failed!!! path does, without ending the process."""
    L = tuple(gold["meta"]["L"])
    vh = L[0] * L[1] * L[2] * L[3] // 2
    prm = q.param(L)
    gd, inp = dev(q, gold["gauge"]), dev(q, gold["b"])
    xd = q.DeviceBuffer(2 * vh * 192)
    sp = q.default_solve_params(tol=1e-12, precision=q.PREC_DOUBLE, max_iter=50)
    q.check(q.lib().qcuSetDefaultSolve(ctypes.byref(sp)))
    q.lib().qcuClearLastError()
    try:
        q.lib().initGridSize(ctypes.byref(q.grid()), ctypes.byref(prm), gd.ptr, None, None)
        q.lib().cg_inverter(inp.ptr, xd.ptr, gd.ptr, ctypes.byref(prm))
        assert q.lib().qcuGetLastError() == 4  # QCU_ERROR_NOT_CONVERGED
    finally:
        sp.max_iter = 0
        q.check(q.lib().qcuSetDefaultSolve(ctypes.byref(sp)))
        q.lib().qcuClearLastError()
    x = xd.to_numpy((2, vh, 4, 3))
    assert np.isfinite(x).all() and np.linalg.norm(x) > 0


def test_clover_export_on_golden(q, gold):
    m = gold["meta"]
    L = tuple(m["L"])
    g = q.Gauge.from_numpy(gold["gauge"], L)
    A, Ainv, nf = g.clover(m["ref_kappa"], m["ref_csw"])
    # at kappa * c_sw = 1 some 6x6 blocks are indefinite: Cholesky fails on them and
    # the pivoted Gauss-Jordan fallback inverts them; 32 of the 512 blocks here, as
    # the oracle counts
    assert nf == 32
    assert rel(A, gold["clover_A_ref"]) < FP64_TOL
    assert rel(Ainv, gold["clover_Ainv_ref"]) < 1e-13


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-11), (2, 1e-11)])
def test_solves_on_golden(q, gold, prec, tol):
    """FP64 and mixed (FP32, FP16 inner) solves; the FP64 true residual bounds the
    distance to the stored solution (both operators are well conditioned)."""
    m = gold["meta"]
    L = tuple(m["L"])
    g = q.Gauge.from_numpy(gold["gauge"], L)
    b = gold["b"]
    sp = q.default_solve_params(tol=tol, precision=prec)
    out = q.apply_m(g, q.Operator(m["wilson_kappa"], 0.0, -1, 18), b)
    assert rel(out, gold["wilson_full"]) < FP64_TOL
    for key, op in (("x_wilson", q.Operator(m["wilson_kappa"], 0.0, -1, 18)),
                    ("x_clover", q.Operator(m["clover_kappa"], m["clover_csw"], -1, 12))):
        x, st = q.solve(g, op, b, sp)
        assert st.converged and st.true_residual < tol, (key, st.as_dict())
        assert rel(x, gold[key]) < 100 * tol, key
