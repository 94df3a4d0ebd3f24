"""GPU parity of the layout converters (qcuLexToEvenOdd / qcuEvenOddToLex and the
reference's shiftVectorStorage / shiftGaugeStorage) against the oracle: pure data
movement is bit-exact; the rebuilt link row 2 is FP64 arithmetic (1e-15)."""
import ctypes

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

# half volumes that are and are not multiples of the 64-site tile
LATTICES = [(4, 4, 4, 4), (6, 3, 5, 2), (2, 2, 2, 2), (8, 8, 8, 8), (16, 16, 16, 32)]


@pytest.fixture(scope="module")
def q():
    import qcu_cg_inverter_amd as q

    q.init(0)
    return q


@pytest.mark.parametrize("L", LATTICES)
def test_lex_eo_spinor_bit_exact(q, L):
    V = orc.volume(L)
    lex = orc.source_gauss(L, 17).reshape(V, 4, 3)  # any data: its order is lexicographic here
    eo = q.lex_to_eo(lex, L, q.FIELD_SPINOR)
    ref = orc.lex_eo(lex.reshape(1, V, 12), L, 12, 1, True).reshape(V, 4, 3)
    np.testing.assert_array_equal(eo, ref)
    np.testing.assert_array_equal(q.eo_to_lex(eo, L, q.FIELD_SPINOR), lex)


@pytest.mark.parametrize("L", LATTICES)
def test_lex_eo_gauge_bit_exact(q, L):
    V = orc.volume(L)
    g = orc.gauge_random(L, 3, 0.3).reshape(4, V, 3, 3)
    eo = q.lex_to_eo(g, L, q.FIELD_GAUGE)
    ref = orc.lex_eo(g.reshape(4, V, 9), L, 9, 4, True).reshape(4, V, 3, 3)
    np.testing.assert_array_equal(eo, ref)
    np.testing.assert_array_equal(q.eo_to_lex(eo, L, q.FIELD_GAUGE), g)


@pytest.mark.parametrize("L", LATTICES)
def test_shift_vector_storage_bit_exact(q, L):
    v = orc.source_gauss(L, 9)[1]
    soa = q.shift_vector_storage(v, q.TO_COALESCE, L)
    np.testing.assert_array_equal(soa, orc.shift_vector(v, 0, L))
    back = q.shift_vector_storage(soa, q.TO_NON_COALESCE, L)
    np.testing.assert_array_equal(back, v)


@pytest.mark.parametrize("L", LATTICES)
def test_shift_gauge_storage(q, L):
    g = orc.gauge_random(L, 21, 0.5)
    soa = q.shift_gauge_storage(g, q.TO_COALESCE, L)
    np.testing.assert_array_equal(soa, orc.shift_gauge(g, 0, L))
    back = q.shift_gauge_storage(soa, q.TO_NON_COALESCE, L)
    np.testing.assert_array_equal(back[..., :2, :], g[..., :2, :])
    assert np.abs(back - orc.shift_gauge(soa, 1, L)).max() < 1e-15
    assert np.abs(back - g).max() < 1e-14


def test_converted_gauge_drives_the_hop(q):
    """A lexicographic gauge and source converted on the device give the same hop as
    the oracle on the even-odd originals."""
    L = (8, 4, 4, 8)
    V = orc.volume(L)
    g_eo = orc.gauge_random(L, 2, 0.3)
    v_eo = orc.source_gauss(L, 4)
    g_lex = orc.lex_eo(g_eo.reshape(4, V, 9), L, 9, 4, False).reshape(4, V, 3, 3)
    v_lex = orc.lex_eo(v_eo.reshape(1, V, 12), L, 12, 1, False).reshape(V, 4, 3)
    g = q.Gauge.from_numpy(q.lex_to_eo(g_lex, L, q.FIELD_GAUGE).reshape(4, 2, V // 2, 3, 3), L)
    v = q.lex_to_eo(v_lex, L, q.FIELD_SPINOR).reshape(2, V // 2, 4, 3)
    out = q.hop(g, v[1], 0, recon=18)
    ref = orc.hop(v_eo[1], g_eo, L, 0, s=1, tbc=1.0, recon=18)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-14


def test_converter_errors(q, monkeypatch):
    L = (4, 4, 4, 4)
    V = orc.volume(L)
    buf = q.DeviceBuffer(V * 192 * 2)
    lib = q.lib()
    # overlapping src and dst, bad field kind: error codes, nothing launched
    assert lib.qcuLexToEvenOdd(buf.ptr, buf.ptr + 16, q._L(L), q.FIELD_SPINOR) == 1
    assert lib.qcuEvenOddToLex(buf.ptr + V * 192, buf.ptr, q._L(L), 7) == 1
    assert lib.qcuLexToEvenOdd(buf.ptr + V * 192, buf.ptr, q._L((3, 4, 4, 4)), q.FIELD_SPINOR) == 1
    # the void storage shifts report through qcuGetLastError when exit is switched off
    monkeypatch.setenv("QCU_EXIT_ON_ERROR", "0")
    lib.qcuClearLastError()
    lib.shiftVectorStorage(buf.ptr + V * 192, buf.ptr, 5, *L)
    assert lib.qcuGetLastError() == 1
    lib.qcuClearLastError()
    lib.shiftGaugeStorage(buf.ptr, buf.ptr, 0, *L)
    assert lib.qcuGetLastError() == 1
    assert b"overlap" in lib.qcuGetLastErrorString()
    lib.qcuClearLastError()
