"""qcuHopNoAlloc, the hop that allocates no device memory, against qcuHop through the
C ABI: bit-identical output on two calls in a row, the second on a smaller lattice,
with the FP64 result also checked against the oracle (1e-14 relative).  The size
limit of its static scratch and the missing-links precondition come back as error
codes, and the calls leave the device's free memory unchanged."""
import ctypes

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

SHAPES = [(0, 18), (0, 12), (1, 18), (1, 12), (2, 18), (2, 12)]  # (precision, recon)


@pytest.fixture(scope="module")
def q():
    import qcu_cg_inverter_amd as q

    q.init(0)
    return q


def free_bytes(q):
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    q.check(q.lib().qcuMemInfo(ctypes.byref(f), ctypes.byref(t)))
    return f.value


def hop_noalloc(q, g, a, o, parity, dagger, tb, recon, prec):
    return q.lib().qcuHopNoAlloc(g.handle, o.ptr, a.ptr, parity, dagger, tb, recon, prec)


def last_error(q):
    return q.lib().qcuGetLastErrorString().decode()


def test_noalloc_equals_hop_two_calls_second_smaller(q):
    big, small = (16, 16, 16, 32), (8, 8, 8, 8)
    cases = []
    for L, seed in ((big, 1), (small, 2)):
        g = q.Gauge.random(L, seed=seed, eps=0.2)
        v = orc.source_gauss(L, seed + 10)
        refs = {}
        for prec, recon in SHAPES:  # qcuHop packs the links the no-alloc call then uses
            for dag in (0, 1):
                refs[(prec, recon, dag)] = q.hop(g, v[1], 0, dagger=dag, t_boundary=-1, recon=recon, precision=prec)
        a = q.DeviceBuffer.from_numpy(np.ascontiguousarray(v[1], np.complex128))
        o = q.DeviceBuffer(g.vh * 192)
        cases.append((L, g, v, refs, a, o))
    f0 = free_bytes(q)
    for L, g, v, refs, a, o in cases:  # the large lattice first, then the small one
        for prec, recon in SHAPES:
            for dag in (0, 1):
                print(f"noalloc L={L} prec={prec} recon={recon} dagger={dag}", flush=True)
                rc = hop_noalloc(q, g, a, o, 0, dag, -1, recon, prec)
                assert rc == 0, last_error(q)
                out = o.to_numpy((g.vh, 4, 3))
                assert np.array_equal(out, refs[(prec, recon, dag)]), (L, prec, recon, dag)
                if prec == 0:
                    ref = orc.hop(v[1], g.to_numpy(), L, 0, s=-1 if dag else 1, tbc=-1.0, recon=recon)
                    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-14
    # nothing was allocated on the device by the no-alloc calls
    assert f0 - free_bytes(q) < (1 << 20)
    for c in cases:
        c[1].destroy()


def test_noalloc_rejects_a_lattice_beyond_its_scratch(q):
    L = (16, 16, 16, 48)  # FP64: two fields of 98,304 half sites need 37.7 MB > 32 MiB
    g = q.Gauge.random(L, seed=3, eps=0.2)
    v = orc.source_gauss(L, 4)
    ref64 = q.hop(g, v[1], 1, t_boundary=1, recon=12, precision=0)
    ref16 = q.hop(g, v[1], 1, t_boundary=1, recon=12, precision=2)
    a = q.DeviceBuffer.from_numpy(np.ascontiguousarray(v[1], np.complex128))
    o = q.DeviceBuffer(g.vh * 192)
    rc = hop_noalloc(q, g, a, o, 1, 0, 1, 12, 0)
    assert rc == 1 and "maximum" in last_error(q)
    # the same lattice in FP16 + norm (52 B per site) fits
    rc = hop_noalloc(q, g, a, o, 1, 0, 1, 12, 2)
    assert rc == 0, last_error(q)
    assert np.array_equal(o.to_numpy((g.vh, 4, 3)), ref16)
    assert ref64 is not None
    g.destroy()


def test_noalloc_needs_packed_links(q):
    L = (4, 4, 4, 4)
    g = q.Gauge.random(L, seed=5, eps=0.3)
    v = orc.source_gauss(L, 6)
    a = q.DeviceBuffer.from_numpy(np.ascontiguousarray(v[0], np.complex128))
    o = q.DeviceBuffer(g.vh * 192)
    rc = hop_noalloc(q, g, a, o, 1, 1, 1, 18, 1)
    assert rc == 7 and "packed links" in last_error(q)
    ref = q.hop(g, v[0], 1, dagger=1, t_boundary=1, recon=18, precision=1)
    rc = hop_noalloc(q, g, a, o, 1, 1, 1, 18, 1)
    assert rc == 0, last_error(q)
    assert np.array_equal(o.to_numpy((g.vh, 4, 3)), ref)
    g.destroy()
