"""qcu_cg_inverter_amd -- MI355X-native even-odd Wilson / Wilson-clover inverter.

The product is ``libqcu.so`` (hand-written gfx950 HIP kernels behind a C ABI, see
``include/qcu.h`` and ``include/qcu_ex.h``).  This module is its ctypes binding: it
loads the in-tree library, declares every exported function and offers numpy
conveniences (device buffers, gauge handles, solves) for tests and benchmarks.

There is no fallback: if ``libqcu.so`` is missing, importing the binding raises.
Build it with ``python -m qcu_cg_inverter_amd.build``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqcu.so")
ROOT = os.path.dirname(HERE)

PREC_DOUBLE, PREC_SINGLE, PREC_HALF = 0, 1, 2

ERRORS = {
    0: "QCU_SUCCESS",
    1: "QCU_ERROR_INVALID_VALUE",
    2: "QCU_ERROR_HIP",
    3: "QCU_ERROR_OUT_OF_MEMORY",
    4: "QCU_ERROR_NOT_CONVERGED",
    5: "QCU_ERROR_BREAKDOWN",
    6: "QCU_ERROR_NOT_FINITE",
    7: "QCU_ERROR_STATE",
    8: "QCU_ERROR_ADMISSION",
}


class QcuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class QcuParam(ctypes.Structure):
    _fields_ = [("lattice_size", ctypes.c_int * 4)]


class QcuGrid(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_int * 4)]


class Operator(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_double), ("csw", ctypes.c_double), ("t_boundary", ctypes.c_int),
                ("recon", ctypes.c_int)]


class SolveParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("precision", ctypes.c_int),
                ("delta", ctypes.c_double), ("graph", ctypes.c_int), ("chunk", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("reliable_updates", ctypes.c_int), ("converged", ctypes.c_int),
                ("breakdown", ctypes.c_int), ("true_residual", ctypes.c_double), ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_batch", ctypes.c_int), ("precision", ctypes.c_int),
                ("chunk", ctypes.c_int), ("graph", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("gauge", ctypes.c_void_p), ("kappa", ctypes.c_double), ("csw", ctypes.c_double),
                ("tol", ctypes.c_double), ("t_boundary", ctypes.c_int), ("recon", ctypes.c_int),
                ("nrhs", ctypes.c_int), ("x", ctypes.POINTER(ctypes.c_void_p)), ("b", ctypes.POINTER(ctypes.c_void_p)),
                ("seed", ctypes.c_uint64), ("max_iter", ctypes.c_int)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("done", ctypes.c_int), ("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("true_residual", ctypes.c_double), ("queue_seconds", ctypes.c_double),
                ("latency_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_sz = ctypes.c_size_t
_i4 = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); every symbol declared in include/*.h
PROTOTYPES = {
    "initGridSize": (None, [ctypes.POINTER(QcuGrid), ctypes.POINTER(QcuParam), _vp, _vp, _vp]),
    "dslashQcu": (None, [_vp, _vp, _vp, ctypes.POINTER(QcuParam), _int]),
    "fullDslashQcu": (None, [_vp, _vp, _vp, ctypes.POINTER(QcuParam), _int]),
    "cg_inverter": (None, [_vp, _vp, _vp, ctypes.POINTER(QcuParam)]),
    "loadQcuGauge": (None, [_vp, ctypes.POINTER(QcuParam)]),
    "qcuInit": (_int, [_int]),
    "qcuFinalize": (_int, []),
    "qcuGetLastError": (_int, []),
    "qcuGetLastErrorString": (ctypes.c_char_p, []),
    "qcuClearLastError": (None, []),
    "qcuVersion": (ctypes.c_char_p, []),
    "qcuMalloc": (_int, [ctypes.POINTER(_vp), _sz]),
    "qcuFree": (_int, [_vp]),
    "qcuMemcpyHtoD": (_int, [_vp, _vp, _sz]),
    "qcuMemcpyDtoH": (_int, [_vp, _vp, _sz]),
    "qcuMemcpyDtoD": (_int, [_vp, _vp, _sz]),
    "qcuMemset": (_int, [_vp, _int, _sz]),
    "qcuSynchronize": (_int, []),
    "qcuMemInfo": (_int, [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "qcuGaugeCreate": (_int, [ctypes.POINTER(_vp), _vp, _i4]),
    "qcuGaugeCreateRandom": (_int, [ctypes.POINTER(_vp), _i4, ctypes.c_uint64, _dbl]),
    "qcuGaugeDestroy": (_int, [_vp]),
    "qcuGaugeExport": (_int, [_vp, _vp]),
    "qcuCloverExport": (_int, [_vp, ctypes.POINTER(Operator), _vp, _vp, _i4]),
    "qcuSourceGaussian": (_int, [_vp, _i4, ctypes.c_uint64]),
    "qcuSourcePoint": (_int, [_vp, _i4, _i4, _int, _int]),
    "qcuLexToEvenOdd": (_int, [_vp, _vp, _i4, _int]),
    "qcuEvenOddToLex": (_int, [_vp, _vp, _i4, _int]),
    "shiftVectorStorage": (None, [_vp, _vp, _int, _int, _int, _int, _int]),
    "shiftGaugeStorage": (None, [_vp, _vp, _int, _int, _int, _int, _int]),
    "qcuHop": (_int, [_vp, _vp, _vp, _int, _int, _int, _int, _int]),
    "qcuHopBatch": (_int, [_vp, _int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _int, _int, _int, _int, _int]),
    "qcuHopNoAlloc": (_int, [_vp, _vp, _vp, _int, _int, _int, _int, _int]),
    "qcuApplyM": (_int, [_vp, ctypes.POINTER(Operator), _vp, _vp, _int]),
    "qcuApplyMhat": (_int, [_vp, ctypes.POINTER(Operator), _vp, _vp, _int, _int]),
    "qcuApplyComposite": (_int, [_vp, ctypes.POINTER(Operator), _int, _vp, _vp, _int]),
    "qcuDefaultSolveParams": (None, [ctypes.POINTER(SolveParams)]),
    "qcuSetDefaultSolve": (_int, [ctypes.POINTER(SolveParams)]),
    "qcuSolve": (_int, [_vp, ctypes.POINTER(Operator), ctypes.POINTER(SolveParams), _vp, _vp,
                        ctypes.POINTER(SolveStats)]),
    "qcuSolveBatch": (_int, [_vp, ctypes.POINTER(Operator), ctypes.POINTER(SolveParams), _int,
                             ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(SolveStats)]),
    "qcuBenchHop": (_int, [_vp, _int, _int, _int, _int, ctypes.POINTER(_dbl)]),
    "qcuBenchCopy": (_int, [_sz, _int, ctypes.POINTER(_dbl)]),
    "qcuBenchLinkProduct": (_int, [_int, ctypes.c_long, _int, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]),
    "qcuTuneHop": (_int, [_vp, _int, _int, _int, _i4, ctypes.POINTER(_dbl)]),
    "qcuSetHopShape": (_int, [_vp, _int, _int, _int, _i4]),
    "qcuGetHopShape": (_int, [_vp, _int, _int, _int, _i4]),
    "qcuDefaultServerConfig": (None, [ctypes.POINTER(ServerConfig)]),
    "qcuServerCreate": (_int, [ctypes.POINTER(_vp), ctypes.POINTER(ServerConfig)]),
    "qcuServerSubmit": (_int, [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    "qcuServerPoll": (_int, [_vp, ctypes.c_uint64, ctypes.POINTER(RequestStatus)]),
    "qcuServerWait": (_int, [_vp, ctypes.c_uint64, ctypes.POINTER(RequestStatus)]),
    "qcuServerDrain": (_int, [_vp]),
    "qcuServerUsage": (_int, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz), _i4]),
    "qcuServerDestroy": (_int, [_vp]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded libqcu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `python -m qcu_cg_inverter_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().qcuGetLastErrorString()
        raise QcuError(rc, msg.decode() if msg else "")
    return rc


def _L(L):
    return (ctypes.c_int * 4)(*[int(v) for v in L])


def half_volume(L):
    return int(L[0]) * int(L[1]) * int(L[2]) * int(L[3]) // 2


class DeviceBuffer:
    """A device allocation made through libqcu.so (qcuMalloc / qcuFree)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().qcuMalloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().qcuMemcpyHtoD(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_numpy(self, shape, dtype=np.complex128) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().qcuMemcpyDtoH(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().qcuMemcpyHtoD(self.ptr, a.ctypes.data, a.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib().qcuFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Gauge:
    """A gauge field resident in HBM (qcuGauge handle)."""

    def __init__(self, handle, L):
        self.handle = handle
        self.L = tuple(int(v) for v in L)

    @classmethod
    def from_numpy(cls, g: np.ndarray, L) -> "Gauge":
        buf = DeviceBuffer.from_numpy(np.ascontiguousarray(g, np.complex128))
        h = _vp()
        check(lib().qcuGaugeCreate(ctypes.byref(h), buf.ptr, _L(L)))
        buf.free()
        return cls(h.value, L)

    @classmethod
    def random(cls, L, seed: int, eps: float) -> "Gauge":
        h = _vp()
        check(lib().qcuGaugeCreateRandom(ctypes.byref(h), _L(L), seed, eps))
        return cls(h.value, L)

    @property
    def vh(self):
        return half_volume(self.L)

    def to_numpy(self) -> np.ndarray:
        buf = DeviceBuffer(4 * 2 * self.vh * 9 * 16)
        check(lib().qcuGaugeExport(self.handle, buf.ptr))
        return buf.to_numpy((4, 2, self.vh, 3, 3))

    def clover(self, kappa, csw):
        """(A, A^-1, nfallback) as complex128 [2][Vh][2][6][6]."""
        op = Operator(kappa, csw, 1, 18)
        n = 2 * self.vh * 72 * 16
        a, ai = DeviceBuffer(n), DeviceBuffer(n)
        nf = ctypes.c_int()
        check(lib().qcuCloverExport(self.handle, ctypes.byref(op), a.ptr, ai.ptr, ctypes.byref(nf)))
        shp = (2, self.vh, 2, 6, 6)
        return a.to_numpy(shp), ai.to_numpy(shp), nf.value

    def destroy(self):
        if self.handle:
            lib().qcuGaugeDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def init(device: int = 0):
    check(lib().qcuInit(device))


def source_gaussian(L, seed) -> np.ndarray:
    vh = half_volume(L)
    buf = DeviceBuffer(2 * vh * 12 * 16)
    check(lib().qcuSourceGaussian(buf.ptr, _L(L), seed))
    return buf.to_numpy((2, vh, 4, 3))


def hop(g: Gauge, inp: np.ndarray, parity: int, dagger=0, t_boundary=1, recon=12, precision=PREC_DOUBLE):
    vh = g.vh
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(inp, np.complex128))
    o = DeviceBuffer(vh * 192)
    check(lib().qcuHop(g.handle, o.ptr, a.ptr, parity, dagger, t_boundary, recon, precision))
    return o.to_numpy((vh, 4, 3))


def hop_batch(g: Gauge, inps, parity: int, dagger=0, t_boundary=1, recon=12, precision=PREC_DOUBLE):
    vh = g.vh
    n = len(inps)
    ins = [DeviceBuffer.from_numpy(np.ascontiguousarray(v, np.complex128)) for v in inps]
    outs = [DeviceBuffer(vh * 192) for _ in range(n)]
    ip = (_vp * n)(*[b.ptr for b in ins])
    op_ = (_vp * n)(*[b.ptr for b in outs])
    check(lib().qcuHopBatch(g.handle, n, op_, ip, parity, dagger, t_boundary, recon, precision))
    return [o.to_numpy((vh, 4, 3)) for o in outs]


def apply_m(g: Gauge, op: Operator, inp: np.ndarray, dagger=0):
    vh = g.vh
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(inp, np.complex128))
    o = DeviceBuffer(2 * vh * 192)
    check(lib().qcuApplyM(g.handle, ctypes.byref(op), o.ptr, a.ptr, dagger))
    return o.to_numpy((2, vh, 4, 3))


COMP_AINV_D, COMP_ONE_MINUS_AINV_D, COMP_AINV, COMP_A, COMP_NORMAL = range(5)


def apply_composite(g: Gauge, op: Operator, kind: int, inp: np.ndarray, dagger=0):
    """Full-lattice composition of the reference's clover path (qcuApplyComposite)."""
    vh = g.vh
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(inp, np.complex128))
    o = DeviceBuffer(2 * vh * 192)
    check(lib().qcuApplyComposite(g.handle, ctypes.byref(op), kind, o.ptr, a.ptr, dagger))
    return o.to_numpy((2, vh, 4, 3))


def apply_mhat(g: Gauge, op: Operator, inp: np.ndarray, dagger=0, precision=PREC_DOUBLE):
    vh = g.vh
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(inp, np.complex128))
    o = DeviceBuffer(vh * 192)
    check(lib().qcuApplyMhat(g.handle, ctypes.byref(op), o.ptr, a.ptr, dagger, precision))
    return o.to_numpy((vh, 4, 3))


def default_solve_params(**kw) -> SolveParams:
    sp = SolveParams()
    lib().qcuDefaultSolveParams(ctypes.byref(sp))
    for k, v in kw.items():
        setattr(sp, k, v)
    return sp


def solve(g: Gauge, op: Operator, b: np.ndarray, params: SolveParams | None = None, raise_on_error=True):
    sp = params or default_solve_params()
    vh = g.vh
    bb = DeviceBuffer.from_numpy(np.ascontiguousarray(b, np.complex128))
    xb = DeviceBuffer(2 * vh * 192)
    st = SolveStats()
    rc = lib().qcuSolve(g.handle, ctypes.byref(op), ctypes.byref(sp), xb.ptr, bb.ptr, ctypes.byref(st))
    if raise_on_error:
        check(rc)
    return xb.to_numpy((2, vh, 4, 3)), st


def solve_batch(g: Gauge, op: Operator, bs, params: SolveParams | None = None, raise_on_error=True):
    sp = params or default_solve_params()
    vh = g.vh
    n = len(bs)
    bbufs = [DeviceBuffer.from_numpy(np.ascontiguousarray(b, np.complex128)) for b in bs]
    xbufs = [DeviceBuffer(2 * vh * 192) for _ in range(n)]
    xs = (_vp * n)(*[x.ptr for x in xbufs])
    bp = (_vp * n)(*[b.ptr for b in bbufs])
    st = (SolveStats * n)()
    rc = lib().qcuSolveBatch(g.handle, ctypes.byref(op), ctypes.byref(sp), n, xs, bp, st)
    if raise_on_error:
        check(rc)
    return [x.to_numpy((2, vh, 4, 3)) for x in xbufs], list(st)


def bench_hop(g: Gauge, precision=PREC_DOUBLE, recon=18, nrhs=1, iters=100) -> float:
    ms = ctypes.c_double()
    check(lib().qcuBenchHop(g.handle, precision, recon, nrhs, iters, ctypes.byref(ms)))
    return ms.value


def source_point_device(L, coords, spin, colour) -> DeviceBuffer:
    """Unit point source as a device full spinor (no host copy)."""
    buf = DeviceBuffer(2 * half_volume(L) * 192)
    check(lib().qcuSourcePoint(buf.ptr, _L(L), _L(coords), spin, colour))
    return buf


FIELD_SPINOR, FIELD_GAUGE = 0, 1
TO_COALESCE, TO_NON_COALESCE = 0, 1


def lex_to_eo(lex: np.ndarray, L, field=FIELD_SPINOR) -> np.ndarray:
    """Lexicographic -> even-odd order on the device (qcuLexToEvenOdd); complex128."""
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(lex, np.complex128))
    o = DeviceBuffer(a.nbytes)
    check(lib().qcuLexToEvenOdd(o.ptr, a.ptr, _L(L), field))
    return o.to_numpy(lex.shape)


def eo_to_lex(eo: np.ndarray, L, field=FIELD_SPINOR) -> np.ndarray:
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(eo, np.complex128))
    o = DeviceBuffer(a.nbytes)
    check(lib().qcuEvenOddToLex(o.ptr, a.ptr, _L(L), field))
    return o.to_numpy(eo.shape)


def shift_vector_storage(src: np.ndarray, direction, L) -> np.ndarray:
    """The reference's shiftVectorStorage on one half spinor: complex128 [Vh][4][3]
    <-> float64 [24][Vh]."""
    vh = half_volume(L)
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(src))
    o = DeviceBuffer(vh * 192)
    lib().qcuClearLastError()
    lib().shiftVectorStorage(o.ptr, a.ptr, direction, *L)
    check(lib().qcuGetLastError())
    if direction == TO_COALESCE:
        return o.to_numpy((24, vh), np.float64)
    return o.to_numpy((vh, 4, 3))


def shift_gauge_storage(src: np.ndarray, direction, L) -> np.ndarray:
    """The reference's shiftGaugeStorage: complex128 [4][2][Vh][3][3] <-> float64
    [4][2][12][Vh] (rows 0-1; the way back rebuilds row 2)."""
    vh = half_volume(L)
    a = DeviceBuffer.from_numpy(np.ascontiguousarray(src))
    if direction == TO_COALESCE:
        o = DeviceBuffer(8 * vh * 12 * 8)
        lib().qcuClearLastError()
        lib().shiftGaugeStorage(o.ptr, a.ptr, direction, *L)
        check(lib().qcuGetLastError())
        return o.to_numpy((4, 2, 12, vh), np.float64)
    o = DeviceBuffer(8 * vh * 18 * 8)
    lib().qcuClearLastError()
    lib().shiftGaugeStorage(o.ptr, a.ptr, direction, *L)
    check(lib().qcuGetLastError())
    return o.to_numpy((4, 2, vh, 3, 3))


def source_gaussian_device(L, seed) -> DeviceBuffer:
    buf = DeviceBuffer(2 * half_volume(L) * 192)
    check(lib().qcuSourceGaussian(buf.ptr, _L(L), seed))
    return buf


def solve_batch_device(g: Gauge, op: Operator, xs, bs, params: SolveParams | None = None, raise_on_error=True):
    """Batched solve on device buffers (DeviceBuffer lists); returns per-RHS stats."""
    sp = params or default_solve_params()
    n = len(bs)
    xp = (_vp * n)(*[x.ptr for x in xs])
    bp = (_vp * n)(*[b.ptr for b in bs])
    st = (SolveStats * n)()
    if n == 1:
        rc = lib().qcuSolve(g.handle, ctypes.byref(op), ctypes.byref(sp), xs[0].ptr, bs[0].ptr, ctypes.byref(st[0]))
    else:
        rc = lib().qcuSolveBatch(g.handle, ctypes.byref(op), ctypes.byref(sp), n, xp, bp, st)
    if raise_on_error:
        check(rc)
    return list(st)


def tune_hop(g: Gauge, precision=PREC_DOUBLE, recon=18, nrhs=1):
    """Autotune the hopping kernel's launch shape; returns (shape dict, ms)."""
    v = (ctypes.c_int * len(SHAPE_KEYS))()
    ms = ctypes.c_double()
    check(lib().qcuTuneHop(g.handle, precision, recon, nrhs, v, ctypes.byref(ms)))
    return dict(zip(SHAPE_KEYS, list(v))), ms.value


SHAPE_KEYS = ("block", "rb", "rpt", "xcd", "tile_z", "tile_t", "group_major", "tile_y", "rpw", "spt", "cached",
              "lds")


def get_hop_shape(g: Gauge, precision, recon, nrhs) -> dict:
    """The launch shape the plain hop uses for (precision, recon, nrhs) on g's volume."""
    v = (ctypes.c_int * len(SHAPE_KEYS))()
    check(lib().qcuGetHopShape(g.handle, precision, recon, nrhs, v))
    return dict(zip(SHAPE_KEYS, list(v)))


def set_hop_shape(g: Gauge, precision, recon, nrhs, shape: dict):
    """Pin the hopping kernel's launch shape for (precision, recon, nrhs) on g's volume."""
    v = (ctypes.c_int * len(SHAPE_KEYS))(*[int(shape.get(k, 1 if k in ("rpw", "spt") else 0)) for k in SHAPE_KEYS])
    check(lib().qcuSetHopShape(g.handle, precision, recon, nrhs, v))


def bench_link_product(mfma: int, site_legs=8 * 1048576, iters=10):
    """(ms, checksum) of the 12-RHS FP64 link product on VALU (0) or MFMA (1)."""
    ms, cs = ctypes.c_double(), ctypes.c_double()
    check(lib().qcuBenchLinkProduct(int(mfma), int(site_legs), iters, ctypes.byref(ms), ctypes.byref(cs)))
    return ms.value, cs.value


def bench_copy(nbytes, iters=20) -> float:
    ms = ctypes.c_double()
    check(lib().qcuBenchCopy(nbytes, iters, ctypes.byref(ms)))
    return ms.value


def param(L) -> QcuParam:
    p = QcuParam()
    for i in range(4):
        p.lattice_size[i] = int(L[i])
    return p


def grid(g=(1, 1, 1, 1)) -> QcuGrid:
    q = QcuGrid()
    for i in range(4):
        q.grid_size[i] = int(g[i])
    return q


class Server:
    """Continuous-batching solve server (qcuServer*)."""

    def __init__(self, precision=PREC_SINGLE, max_batch=12, chunk=8, graph=1, hbm_budget=0):
        cfg = ServerConfig()
        lib().qcuDefaultServerConfig(ctypes.byref(cfg))
        cfg.precision, cfg.max_batch, cfg.chunk, cfg.graph, cfg.hbm_budget = precision, max_batch, chunk, graph, hbm_budget
        h = _vp()
        check(lib().qcuServerCreate(ctypes.byref(h), ctypes.byref(cfg)))
        self.handle = h.value
        self._keep = {}

    def submit(self, gauge: Gauge, kappa, csw, tol, nrhs, seed=0, xs=None, bs=None, t_boundary=-1, recon=12,
               max_iter=0) -> int:
        r = Request()
        r.gauge, r.kappa, r.csw, r.tol = gauge.handle, kappa, csw, tol
        r.t_boundary, r.recon, r.nrhs, r.seed, r.max_iter = t_boundary, recon, nrhs, seed, max_iter
        keep = []
        if xs is not None:
            arr = (_vp * nrhs)(*[x.ptr for x in xs])
            r.x = ctypes.cast(arr, ctypes.POINTER(_vp))
            keep.append(arr)
        if bs is not None:
            arr = (_vp * nrhs)(*[b.ptr for b in bs])
            r.b = ctypes.cast(arr, ctypes.POINTER(_vp))
            keep.append(arr)
        rid = ctypes.c_uint64()
        check(lib().qcuServerSubmit(self.handle, ctypes.byref(r), ctypes.byref(rid)))
        self._keep[rid.value] = keep
        return rid.value

    def poll(self, rid) -> RequestStatus:
        st = RequestStatus()
        check(lib().qcuServerPoll(self.handle, rid, ctypes.byref(st)))
        return st

    def wait(self, rid, raise_on_error=True) -> RequestStatus:
        st = RequestStatus()
        rc = lib().qcuServerWait(self.handle, rid, ctypes.byref(st))
        if raise_on_error:
            check(rc)
        return st

    def drain(self):
        check(lib().qcuServerDrain(self.handle))

    def usage(self):
        b, p, n = _sz(), _sz(), ctypes.c_int()
        check(lib().qcuServerUsage(self.handle, ctypes.byref(b), ctypes.byref(p), ctypes.byref(n)))
        return {"bytes": b.value, "peak_bytes": p.value, "batches": n.value}

    def destroy(self):
        if self.handle:
            check(lib().qcuServerDestroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
