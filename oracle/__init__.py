"""CPU FP64 oracle for the qcu even-odd Wilson / Wilson-clover path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import it.  The
product (``qcu_cg_inverter_amd``, ``libqcu.so``) never imports, links or calls it.

It wraps ``liboracle.so`` (built from ``qcu_oracle.c`` by ``make -C oracle``), a C
restatement of the reference arithmetic; each C function cites the reference
file:line it follows.  Arrays are numpy complex128 in the reference's interface
layouts:

* gauge       ``[4][2][Vh][3][3]``   (mu, parity, half-site, row, col)
* half spinor ``[Vh][4][3]``
* full spinor ``[2][Vh][4][3]``      (parity 0 first)
* clover      ``[2][Vh][2][6][6]``
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _declare(ctypes.CDLL(_LIB_PATH))
    return _lib


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcSolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("true_residual", ctypes.c_double),
        ("normal_residual", ctypes.c_double),
        ("converged", ctypes.c_int),
    ]


def _declare(L):
    L.orc_coords.argtypes = [_ip, ctypes.c_int, ctypes.c_long, _ip]
    L.orc_index.argtypes = [_ip, _ip, _ip]
    L.orc_index.restype = ctypes.c_long
    L.orc_neighbor.argtypes = [_ip, ctypes.c_int, ctypes.c_long, ctypes.c_int, ctypes.c_int, _ip]
    L.orc_neighbor.restype = ctypes.c_long
    L.orc_reconstruct12.argtypes = [_dp]
    L.orc_shift_vector.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_long]
    L.orc_shift_gauge.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_long]
    L.orc_lex_eo.argtypes = [_dp, _dp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_hop.argtypes = [_dp, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.orc_clover_T.argtypes = [_dp, _dp, _ip, ctypes.c_int]
    L.orc_clover_A.argtypes = [_dp, _dp, ctypes.c_long, ctypes.c_double]
    L.orc_invert6_gj.argtypes = [_dp, _dp]
    L.orc_invert6_chol.argtypes = [_dp, _dp]
    L.orc_invert6_chol.restype = ctypes.c_int
    L.orc_clover_invert.argtypes = [_dp, _dp, ctypes.c_long, ctypes.c_int]
    L.orc_clover_invert.restype = ctypes.c_int
    L.orc_clover_apply.argtypes = [_dp, _dp, _dp, ctypes.c_long]
    L.orc_full_op.argtypes = [_dp, _dp, _dp, _dp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                              ctypes.c_int]
    L.orc_schur_op.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, ctypes.c_double, ctypes.c_int,
                               ctypes.c_double, ctypes.c_int]
    L.orc_cg_eo.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                            ctypes.c_double, ctypes.c_int, ctypes.POINTER(OrcSolveStats)]
    L.orc_cg_eo.restype = ctypes.c_int
    L.orc_cg_ref.argtypes = [_dp, _dp, _dp, _dp, _ip, ctypes.c_double, ctypes.c_int, _ip]
    L.orc_cg_ref.restype = ctypes.c_int
    L.orc_rng_u64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.orc_rng_u64.restype = ctypes.c_uint64
    L.orc_rng_gauss.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.orc_rng_gauss.restype = ctypes.c_double
    L.orc_gauge_random.argtypes = [_dp, _ip, ctypes.c_uint64, ctypes.c_double]
    L.orc_source_gauss.argtypes = [_dp, ctypes.c_long, ctypes.c_uint64]
    L.orc_source_point.argtypes = [_dp, _ip, _ip, ctypes.c_int, ctypes.c_int]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    return L


def _L(L):
    return (ctypes.c_int * 4)(*[int(v) for v in L])


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    assert a.dtype in (np.complex128, np.float64)
    return a.ctypes.data_as(_dp)


def volume(L):
    return int(L[0]) * int(L[1]) * int(L[2]) * int(L[3])


def half_volume(L):
    return volume(L) // 2


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


# ---- geometry --------------------------------------------------------------
def coords(L, parity, idx):
    c = (ctypes.c_int * 4)()
    lib().orc_coords(_L(L), parity, idx, c)
    return tuple(c)


def index(L, c):
    par = ctypes.c_int()
    i = lib().orc_index(_L(L), (ctypes.c_int * 4)(*c), ctypes.byref(par))
    return par.value, i


def neighbor(L, parity, idx, mu, fwd):
    w = ctypes.c_int()
    n = lib().orc_neighbor(_L(L), parity, idx, mu, 1 if fwd else 0, ctypes.byref(w))
    return n, w.value


# ---- fields ------------------------------------------------------------------
def zeros_half(L):
    return np.zeros((half_volume(L), 4, 3), np.complex128)


def zeros_full(L):
    return np.zeros((2, half_volume(L), 4, 3), np.complex128)


def gauge_random(L, seed, eps):
    g = np.empty((4, 2, half_volume(L), 3, 3), np.complex128)
    lib().orc_gauge_random(_p(g), _L(L), seed, eps)
    return g


def gauge_unit(L):
    g = np.zeros((4, 2, half_volume(L), 3, 3), np.complex128)
    g[..., 0, 0] = g[..., 1, 1] = g[..., 2, 2] = 1.0
    return g


def source_gauss(L, seed):
    v = zeros_full(L)
    lib().orc_source_gauss(_p(v), volume(L), seed)
    return v


def source_point(L, c, spin, colour):
    v = zeros_full(L)
    lib().orc_source_point(_p(v), _L(L), (ctypes.c_int * 4)(*c), spin, colour)
    return v


def rng_u64(seed, ctr):
    return lib().orc_rng_u64(seed, ctr)


def rng_gauss(seed, ctr):
    return lib().orc_rng_gauss(seed, ctr)


def reconstruct12(u):
    u = np.ascontiguousarray(u, np.complex128).copy()
    lib().orc_reconstruct12(_p(u))
    return u


# ---- layouts -----------------------------------------------------------------
def shift_vector(src, direction, L):
    """Half spinor: complex128 [Vh][4][3] -> float64 [24][Vh] (direction 0) or back."""
    vh = volume(L) // 2
    src = np.ascontiguousarray(src)
    if direction == 0:
        out = np.empty((24, vh), np.float64)
    else:
        out = np.empty((vh, 4, 3), np.complex128)
    lib().orc_shift_vector(_p(out), _p(src), direction, vh)
    return out


def shift_gauge(src, direction, L):
    """Gauge: complex128 [4][2][Vh][3][3] -> float64 [4][2][12][Vh] (direction 0) or
    back with row 2 rebuilt."""
    vh = volume(L) // 2
    src = np.ascontiguousarray(src)
    if direction == 0:
        out = np.empty((4, 2, 12, vh), np.float64)
    else:
        out = np.empty((4, 2, vh, 3, 3), np.complex128)
    lib().orc_shift_gauge(_p(out), _p(src), direction, vh)
    return out


def lex_eo(src, L, nc, nf, to_eo):
    """nf fields of nc complex per site, lexicographic <-> even-odd (same shape out)."""
    src = np.ascontiguousarray(src, np.complex128)
    out = np.empty_like(src)
    lib().orc_lex_eo(_p(out), _p(src), _L(L), nc, nf, int(to_eo))
    return out


# ---- operators ---------------------------------------------------------------
def hop(inp, gauge, L, parity, s=1, tbc=1.0, recon=12):
    """out(parity) = D^s_{p,1-p} inp(1-parity); inp/out are half spinors."""
    out = zeros_half(L)
    lib().orc_hop(_p(out), _p(np.ascontiguousarray(inp)), _p(gauge), _L(L), parity, s, tbc, recon)
    return out


def clover_T(gauge, L, recon=12):
    T = np.empty((2, half_volume(L), 2, 6, 6), np.complex128)
    lib().orc_clover_T(_p(T), _p(gauge), _L(L), recon)
    return T


def clover_A(T, ck):
    A = np.empty_like(T)
    lib().orc_clover_A(_p(A), _p(np.ascontiguousarray(T)), T.size // 72, ck)
    return A


def clover_invert(A, chol=True):
    inv = np.empty_like(A)
    bad = lib().orc_clover_invert(_p(inv), _p(np.ascontiguousarray(A)), A.size // 72, 1 if chol else 0)
    return inv, bad


def invert6_gj(m):
    m = np.ascontiguousarray(m, np.complex128)
    r = np.empty_like(m)
    lib().orc_invert6_gj(_p(m), _p(r))
    return r


def invert6_chol(m):
    m = np.ascontiguousarray(m, np.complex128)
    r = np.empty_like(m)
    rc = lib().orc_invert6_chol(_p(m), _p(r))
    return r, rc


def clover_apply(inp, M):
    """inp half spinor [Vh][4][3], M half-lattice clover [Vh][2][6][6]."""
    out = np.empty_like(inp)
    lib().orc_clover_apply(_p(out), _p(np.ascontiguousarray(inp)), _p(np.ascontiguousarray(M)), inp.shape[0])
    return out


def full_op(inp, gauge, L, kappa, A=None, s=1, tbc=1.0, recon=12):
    out = zeros_full(L)
    lib().orc_full_op(_p(out), _p(np.ascontiguousarray(inp)), _p(gauge), _p(A), _L(L), kappa, s, tbc, recon)
    return out


def schur_op(inp, gauge, L, kappa, Aee=None, Ainv_oo=None, s=1, tbc=1.0, recon=12):
    out = zeros_half(L)
    lib().orc_schur_op(_p(out), _p(np.ascontiguousarray(inp)), _p(gauge), _p(Aee), _p(Ainv_oo), _L(L), kappa,
                       s, tbc, recon)
    return out


def cg_eo(b, gauge, L, kappa, A=None, Ainv=None, tbc=1.0, recon=12, tol=1e-10, max_iter=10000):
    x = zeros_full(L)
    st = OrcSolveStats()
    lib().orc_cg_eo(_p(x), _p(np.ascontiguousarray(b)), _p(gauge), _p(A), _p(Ainv), _L(L), kappa, tbc, recon, tol,
                    max_iter, ctypes.byref(st))
    return x, st


def cg_ref(b, gauge, L, Ainv=None, stop_rr=1e-23, max_iter=None):
    x = zeros_full(L)
    it = ctypes.c_int()
    if max_iter is None:
        max_iter = volume(L)
    rc = lib().orc_cg_ref(_p(x), _p(np.ascontiguousarray(b)), _p(gauge), _p(Ainv), _L(L), stop_rr, max_iter,
                          ctypes.byref(it))
    return x, it.value, rc


def clover_fields(gauge, L, kappa, csw, recon=12):
    """(A, A^-1) for both parities with A = 1 + kappa*c_sw*T."""
    T = clover_T(gauge, L, recon)
    A = clover_A(T, kappa * csw)
    Ainv, _ = clover_invert(A, chol=True)
    return A, Ainv
