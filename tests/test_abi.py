"""CPU-side checks of the C ABI: libqcu.so loads, exports every symbol that
include/*.h declares, and carries no link to the oracle.  No GPU call is made."""
import glob
import os
import re
import subprocess

import pytest

import qcu_cg_inverter_amd as q

HEADERS = sorted(glob.glob(os.path.join(q.ROOT, "include", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", src, flags=re.M):
            name = m.group(1)
            if name not in ("if", "for", "while", "sizeof", "return"):
                names.add(name)
    return names


def test_library_built():
    assert os.path.exists(q.LIB_PATH), "build libqcu.so first"


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert {"initGridSize", "dslashQcu", "fullDslashQcu", "cg_inverter", "loadQcuGauge"} <= names
    L = q.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes binding declares exactly the header's functions
    assert set(q.PROTOTYPES) == names


def test_dynamic_symbol_table():
    out = subprocess.run(["nm", "-D", "--defined-only", q.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in declared_functions():
        assert n in exported, n


def test_no_oracle_or_torch_dependency():
    out = subprocess.run(["ldd", q.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "torch" not in out
    nm = subprocess.run(["nm", "-D", q.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in nm


def test_version_string_without_gpu():
    assert b"gfx950" in q.lib().qcuVersion()


def test_struct_layouts():
    import ctypes

    assert ctypes.sizeof(q.QcuParam) == 16 and ctypes.sizeof(q.QcuGrid) == 16
    assert ctypes.sizeof(q.Operator) == 24
    assert ctypes.sizeof(q.SolveStats) == 32


@pytest.mark.parametrize("d", [1, 2, 3, 5, 6, 7, 12, 16, 24, 31, 48, 64, 97, 1000])
def test_fastdiv_host_model(d):
    """The kernels divide by Lx/2, Ly, Lz with a multiply-high (round-up method);
    this restates its constants in Python and checks exactness."""
    l = 0
    while (1 << l) < d:
        l += 1
    m = ((((1 << l) - d) << 32) // d + 1) & 0xFFFFFFFF
    if d == 1:
        m, l = 0, 0
    for n in list(range(0, 5000)) + [2**20 - 1, 2**24 + 7, 2**30 + 12345, 2**31 - 1]:
        hi = (n * m) >> 32
        assert (hi + n) >> l == n // d
