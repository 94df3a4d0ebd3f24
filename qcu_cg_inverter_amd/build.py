"""Build libqcu.so (hand-written HIP kernels for gfx950 + the C ABI) in-tree.

Every source under csrc/ is compiled by hipcc for gfx950 into _build/, in parallel,
and linked into qcu_cg_inverter_amd/libqcu.so.  Run ``python -m
qcu_cg_inverter_amd.build`` or call ``build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libqcu.so")

SOURCES = [
    *[f"k_hop_{p}_{r}.hip" for p in ("f64", "f32", "f16") for r in (1, 2, 3, 4)],
    "k_hop.hip",
    "k_hop_iface.hip",
    "k_layout.hip",
    "k_blas.hip",
    "k_clover.hip",
    "k_gen.hip",
    "k_linkprod.hip",
    "api.cpp",
    "solver.cpp",
    "cg_batch.cpp",
    "tune.cpp",
    "serve.cpp",
    "noalloc.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed-FP32 instructions are disabled for all device code.  No implicit FMA
# contraction: where the kernels fuse, they say so (cx ops in qcu_fields.h), so every
# instantiation of a kernel (right-hand sides per thread, epilogue, shape) rounds the
# same way and batched results equal single ones bit for bit.
CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l
                    and "ignoring feature" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
