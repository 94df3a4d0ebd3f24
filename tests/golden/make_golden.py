"""Writes tests/golden/golden_4x4x4x4.npz: seeded inputs and the oracle's outputs for
every function of the path on a 4^4 lattice.

The reference ships no vectors of its own (its one test, src/test/
test_qcu_complex_computation.cu, has no fixtures or assertions), so these are written
by the oracle, after the oracle is pinned by the textbook, closed-form and dense-solve
checks of tests/test_oracle_*.py.  They hold the oracle and libqcu.so to fixed numbers
across changes: tests/test_golden.py recomputes them on the CPU, the GPU tests compare
libqcu.so with them.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as orc  # noqa: E402

L = (4, 4, 4, 4)
META = {
    "L": list(L),
    "gauge_seed": 20261, "gauge_eps": 0.3,
    "source_seed": 77,
    # the reference's own normalisation (dslashQcu, fullDslashQcu, cg_inverter):
    # kappa = 1, kappa * c_sw = 1, periodic, reconstruct-12
    "ref_kappa": 1.0, "ref_csw": 1.0,
    # a physical Wilson and clover operator, antiperiodic in t
    "wilson_kappa": 0.124,
    "clover_kappa": 0.122, "clover_csw": 1.0,
    "cg_tol": 1e-13,
}


def main():
    m = META
    orc.set_threads(1)  # fixed summation order for the CG's reductions
    g = orc.gauge_random(L, m["gauge_seed"], m["gauge_eps"])
    b = orc.source_gauss(L, m["source_seed"])
    A_ref, Ainv_ref = orc.clover_fields(g, L, m["ref_kappa"], m["ref_csw"], recon=12)
    A_cl, Ainv_cl = orc.clover_fields(g, L, m["clover_kappa"], m["clover_csw"], recon=12)
    out = {
        "gauge": g,
        "b": b,
        # dslashQcu: D_{p,1-p}, s = +1, kappa = 1, periodic (src/dslash.cu:62-77)
        "dslash_p0": orc.hop(b[1], g, L, 0, s=1, tbc=1.0, recon=12),
        "dslash_p1": orc.hop(b[0], g, L, 1, s=1, tbc=1.0, recon=12),
        # fullDslashQcu: (A - D_s) with A = 1 + T (src/qcu_clover_dslash.cu:948-987)
        "full_ref_s1": orc.full_op(b, g, L, 1.0, A=A_ref, s=1, tbc=1.0, recon=12),
        "full_ref_sm1": orc.full_op(b, g, L, 1.0, A=A_ref, s=-1, tbc=1.0, recon=12),
        "clover_A_ref": A_ref,
        "clover_Ainv_ref": Ainv_ref,
        "wilson_full": orc.full_op(b, g, L, m["wilson_kappa"], A=None, s=1, tbc=-1.0, recon=18),
    }
    x, st = orc.cg_eo(b, g, L, 1.0, A=A_ref, Ainv=Ainv_ref, tbc=1.0, recon=12, tol=m["cg_tol"], max_iter=20000)
    assert st.converged, "reference-normalised solve did not converge"
    out["x_ref"] = x
    m["x_ref_iterations"] = st.iterations
    x, st = orc.cg_eo(b, g, L, m["wilson_kappa"], tbc=-1.0, recon=18, tol=m["cg_tol"])
    assert st.converged
    out["x_wilson"] = x
    m["x_wilson_iterations"] = st.iterations
    x, st = orc.cg_eo(b, g, L, m["clover_kappa"], A=A_cl, Ainv=Ainv_cl, tbc=-1.0, recon=12, tol=m["cg_tol"])
    assert st.converged
    out["x_clover"] = x
    m["x_clover_iterations"] = st.iterations
    out["meta"] = np.array(json.dumps(m, sort_keys=True))
    path = os.path.join(ROOT, "tests", "golden", "golden_4x4x4x4.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes;", {k: m[k] for k in m if k.endswith("iterations")})


if __name__ == "__main__":
    main()
