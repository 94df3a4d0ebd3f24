"""The oracle against the committed golden vectors (tests/golden/golden_4x4x4x4.npz,
written by tests/golden/make_golden.py): inputs regenerate bit for bit from their
seeds, per-site outputs recompute bit for bit, and the stored solutions solve their
systems to the stored tolerance."""
import json
import os

import numpy as np
import pytest

import oracle as orc

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_4x4x4x4.npz")


@pytest.fixture(scope="module")
def gold():
    z = np.load(PATH, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(str(d["meta"]))
    return d


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_inputs_regenerate(gold):
    m = gold["meta"]
    L = tuple(m["L"])
    np.testing.assert_array_equal(orc.gauge_random(L, m["gauge_seed"], m["gauge_eps"]), gold["gauge"])
    np.testing.assert_array_equal(orc.source_gauss(L, m["source_seed"]), gold["b"])


def test_per_site_outputs_recompute_exactly(gold):
    m = gold["meta"]
    L = tuple(m["L"])
    g, b = gold["gauge"], gold["b"]
    np.testing.assert_array_equal(orc.hop(b[1], g, L, 0, s=1, tbc=1.0, recon=12), gold["dslash_p0"])
    np.testing.assert_array_equal(orc.hop(b[0], g, L, 1, s=1, tbc=1.0, recon=12), gold["dslash_p1"])
    A, Ainv = orc.clover_fields(g, L, m["ref_kappa"], m["ref_csw"], recon=12)
    np.testing.assert_array_equal(A, gold["clover_A_ref"])
    np.testing.assert_array_equal(Ainv, gold["clover_Ainv_ref"])
    np.testing.assert_array_equal(orc.full_op(b, g, L, 1.0, A=A, s=1, tbc=1.0, recon=12), gold["full_ref_s1"])
    np.testing.assert_array_equal(orc.full_op(b, g, L, 1.0, A=A, s=-1, tbc=1.0, recon=12), gold["full_ref_sm1"])
    np.testing.assert_array_equal(orc.full_op(b, g, L, m["wilson_kappa"], s=1, tbc=-1.0, recon=18),
                                  gold["wilson_full"])


def test_stored_solutions_solve(gold):
    m = gold["meta"]
    L = tuple(m["L"])
    g, b = gold["gauge"], gold["b"]
    tol = 10 * m["cg_tol"]
    A = gold["clover_A_ref"]
    assert rel(orc.full_op(gold["x_ref"], g, L, 1.0, A=A, s=1, tbc=1.0, recon=12), b) < tol
    assert rel(orc.full_op(gold["x_wilson"], g, L, m["wilson_kappa"], tbc=-1.0, recon=18), b) < tol
    A_cl = orc.clover_fields(g, L, m["clover_kappa"], m["clover_csw"], recon=12)[0]
    assert rel(orc.full_op(gold["x_clover"], g, L, m["clover_kappa"], A=A_cl, tbc=-1.0, recon=12), b) < tol


def test_golden_known_answers(gold):
    """Properties the mathematics fixes, checked on the stored vectors themselves:
    A is Hermitian with unit-plus-traceless blocks, A*A^-1 = 1, and D with s = +1 and
    s = -1 are adjoint across parities."""
    A, Ainv = gold["clover_A_ref"], gold["clover_Ainv_ref"]
    np.testing.assert_allclose(A, np.conj(np.swapaxes(A, -1, -2)), atol=1e-15)
    eye = np.eye(6)
    assert np.abs(A @ Ainv - eye).max() < 1e-13
    assert np.abs(np.trace(A, axis1=-2, axis2=-1) - 6).max() < 1e-13  # tr(sigma.F) = 0
    m = gold["meta"]
    L = tuple(m["L"])
    g, b = gold["gauge"], gold["b"]
    lhs = np.vdot(b[0], gold["dslash_p0"])  # <phi_e, D^+ psi_o>
    rhs = np.vdot(orc.hop(b[0], g, L, 1, s=-1, tbc=1.0, recon=12), b[1])
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
