"""Pins the oracle's CG solvers and seeded generators."""
import numpy as np
import pytest

import textbook as tb


@pytest.fixture(scope="module")
def setup44(orc):
    L = (4, 4, 4, 4)
    g = orc.gauge_random(L, seed=31, eps=0.25)
    b = orc.source_gauss(L, seed=7)
    return L, g, b


@pytest.mark.parametrize("clover", [False, True])
def test_cg_eo_matches_dense_solve(orc, setup44, clover):
    L, g, b = setup44
    kappa, csw, tbc = 0.122, 1.0, -1.0
    A = Ainv = None
    if clover:
        A, Ainv = orc.clover_fields(g, L, kappa, csw)
    x, st = orc.cg_eo(b, g, L, kappa, A=A, Ainv=Ainv, tbc=tbc, tol=1e-12)
    assert st.converged and st.true_residual < 1e-11
    M = tb.dense_operator(tb.eo_to_lex_gauge(g, L), L, kappa, s=1, tbc=tbc, A_blocks=A)
    xr = np.linalg.solve(M, b.reshape(-1)).reshape(b.shape)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-10


def test_cg_eo_reference_normalisation(orc, setup44):
    """kappa = c_sw = 1, periodic: the system cg_inverter solves, (A - D) x = b."""
    L, g, b = setup44
    A, Ainv = orc.clover_fields(g, L, 1.0, 1.0)
    x, st = orc.cg_eo(b, g, L, 1.0, A=A, Ainv=Ainv, tbc=1.0, tol=1e-10, max_iter=20000)
    assert st.converged and st.true_residual <= 1e-10
    M = tb.dense_operator(tb.eo_to_lex_gauge(g, L), L, 1.0, s=1, tbc=1.0, A_blocks=A)
    r = b.reshape(-1) - M @ x.reshape(-1)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.0001e-10


def test_cg_ref_wilson_matches_dense(orc, setup44):
    """The reference's full-lattice CG (exact CG when A = 1) solves (1 - D) x = b."""
    L, g, b = setup44
    x, it, rc = orc.cg_ref(b, g, L, Ainv=None, stop_rr=1e-23)
    assert rc == 0 and it > 0
    M = tb.dense_operator(tb.eo_to_lex_gauge(g, L), L, 1.0, s=1, tbc=1.0)
    xr = np.linalg.solve(M, b.reshape(-1)).reshape(b.shape)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
    # the even-odd CGLS solve of the same system gives the same solution
    xe, st = orc.cg_eo(b, g, L, 1.0, tbc=1.0, tol=1e-12, max_iter=20000)
    assert np.linalg.norm(xe - xr) / np.linalg.norm(xr) < 1e-9


def test_cg_config1_cpu_solve(orc):
    """8^4 Wilson, kappa = 0.12, eps = 0.1, antiperiodic t, one Gaussian source, 1e-10."""
    L = (8, 8, 8, 8)
    g = orc.gauge_random(L, seed=1, eps=0.1)
    b = orc.source_gauss(L, seed=2)
    x, st = orc.cg_eo(b, g, L, 0.12, tbc=-1.0, tol=1e-10)
    assert st.converged and st.true_residual <= 1e-10 and 5 < st.iterations < 500


def test_point_source(orc):
    L = (4, 4, 2, 2)
    v = orc.source_point(L, (1, 2, 1, 0), 2, 1)
    p, i = orc.index(L, (1, 2, 1, 0))
    assert v[p, i, 2, 1] == 1.0 and np.count_nonzero(v) == 1


def test_gauge_generator_su3(orc):
    L = (4, 4, 2, 2)
    g = orc.gauge_random(L, seed=5, eps=0.25).reshape(-1, 3, 3)
    assert np.abs(g @ tb.dag(g) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(g) - 1).max() < 1e-14
    # near-unit: |U - 1| grows with eps
    g2 = orc.gauge_random(L, seed=5, eps=0.05).reshape(-1, 3, 3)
    assert np.abs(g2 - np.eye(3)).mean() < np.abs(g - np.eye(3)).mean()
    g0 = orc.gauge_random(L, seed=5, eps=0.0).reshape(-1, 3, 3)
    assert np.abs(g0 - np.eye(3)).max() < 1e-15


def test_gauge_generator_matches_matrix_exponential(orc):
    """U = exp(i eps H), H = sum_a h_a lambda_a / 2 (already in SU(3), so the
    projection changes nothing), with h_a the generator's Gaussians."""
    import scipy.linalg

    L = (2, 2, 2, 2)
    eps, seed = 0.3, 77
    g = orc.gauge_random(L, seed=seed, eps=eps).reshape(-1, 3, 3)
    lam = np.zeros((8, 3, 3), complex)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1], lam[1][1, 0] = -1j, 1j
    lam[2][0, 0], lam[2][1, 1] = 1, -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2], lam[4][2, 0] = -1j, 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2], lam[6][2, 1] = -1j, 1j
    lam[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    for link in range(g.shape[0]):
        h = [orc.rng_gauss(seed, link * 8 + a) for a in range(8)]
        H = sum(h[a] * lam[a] / 2 for a in range(8))
        assert np.abs(scipy.linalg.expm(1j * eps * H) - g[link]).max() < 1e-14


def test_rng_determinism_and_moments(orc):
    a = [orc.rng_u64(3, i) for i in range(5)]
    assert a == [orc.rng_u64(3, i) for i in range(5)]
    assert a != [orc.rng_u64(4, i) for i in range(5)]
    L = (8, 8, 8, 8)
    v = orc.source_gauss(L, seed=11).view(np.float64).ravel()
    assert abs(v.mean()) < 0.01 and abs(v.var() - 1) < 0.01
