"""Pins the oracle's clover term, 6x6 inverses and operator compositions."""
import numpy as np
import pytest

import textbook as tb


def test_clover_unit_gauge_is_zero(orc):
    L = (4, 4, 4, 4)
    T = orc.clover_T(orc.gauge_unit(L), L)
    assert np.abs(T).max() == 0.0


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (4, 4, 4, 4), (4, 2, 6, 2)])
def test_clover_matches_textbook_sigma_F(orc, L):
    g = orc.gauge_random(L, seed=13, eps=0.8)
    T = orc.clover_T(g, L)
    ref, off = tb.T_to_blocks(tb.clover_sigma_F(tb.eo_to_lex_gauge(g, L)), L)
    assert off < 1e-14  # textbook T is block diagonal in chirality
    assert np.abs(T - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())


def test_clover_coefficient_table_is_gamma_product():
    """The reference tabulates g_mu g_nu per plane (src/qcu_clover_dslash.cu:210-267);
    the oracle restates that table; here it is rebuilt from the hop's gammas."""
    planes = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    table = {
        (0, 1): {(0, 0): -1j, (1, 1): 1j, (2, 2): -1j, (3, 3): 1j},
        (0, 2): {(0, 1): -1, (1, 0): 1, (2, 3): -1, (3, 2): 1},
        (0, 3): {(0, 1): 1j, (1, 0): 1j, (2, 3): -1j, (3, 2): -1j},
        (1, 2): {(0, 1): -1j, (1, 0): -1j, (2, 3): -1j, (3, 2): -1j},
        (1, 3): {(0, 1): -1, (1, 0): 1, (2, 3): 1, (3, 2): -1},
        (2, 3): {(0, 0): 1j, (1, 1): -1j, (2, 2): -1j, (3, 3): 1j},
    }
    for mu, nu in planes:
        gg = tb.GAMMA[mu] @ tb.GAMMA[nu]
        m = np.zeros((4, 4), complex)
        for (r, c), v in table[(mu, nu)].items():
            m[r, c] = v
        assert np.allclose(gg, m)


def test_clover_hermitian_and_covariant(orc):
    L = (4, 4, 2, 4)
    rng = np.random.default_rng(5)
    g = orc.gauge_random(L, seed=3, eps=0.6)
    T = orc.clover_T(g, L)
    assert np.abs(T - np.conj(np.swapaxes(T, -1, -2))).max() < 1e-15
    U = tb.eo_to_lex_gauge(g, L)
    G = tb.random_su3(rng, int(np.prod(L))).reshape(L[3], L[2], L[1], L[0], 3, 3)
    Up = np.stack([G @ U[mu] @ tb.dag(tb.shift(G, mu, 1)) for mu in range(4)])
    Tp = orc.clover_T(tb.lex_to_eo_gauge(Up, L), L)
    vh = orc.half_volume(L)
    for p in range(2):
        x, y, z, t = tb.eo_coords(L, p, vh)
        Gs = G[t, z, y, x]
        for b in range(2):
            for rs in range(2):
                for cs in range(2):
                    blk = T[p, :, b, rs * 3:rs * 3 + 3, cs * 3:cs * 3 + 3]
                    blkp = Tp[p, :, b, rs * 3:rs * 3 + 3, cs * 3:cs * 3 + 3]
                    assert np.abs(Gs @ blk @ tb.dag(Gs) - blkp).max() < 1e-13


def test_clover_single_link_support(orc):
    """Changing U_mu(y) changes T exactly at the corners of the plaquettes holding it:
    y, y+mu and y+nu, y+mu+nu, y-nu, y+mu-nu for every nu != mu."""
    L = (6, 6, 6, 6)
    g = orc.gauge_random(L, seed=6, eps=0.3)
    T0 = orc.clover_T(g, L)
    for mu, y in [(0, (2, 3, 1, 4)), (3, (5, 0, 2, 0)), (1, (0, 5, 5, 3))]:
        p, i = orc.index(L, y)
        g2 = g.copy()
        g2[mu, p, i] = orc.gauge_random((2, 2, 2, 2), seed=99, eps=1.0)[0, 0, 0]
        T1 = orc.clover_T(g2, L)
        changed = set()
        for q in range(2):
            d = np.abs(T1[q] - T0[q]).reshape(T0.shape[1], -1).max(axis=1)
            changed |= {(q, int(j)) for j in np.nonzero(d > 1e-14)[0]}
        expect = set()
        for nu in range(4):
            if nu == mu:
                continue
            for a, bb in [(0, 0), (1, 0), (0, 1), (1, 1), (0, -1), (1, -1)]:
                c = list(y)
                c[mu] += a
                c[nu] += bb
                expect.add(orc.index(L, c))
        assert changed == expect


def test_invert6_methods(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        h = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
        m = np.eye(6) + 0.05 * (h + h.conj().T)
        assert np.linalg.eigvalsh(m).min() > 0
        gj = orc.invert6_gj(m)
        ch, rc = orc.invert6_chol(m)
        assert rc == 0
        ref = np.linalg.inv(m)
        assert np.abs(gj - ref).max() < 1e-14
        assert np.abs(ch - ref).max() < 1e-14
        assert np.abs(m @ ch - np.eye(6)).max() < 1e-14
    bad = np.diag([1, 1, -1, 1, 1, 1]).astype(complex)
    _, rc = orc.invert6_chol(bad)
    assert rc == -1


def test_clover_invert_fields(orc):
    L = (4, 4, 4, 4)
    g = orc.gauge_random(L, seed=1, eps=0.25)
    A, Ainv = orc.clover_fields(g, L, kappa=0.122, csw=1.0)
    prod = np.einsum("...ij,...jk->...ik", A, Ainv)
    assert np.abs(prod - np.eye(6)).max() < 1e-14
    # reference normalisation (kappa * c_sw = 1): some blocks of A = 1 + T are
    # indefinite, so Cholesky must fall back to Gauss-Jordan on exactly those
    A1, _ = orc.clover_fields(g, L, kappa=1.0, csw=1.0)
    nneg = int((np.linalg.eigvalsh(A1.reshape(-1, 6, 6)).min(axis=1) <= 0).sum())
    gj, _ = orc.clover_invert(A1, chol=False)
    ch, nbad = orc.clover_invert(A1, chol=True)
    assert nbad == nneg > 0
    # near-singular blocks (condition number ~1e3) bound the agreement
    assert np.abs(gj - ch).max() < 1e-13 * np.linalg.cond(A1.reshape(-1, 6, 6)).max() * np.abs(gj).max()
    assert np.abs(np.einsum("...ij,...jk->...ik", A1, gj) - np.eye(6)).max() < 1e-12


def test_clover_apply_matches_dense(orc):
    L = (2, 2, 4, 2)
    g = orc.gauge_random(L, seed=1, eps=0.5)
    A, _ = orc.clover_fields(g, L, 0.13, 1.0)
    v = orc.source_gauss(L, seed=4)
    out = orc.clover_apply(v[1], A[1])
    vf = v[1].reshape(-1, 12)
    ref = np.stack([np.concatenate([A[1, i, 0] @ vf[i, :6], A[1, i, 1] @ vf[i, 6:]]) for i in range(vf.shape[0])])
    assert np.abs(out.reshape(-1, 12) - ref).max() < 1e-14


@pytest.mark.parametrize("clover", [False, True])
@pytest.mark.parametrize("s", [1, -1])
def test_full_op_matches_dense(orc, clover, s):
    L = (2, 2, 2, 4)
    kappa, csw, tbc = 0.124, 1.0, -1.0
    g = orc.gauge_random(L, seed=17, eps=0.5)
    A = orc.clover_fields(g, L, kappa, csw)[0] if clover else None
    v = orc.source_gauss(L, seed=2)
    out = orc.full_op(v, g, L, kappa, A=A, s=s, tbc=tbc)
    M = tb.dense_operator(tb.eo_to_lex_gauge(g, L), L, kappa, s=s, tbc=tbc, A_blocks=A)
    ref = (M @ v.reshape(-1)).reshape(v.shape)
    assert np.abs(out - ref).max() < 1e-13


def test_schur_identity(orc):
    """With x_o = A_oo^-1 kappa D_oe x_e the odd rows of M x vanish and the even rows
    equal Mhat x_e."""
    L = (4, 4, 2, 4)
    kappa = 0.122
    g = orc.gauge_random(L, seed=23, eps=0.25)
    A, Ainv = orc.clover_fields(g, L, kappa, 1.0)
    xe = orc.source_gauss(L, seed=5)[0]
    xo = orc.clover_apply(kappa * orc.hop(xe, g, L, 1, tbc=-1.0), Ainv[1])
    full = orc.full_op(np.stack([xe, xo]), g, L, kappa, A=A, tbc=-1.0)
    mh = orc.schur_op(xe, g, L, kappa, Aee=A[0], Ainv_oo=Ainv[1], tbc=-1.0)
    assert np.abs(full[1]).max() < 1e-13
    assert np.abs(full[0] - mh).max() < 1e-13
    # adjoint: <a, Mhat b> = <Mhat^dag a, b>
    a = orc.source_gauss(L, seed=6)[0]
    lhs = np.vdot(a, orc.schur_op(xe, g, L, kappa, A[0], Ainv[1], s=1, tbc=-1.0))
    rhs = np.vdot(orc.schur_op(a, g, L, kappa, A[0], Ainv[1], s=-1, tbc=-1.0), xe)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
