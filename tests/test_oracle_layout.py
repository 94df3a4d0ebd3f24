"""Oracle layout functions pinned by coordinate-coded known answers.

The reference ships no vectors for its storage shifts (src/qcu_shift_storage.cu) or
its even-odd order (include/qcu_point.cuh:20-93).  These tests build fields whose
every entry names its own site and component, derive the expected position of each
entry from numpy's lexicographic grid (independently of the oracle's index code), and
check the oracle moves every entry there; round trips and the SU(3) row-2 closed form
cover the rest."""
import numpy as np
import pytest

import oracle as orc

LATTICES = [(4, 4, 4, 4), (2, 2, 2, 2), (6, 4, 2, 8), (4, 3, 5, 2), (8, 2, 2, 6)]


def coded_lex(L, nc, nf=1):
    """Lexicographic field [nf][Lt][Lz][Ly][Lx][nc]: entry = site code + 1j*(comp + 100 field)."""
    Lx, Ly, Lz, Lt = L
    t, z, y, x = np.indices((Lt, Lz, Ly, Lx))
    code = (x + 10 * y + 100 * z + 1000 * t).astype(np.float64)
    f = np.arange(nf)[:, None, None, None, None, None]
    k = np.arange(nc)[None, None, None, None, None, :]
    return (code[None, ..., None] + 1j * (k + 100 * f)).astype(np.complex128)


def expected_eo(lex, L):
    """Even-odd [nf][2][Vh][nc] from the definition: half-site (t,z,y,xh) of parity p
    holds x = 2*xh + ((p+y+z+t) & 1)."""
    Lx, Ly, Lz, Lt = L
    nf, nc = lex.shape[0], lex.shape[-1]
    t, z, y, xh = np.indices((Lt, Lz, Ly, Lx // 2))
    out = np.empty((nf, 2, Lt, Lz, Ly, Lx // 2, nc), np.complex128)
    for p in range(2):
        x = 2 * xh + ((p + y + z + t) & 1)
        out[:, p] = lex[:, t, z, y, x]
    return out.reshape(nf, 2, -1, nc)


@pytest.mark.parametrize("L", LATTICES)
@pytest.mark.parametrize("nc,nf", [(12, 1), (9, 4)])
def test_lex_eo_known_answer(L, nc, nf):
    lex = coded_lex(L, nc, nf)
    eo = orc.lex_eo(lex.reshape(nf, -1, nc), L, nc, nf, True)
    np.testing.assert_array_equal(eo.reshape(nf, 2, -1, nc), expected_eo(lex, L))
    back = orc.lex_eo(eo, L, nc, nf, False)
    np.testing.assert_array_equal(back, lex.reshape(nf, -1, nc))


@pytest.mark.parametrize("L", LATTICES)
def test_point_source_lands_on_its_lexicographic_site(L):
    rng = np.random.default_rng(sum(L))
    for _ in range(4):
        c = [int(rng.integers(0, n)) for n in L]
        s, col = int(rng.integers(0, 4)), int(rng.integers(0, 3))
        eo = orc.source_point(L, c, s, col)
        lex = orc.lex_eo(eo.reshape(1, -1, 12), L, 12, 1, False).reshape(L[3], L[2], L[1], L[0], 4, 3)
        nz = np.argwhere(lex != 0)
        assert nz.tolist() == [[c[3], c[2], c[1], c[0], s, col]]


@pytest.mark.parametrize("L", LATTICES)
def test_shift_vector_known_answer(L):
    vh = orc.volume(L) // 2
    src = (np.arange(vh * 12) + 1j * (np.arange(vh * 12) + 0.5)).reshape(vh, 4, 3)
    soa = orc.shift_vector(src, 0, L)
    d = src.view(np.float64).reshape(vh, 24)
    for i in (0, 1, 7, 23):
        np.testing.assert_array_equal(soa[i], d[:, i])
    assert sorted(soa.ravel().tolist()) == sorted(d.ravel().tolist())
    np.testing.assert_array_equal(orc.shift_vector(soa, 1, L), src)


@pytest.mark.parametrize("L", LATTICES[:3])
def test_shift_gauge_known_answer_and_row2(L):
    vh = orc.volume(L) // 2
    g = orc.gauge_random(L, 5, 0.4)
    soa = orc.shift_gauge(g, 0, L)
    d = g.view(np.float64).reshape(4, 2, vh, 18)
    for mu in range(4):
        for p in range(2):
            for j in range(12):
                np.testing.assert_array_equal(soa[mu, p, j], d[mu, p, :, j])
    back = orc.shift_gauge(soa, 1, L)
    # rows 0-1 are moved exactly; row 2 = conj(row0 x row1) holds on SU(3) to rounding
    np.testing.assert_array_equal(back[..., :2, :], g[..., :2, :])
    assert np.abs(back - g).max() < 1e-14
    u = g[1, 0, 3]
    np.testing.assert_array_equal(back[1, 0, 3], orc.reconstruct12(u))
