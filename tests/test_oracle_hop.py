"""Pins the oracle's geometry, reconstruct-12 and hopping term against the numpy
textbook formulation, closed forms and symmetries (no reference golden vectors exist
for this path)."""
import itertools

import numpy as np
import pytest

import textbook as tb

LATTICES = [(4, 4, 4, 4), (2, 2, 2, 2), (6, 4, 2, 4), (4, 2, 6, 2)]


@pytest.mark.parametrize("L", LATTICES)
def test_neighbor_matches_bruteforce(orc, L):
    vh = orc.half_volume(L)
    for p in range(2):
        for i in range(vh):
            c = orc.coords(L, p, i)
            assert sum(c) % 2 == p
            assert orc.index(L, c) == (p, i)
            for mu in range(4):
                for fwd in (True, False):
                    d = list(c)
                    d[mu] += 1 if fwd else -1
                    wrapped = d[mu] < 0 or d[mu] >= L[mu]
                    q, j = orc.index(L, d)
                    n, w = orc.neighbor(L, p, i, mu, fwd)
                    assert (q, n, w) == (1 - p, j, int(wrapped)), (L, p, i, mu, fwd)


def test_gamma_algebra():
    g = tb.GAMMA
    for mu in range(4):
        assert np.allclose(g[mu], g[mu].conj().T)
        for nu in range(4):
            ac = g[mu] @ g[nu] + g[nu] @ g[mu]
            assert np.allclose(ac, 2 * np.eye(4) * (mu == nu))
    g5 = g[0] @ g[1] @ g[2] @ g[3]
    assert np.allclose(g5, np.diag([1, 1, -1, -1]))


def test_reconstruct12_exact_on_su3(orc):
    rng = np.random.default_rng(1)
    for u in tb.random_su3(rng, 50):
        v = u.copy()
        v[2] = 0
        assert np.allclose(orc.reconstruct12(v), u, atol=1e-14)


@pytest.mark.parametrize("L", LATTICES)
@pytest.mark.parametrize("s", [1, -1])
@pytest.mark.parametrize("tbc", [1.0, -1.0])
def test_hop_matches_textbook(orc, L, s, tbc):
    rng = np.random.default_rng(7)
    g = orc.gauge_random(L, seed=11, eps=0.7)
    v = orc.source_gauss(L, seed=5)
    U = tb.eo_to_lex_gauge(g, L)
    psi = tb.eo_to_lex_spinor(v, L)
    ref = tb.lex_to_eo_spinor(tb.hop_lex(psi, U, s=s, tbc=tbc), L)
    for p in range(2):
        for recon in (12, 18):
            out = orc.hop(v[1 - p], g, L, p, s=s, tbc=tbc, recon=recon)
            err = np.abs(out - ref[p]).max() / np.abs(ref[p]).max()
            assert err < 1e-14, (p, recon, err)
    del rng


@pytest.mark.parametrize("tbc", [1.0, -1.0])
def test_hop_free_field_plane_wave(orc, tbc):
    """U = 1, psi = e^{ip.x} chi:  D^s psi = sum_mu (2 cos p_mu - 2 i s g_mu sin p_mu) psi."""
    L = (4, 6, 4, 8)
    g = orc.gauge_unit(L)
    chi = (np.arange(12) + 1j * np.arange(12)[::-1]).reshape(4, 3) / 7.0
    tshift = 0.5 if tbc < 0 else 0.0
    n = (1, 2, 3, 1)
    pm = [2 * np.pi * n[0] / L[0], 2 * np.pi * n[1] / L[1], 2 * np.pi * n[2] / L[2],
          2 * np.pi * (n[3] + tshift) / L[3]]
    vh = orc.half_volume(L)
    for s in (1, -1):
        K = sum(2 * np.cos(pm[m]) * np.eye(4) - 2j * s * tb.GAMMA[m] * np.sin(pm[m]) for m in range(4))
        for p in range(2):
            x = np.array(tb.eo_coords(L, 1 - p, vh))
            ph = np.exp(1j * (pm[0] * x[0] + pm[1] * x[1] + pm[2] * x[2] + pm[3] * x[3]))
            psi = ph[:, None, None] * chi[None]
            out = orc.hop(psi, g, L, p, s=s, tbc=tbc)
            xo = np.array(tb.eo_coords(L, p, vh))
            pho = np.exp(1j * (pm[0] * xo[0] + pm[1] * xo[1] + pm[2] * xo[2] + pm[3] * xo[3]))
            expect = pho[:, None, None] * (K @ chi)[None]
            assert np.abs(out - expect).max() < 1e-13


def test_hop_gauge_covariance(orc):
    L = (4, 4, 2, 6)
    rng = np.random.default_rng(3)
    g = orc.gauge_random(L, seed=2, eps=0.5)
    v = orc.source_gauss(L, seed=9)
    U = tb.eo_to_lex_gauge(g, L)
    G = tb.random_su3(rng, int(np.prod(L))).reshape(L[3], L[2], L[1], L[0], 3, 3)
    Up = np.stack([G @ U[mu] @ tb.dag(tb.shift(G, mu, 1)) for mu in range(4)])
    gp = tb.lex_to_eo_gauge(Up, L)
    psi = tb.eo_to_lex_spinor(v, L)
    vp = tb.lex_to_eo_spinor(tb.colour_mul(G, psi), L)
    for p in range(2):
        for s in (1, -1):
            a = orc.hop(vp[1 - p], gp, L, p, s=s, tbc=-1.0)
            b = orc.hop(v[1 - p], g, L, p, s=s, tbc=-1.0)
            full = np.zeros_like(v)
            full[p] = b
            gb = tb.lex_to_eo_spinor(tb.colour_mul(G, tb.eo_to_lex_spinor(full, L)), L)[p]
            assert np.abs(a - gb).max() < 1e-13


def test_hop_adjointness(orc):
    """<phi_p, D^{+} psi_{1-p}> = <D^{-}_{1-p,p} phi_p, psi_{1-p}>"""
    L = (4, 4, 4, 4)
    g = orc.gauge_random(L, seed=21, eps=1.0)
    a = orc.source_gauss(L, seed=1)
    b = orc.source_gauss(L, seed=2)
    for p in range(2):
        for tbc in (1.0, -1.0):
            lhs = np.vdot(a[p], orc.hop(b[1 - p], g, L, p, s=1, tbc=tbc))
            rhs = np.vdot(orc.hop(a[p], g, L, 1 - p, s=-1, tbc=tbc), b[1 - p])
            assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_hop_gamma5_hermiticity(orc):
    L = (4, 2, 4, 2)
    g = orc.gauge_random(L, seed=4, eps=0.3)
    v = orc.source_gauss(L, seed=3)
    g5 = np.array([1, 1, -1, -1])[None, :, None]
    for p in range(2):
        a = orc.hop(v[1 - p], g, L, p, s=-1)
        b = g5 * orc.hop(g5 * v[1 - p], g, L, p, s=1)
        assert np.abs(a - b).max() < 1e-14


def test_hop_linearity(orc):
    L = (4, 4, 2, 2)
    g = orc.gauge_random(L, seed=8, eps=0.4)
    a = orc.source_gauss(L, seed=1)[0]
    b = orc.source_gauss(L, seed=2)[0]
    z = 0.3 - 1.7j
    lhs = orc.hop(a + z * b, g, L, 1)
    rhs = orc.hop(a, g, L, 1) + z * orc.hop(b, g, L, 1)
    assert np.abs(lhs - rhs).max() < 1e-13


def test_hop_single_site_support(orc):
    """A delta source at one odd site reaches exactly its 8 (distinct) even neighbours."""
    L = (6, 4, 4, 4)
    g = orc.gauge_random(L, seed=5, eps=0.2)
    vh = orc.half_volume(L)
    for j in (0, 17, vh - 1):
        v = np.zeros((vh, 4, 3), np.complex128)
        v[j, 2, 1] = 1.0
        out = orc.hop(v, g, L, 0)
        hit = set(np.nonzero(np.abs(out).reshape(vh, -1).max(axis=1))[0])
        c = orc.coords(L, 1, j)
        expect = set()
        for mu, st in itertools.product(range(4), (1, -1)):
            d = list(c)
            d[mu] += st
            expect.add(orc.index(L, d)[1])
        assert hit == expect
