"""Textbook lattice-QCD formulas in numpy, written independently of the oracle.

Used only by tests to pin the oracle.  Everything here works on lexicographic
fields (``psi[t, z, y, x, spin, colour]``, ``U[mu, t, z, y, x, 3, 3]``), builds
neighbours with ``np.roll`` and uses full 4x4 gamma matrices.  It shares none of the
oracle's code paths: no even-odd indexing, no spin projection, no coefficient tables.
"""
from __future__ import annotations

import numpy as np

# DeGrand-Rossi gamma matrices, as listed for the reference's hop projections
GAMMA = np.array(
    [
        [[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]],
        [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
        [[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]],
        [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],
    ],
    dtype=np.complex128,
)
# lexicographic array axis of direction mu for fields [t, z, y, x, ...]
AXIS = {0: 3, 1: 2, 2: 1, 3: 0}


def eo_coords(L, parity, n):
    """Full coordinates of half-sites 0..n-1 of one parity (vectorised)."""
    lxh = L[0] // 2
    idx = np.arange(n)
    xh = idx % lxh
    r = idx // lxh
    y = r % L[1]
    r //= L[1]
    z = r % L[2]
    t = r // L[2]
    x = 2 * xh + ((parity + y + z + t) & 1)
    return x, y, z, t


def eo_to_lex_spinor(v, L):
    """full even-odd spinor [2][Vh][4][3] -> lexicographic [t][z][y][x][4][3]."""
    vh = v.shape[1]
    out = np.zeros((L[3], L[2], L[1], L[0], 4, 3), np.complex128)
    for p in range(2):
        x, y, z, t = eo_coords(L, p, vh)
        out[t, z, y, x] = v[p]
    return out


def lex_to_eo_spinor(w, L):
    vh = np.prod(L) // 2
    out = np.zeros((2, vh, 4, 3), np.complex128)
    for p in range(2):
        x, y, z, t = eo_coords(L, p, vh)
        out[p] = w[t, z, y, x]
    return out


def eo_to_lex_gauge(g, L):
    """gauge [4][2][Vh][3][3] -> [4][t][z][y][x][3][3]"""
    vh = g.shape[2]
    out = np.zeros((4, L[3], L[2], L[1], L[0], 3, 3), np.complex128)
    for p in range(2):
        x, y, z, t = eo_coords(L, p, vh)
        for mu in range(4):
            out[mu, t, z, y, x] = g[mu, p]
    return out


def lex_to_eo_gauge(u, L):
    vh = np.prod(L) // 2
    out = np.zeros((4, 2, vh, 3, 3), np.complex128)
    for p in range(2):
        x, y, z, t = eo_coords(L, p, vh)
        for mu in range(4):
            out[mu, p] = u[mu, t, z, y, x]
    return out


def dag(m):
    return np.conj(np.swapaxes(m, -1, -2))


def shift(f, mu, step, tbc=1.0, spinor=False):
    """f(x + step*mu) with periodic wrap; antiperiodic factor tbc applied to the
    values that crossed the t edge (spinors only)."""
    ax = AXIS[mu]
    out = np.roll(f, -step, axis=ax)
    if spinor and mu == 3 and tbc != 1.0:
        out = out.copy()
        if step == 1:
            out[-1] *= tbc
        else:
            out[0] *= tbc
    return out


def colour_mul(U, psi):
    return np.einsum("...ab,...sb->...sa", U, psi)


def spin_mul(P, psi):
    return np.einsum("st,...tc->...sc", P, psi)


def hop_lex(psi, U, s=1, tbc=1.0):
    """D^s psi(x) = sum_mu U_mu(x)(1 - s g_mu) psi(x+mu) + U_mu(x-mu)^dag (1 + s g_mu) psi(x-mu)."""
    one = np.eye(4)
    out = np.zeros_like(psi)
    for mu in range(4):
        fwd = shift(psi, mu, 1, tbc, spinor=True)
        out += colour_mul(U[mu], spin_mul(one - s * GAMMA[mu], fwd))
        bwd = shift(psi, mu, -1, tbc, spinor=True)
        ub = dag(shift(U[mu], mu, -1))
        out += colour_mul(ub, spin_mul(one + s * GAMMA[mu], bwd))
    return out


def plaquette_leaves(U, mu, nu):
    """Q_munu(x): sum of the four oriented plaquettes in the (mu,nu) plane that
    start and end at x."""
    Um, Un = U[mu], U[nu]
    m = lambda f, a, b: shift(shift(f, mu, a), nu, b)  # f(x + a mu + b nu)
    q = Um @ m(Un, 1, 0) @ dag(m(Um, 0, 1)) @ dag(Un)
    q = q + Un @ dag(m(Um, -1, 1)) @ dag(m(Un, -1, 0)) @ m(Um, -1, 0)
    q = q + dag(m(Um, -1, 0)) @ dag(m(Un, -1, -1)) @ m(Um, -1, -1) @ m(Un, 0, -1)
    q = q + dag(m(Un, 0, -1)) @ m(Um, 0, -1) @ m(Un, 1, -1) @ dag(Um)
    return q


def clover_sigma_F(U):
    """T(x) = -(1/2) sum_{mu,nu} sigma_munu F_munu as a 12x12 (spin x colour) matrix
    per site, with sigma_munu = (i/2)[g_mu, g_nu] and F_munu = (Q - Q^dag)/(8i)."""
    shape = U.shape[1:5]
    T = np.zeros(shape + (12, 12), np.complex128)
    for mu in range(4):
        for nu in range(4):
            if mu == nu:
                continue
            sig = 0.5j * (GAMMA[mu] @ GAMMA[nu] - GAMMA[nu] @ GAMMA[mu])
            Q = plaquette_leaves(U, mu, nu)
            F = (Q - dag(Q)) / 8j
            T += -0.5 * np.einsum("st,...ab->...satb", sig, F).reshape(shape + (12, 12))
    return T


def T_to_blocks(T12, L):
    """12x12 per-site lexicographic clover -> even-odd [2][Vh][2][6][6] chiral blocks
    (checks that the off-diagonal chiral blocks vanish)."""
    vh = np.prod(L) // 2
    out = np.zeros((2, vh, 2, 6, 6), np.complex128)
    off = 0.0
    for p in range(2):
        x, y, z, t = eo_coords(L, p, vh)
        m = T12[t, z, y, x]
        out[p, :, 0] = m[:, 0:6, 0:6]
        out[p, :, 1] = m[:, 6:12, 6:12]
        off = max(off, np.abs(m[:, 0:6, 6:12]).max(), np.abs(m[:, 6:12, 0:6]).max())
    return out, off


def dense_operator(U, L, kappa, s=1, tbc=1.0, A_blocks=None):
    """Dense (A - kappa D^s) on the even-odd ordered full lattice, index
    ((parity*Vh + idx)*4 + spin)*3 + colour.  A_blocks: [2][Vh][2][6][6] or None."""
    vh = np.prod(L) // 2
    n = 2 * vh * 12
    M = np.zeros((n, n), np.complex128)
    lex_index = np.zeros((L[3], L[2], L[1], L[0]), np.int64)
    for p in range(2):
        x, y, z, t = eo_coords(L, p, vh)
        lex_index[t, z, y, x] = p * vh + np.arange(vh)
    one = np.eye(4)
    for t in range(L[3]):
        for z in range(L[2]):
            for y in range(L[1]):
                for x in range(L[0]):
                    row = lex_index[t, z, y, x]
                    c = [x, y, z, t]
                    for mu in range(4):
                        for step in (1, -1):
                            d = list(c)
                            d[mu] += step
                            sign = 1.0
                            if mu == 3 and (d[3] < 0 or d[3] >= L[3]):
                                sign = tbc
                            d = [d[k] % L[k] for k in range(4)]
                            col = lex_index[d[3], d[2], d[1], d[0]]
                            if step == 1:
                                blk = np.kron(one - s * GAMMA[mu], U[mu, t, z, y, x])
                            else:
                                blk = np.kron(one + s * GAMMA[mu], dag(U[mu, d[3], d[2], d[1], d[0]]))
                            M[row * 12:(row + 1) * 12, col * 12:(col + 1) * 12] -= kappa * sign * blk
    for site in range(2 * vh):
        sl = slice(site * 12, (site + 1) * 12)
        if A_blocks is None:
            M[sl, sl] += np.eye(12)
        else:
            p, i = divmod(site, vh)
            M[site * 12:site * 12 + 6, site * 12:site * 12 + 6] += A_blocks[p, i, 0]
            M[site * 12 + 6:site * 12 + 12, site * 12 + 6:site * 12 + 12] += A_blocks[p, i, 1]
    return M


def random_su3(rng, n):
    """n Haar-ish random SU(3) matrices (QR of complex Gaussians, det fixed to 1)."""
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q = q / (det ** (1.0 / 3.0))[:, None, None]
    return q
