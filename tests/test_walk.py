"""CPU model of the hopping kernel's tiled site walk (walk_site in k_hop.inc,
make_walk in qcu_common.h): a permutation of the half lattice whose 64-site waves
are aligned storage chunks, so reductions keep their grouping."""
import pytest


def make_walk(L, ty, tz, tt):
    lxh = L[0] // 2
    if ty <= 0:
        ty = L[1]
    sub = lxh * ty
    if tz <= 0 or tt <= 0 or L[1] % ty or sub % 64 or L[2] % tz or L[3] % tt:
        return None
    return (ty, tz, tt, L[1] // ty, L[2] // tz)


def walk_site(L, w, l):
    lxh = L[0] // 2
    sy, sz, st = lxh, lxh * L[1], lxh * L[1] * L[2]
    if w is None:
        return l
    ty, tz, tt, ny, ntz = w
    P, in_sub = divmod(l, lxh * ty)
    P2, t_in = divmod(P, tt)
    tile, z_in = divmod(P2, tz)
    rest, ytile = divmod(tile, ny)
    ttile, ztile = divmod(rest, ntz)
    z, t = ztile * tz + z_in, ttile * tt + t_in
    return t * st + z * sz + ytile * ty * sy + in_sub


def coords(L, s):
    lxh = L[0] // 2
    xh, r = s % lxh, s // lxh
    y, r = r % L[1], r // L[1]
    z, t = r % L[2], r // L[2]
    return xh, y, z, t


@pytest.mark.parametrize("L", [(16, 16, 16, 32), (32, 32, 32, 64), (24, 24, 24, 48), (16, 8, 8, 8), (8, 16, 4, 12),
                               (16, 16, 8, 16)])
@pytest.mark.parametrize("tile", [(0, 4, 4), (0, 8, 8), (0, 2, 4), (4, 4, 4), (8, 4, 4), (16, 4, 4), (8, 8, 8),
                                  (4, 2, 4)])
def test_walk_is_aligned_permutation(L, tile):
    vh = L[0] * L[1] * L[2] * L[3] // 2
    w = make_walk(L, *tile)
    if w is None:
        return
    step = max(1, vh // 20000)
    for l0 in range(0, vh, 64 * step):
        base = walk_site(L, w, l0)
        assert base % 64 == 0
        for k in range(64):
            s = walk_site(L, w, l0 + k)
            assert s == base + k and 0 <= s < vh
    if vh <= 1 << 17:
        assert {walk_site(L, w, l) for l in range(vh)} == set(range(vh))


@pytest.mark.parametrize("tile", [(4, 4, 4), (8, 2, 4), (0, 4, 4)])
def test_walk_keeps_neighbours_close(tile):
    """Inside a tile the t-neighbour of a site is one sub-plane later in walk order
    and the z-neighbour tile_t sub-planes later (the point of the tiling)."""
    L = (32, 32, 32, 64)
    w = make_walk(L, *tile)
    ty, tz, tt = w[0], w[1], w[2]
    sub = L[0] // 2 * ty
    pos = {}
    for l in range(0, 4 * tt * tz * sub):
        pos[walk_site(L, w, l)] = l
    for s, l in pos.items():
        xh, y, z, t = coords(L, s)
        if t % tt < tt - 1:
            st = s + (L[0] // 2) * L[1] * L[2]
            assert pos[st] - l == sub
        if z % tz < tz - 1:
            sz = s + (L[0] // 2) * L[1]
            assert pos[sz] - l == tt * sub


def test_walk_rejects_bad_tiles():
    assert make_walk((4, 4, 4, 4), 0, 4, 4) is None      # plane of 8 sites
    assert make_walk((16, 16, 6, 32), 0, 4, 4) is None   # 4 does not divide 6
    assert make_walk((16, 16, 16, 32), 0, 0, 0) is None
    assert make_walk((16, 16, 16, 32), 4, 4, 4) is None  # 32-site sub-planes
    assert make_walk((24, 24, 24, 48), 8, 4, 4) is None  # 96-site sub-planes
