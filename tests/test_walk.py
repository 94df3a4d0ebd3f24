"""CPU model of the hopping kernel's tiled site walk (walk_site in k_hop.inc,
make_walk in qcu_common.h): a permutation of the half lattice whose 64-site waves
are aligned storage chunks, so reductions keep their grouping."""
import pytest


def make_walk(L, tz, tt):
    lxh = L[0] // 2
    plane = lxh * L[1]
    if tz <= 0 or tt <= 0 or plane % 64 or L[2] % tz or L[3] % tt:
        return None
    return (tz, tt, L[2] // tz)


def walk_site(L, w, l):
    lxh = L[0] // 2
    plane = lxh * L[1]
    sz, st = plane, plane * L[2]
    if w is None:
        return l
    tz, tt, ntz = w
    P, in_plane = divmod(l, plane)
    P2, t_in = divmod(P, tt)
    tile, z_in = divmod(P2, tz)
    ttile, ztile = divmod(tile, ntz)
    z, t = ztile * tz + z_in, ttile * tt + t_in
    return t * st + z * sz + in_plane


@pytest.mark.parametrize("L", [(16, 16, 16, 32), (32, 32, 32, 64), (24, 24, 24, 48), (16, 8, 8, 8), (8, 16, 4, 12)])
@pytest.mark.parametrize("tile", [(4, 4), (8, 8), (2, 4)])
def test_walk_is_aligned_permutation(L, tile):
    vh = L[0] * L[1] * L[2] * L[3] // 2
    w = make_walk(L, *tile)
    if w is None:
        return
    step = max(1, vh // 20000)
    seen = set()
    for l0 in range(0, vh, 64 * step):
        base = walk_site(L, w, l0)
        assert base % 64 == 0
        for k in range(64):
            s = walk_site(L, w, l0 + k)
            assert s == base + k and 0 <= s < vh
    if vh <= 1 << 17:
        seen = {walk_site(L, w, l) for l in range(vh)}
        assert seen == set(range(vh))


def test_walk_rejects_bad_tiles():
    assert make_walk((4, 4, 4, 4), 4, 4) is None      # plane of 8 sites
    assert make_walk((16, 16, 6, 32), 4, 4) is None   # 4 does not divide 6
    assert make_walk((16, 16, 16, 32), 0, 0) is None
