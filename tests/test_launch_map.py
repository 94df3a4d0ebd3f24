"""CPU model of the hopping kernel's thread -> (site, right-hand side) map (hop_kernel
and launch_hop_prec in k_hop.inc, remap_block in qcu_fields.h): for every launch shape
each (site, right-hand side) pair is computed by exactly one live thread, every load
index stays inside the batch, and each 16-site reduction partial is written once."""
import itertools

import pytest


def remap_block(b, nb, on):
    if not on:
        return b
    x, k, q, r = b & 7, b >> 3, nb >> 3, nb & 7
    return x * q + (x if x < r else r) + k


def launch(vh, nrhs, block, rb, rpw, rpt, gm, xcd, spt=1):
    if rpw not in (2, 4) or rpt > 1:
        rpw = 1
    sg = block // rb
    per_block = rb * rpw * rpt
    ngroups = -(-nrhs // per_block)
    sb = (sg // 64) * (64 // rpw)
    nsb = -(-vh // sb)
    nbx = -(-nsb // spt)
    grid = nbx * ngroups
    covered, partials = {}, {}
    for bid, tid, it in itertools.product(range(grid), range(block), range(spt)):
        lb = remap_block(bid, grid, xcd)
        if gm:
            grp, bx0 = lb // nbx, lb % nbx
        else:
            bx0, grp = lb // ngroups, lb % ngroups
        bx = bx0 + it * nbx
        if bx >= nsb:
            continue
        rl, tg = tid // sg, tid % sg
        lane = tg & 63
        spw = 64 // rpw
        sj = lane // spw
        rq = ((grp * rb + rl) * rpw + sj) * rpt
        have = rq < nrhs
        r0 = rq if have else 0
        nr = min(nrhs - r0, rpt) if have else 1
        assert 0 <= r0 and r0 + nr <= nrhs          # loads stay in the batch
        l0 = bx * sb + (tg >> 6) * spw + (lane - sj * spw)
        site = l0 if l0 < vh else vh - 1            # storage-order walk
        assert 0 <= site < vh
        live = have and l0 < vh
        if live:
            for j in range(nr):
                key = (site, r0 + j)
                assert key not in covered
                covered[key] = (bid, tid)
                if (lane & 15) == 0:
                    pk = (r0 + j, site // 16)
                    assert pk not in partials and site % 16 == 0
                    partials[pk] = True
    assert len(covered) == vh * nrhs
    assert len(partials) == nrhs * (-(-vh // 16))


SHAPES = [(64, 1, 1, 1, 1), (256, 4, 1, 1, 2), (192, 3, 4, 1, 1), (64, 1, 4, 1, 4), (128, 2, 2, 1, 3),
          (256, 4, 4, 1, 1), (128, 1, 2, 2, 2), (256, 2, 4, 3, 1), (64, 1, 1, 4, 4)]


@pytest.mark.parametrize("vh", [8, 48, 1000, 2048])
@pytest.mark.parametrize("nrhs", [1, 2, 5, 12, 16])
@pytest.mark.parametrize("shape", SHAPES)
def test_launch_covers_each_pair_once(vh, nrhs, shape):
    block, rb, rpw, rpt, spt = shape
    for gm, xcd in itertools.product((0, 1), (0, 1)):
        launch(vh, nrhs, block, rb, rpw, rpt, gm, xcd, spt)


def colgroup_block(b, w, q, r, np_):
    """colgroup_block in qcu_fields.h: XCD x = b % 8 takes w consecutive chunks of every
    q-chunk sub-plane in one of r ranges of the np sub-planes."""
    x, k = b & 7, b >> 3
    per = q // w
    rr, gx = x // per, x % per
    pl, i = k // w, k % w
    P = rr * (np_ // r) + pl
    return P * q + gx * w + i


def colgroup_params(vh, sub, w, spt=1):
    """The launcher's validity check (launch_hop_prec in k_hop.inc); None: round-robin."""
    q = sub // 64
    np_ = vh // sub if sub % 64 == 0 else 0
    nsb = -(-vh // 64)
    if sub % 64 or q % w or (8 * w) % q or np_ % ((8 * w) // q) or spt != 1 or np_ * q != nsb:
        return None
    return w, q, (8 * w) // q, np_


@pytest.mark.parametrize("L", [(32, 32, 32, 64), (16, 16, 16, 32), (32, 16, 8, 16), (16, 16, 8, 16)])
@pytest.mark.parametrize("w", [2, 4])
def test_xcd_column_groups(L, w):
    """The XCD column-group order of the LDS kernel is a permutation of the 64-site
    chunks, and each XCD gets w consecutive chunks of every sub-plane of its range."""
    lxh = L[0] // 2
    vh = lxh * L[1] * L[2] * L[3]
    sub = lxh * L[1]  # whole xy-planes (the tiles of all z by tt t-planes)
    p = colgroup_params(vh, sub, w)
    if p is None:
        pytest.skip("column groups do not tile this walk; the launcher uses round-robin")
    w_, q, r, np_ = p
    nsb = vh // 64
    img = [colgroup_block(b, w_, q, r, np_) for b in range(nsb)]
    assert sorted(img) == list(range(nsb))
    for x in range(8):
        mine = sorted(img[b] for b in range(x, nsb, 8))
        by_plane = {}
        for c in mine:
            by_plane.setdefault(c // q, []).append(c % q)
        assert all(len(v) == w and v == list(range(v[0], v[0] + w)) for v in by_plane.values())
        planes = sorted(by_plane)
        assert planes == list(range(planes[0], planes[0] + len(planes)))  # one range of sub-planes
