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
