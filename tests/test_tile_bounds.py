"""Host-side bounds check of the addresses the sweep kernel's FP64-MFMA tile loops form
(ttns_lib_amd/csrc/device/wg_ops.hpp: t_mode_contract, t_mul_transposed, t_mul,
wg_outer_contract with t_outer_contract's K splits over team members; device/zgemm.hpp: ztile
reads A(i, k) at a + i sar + k sak for i < mv, k < K and B(k, j) at b + k sbk + j sbc for j < nv,
ztile_store writes C(row, col) at c + row scr + col scc).  The index arithmetic is restated here
line for line and every tile's largest address is compared with the extent of the buffer the
call sites pass: node tensors and work vectors hold L d R = the node's size, operator and
environment matrices d x d, a member's partial product TILE_MAX = 64 x 64 entries.  Shapes: every
node and child mode of the 64 batched instances, the single larger run's tree, and the 64^3
contraction node, for team sizes 1, 2, 5, 16 and 32.  All strides are positive, so a tile's
largest address is at its last row and last k; its smallest is its base, never negative."""
import numpy as np
import pytest

from helpers import tree_children
from ttns_lib_amd.capi import Library, sb_problem

TILE_MAX = 64 * 64


def ztile_max(a, sar, sak, mv, b, sbk, sbc, nv, K):
    return a + (mv - 1) * sar + (K - 1) * sak, b + (K - 1) * sbk + (nv - 1) * sbc


def mode_contract(L, d, R):
    """(op, src, dst) largest offsets of t_mode_contract's tiles"""
    mt, nt = (d + 15) >> 4, (R + 15) >> 4
    m, n, l = np.meshgrid(np.arange(mt), np.arange(nt), np.arange(L), indexing="ij")
    mv = np.minimum(16, d - 16 * m)
    nv = np.minimum(16, R - 16 * n)
    s = l * d * R + 16 * n
    oa, ob = ztile_max(16 * m * d, d, 1, mv, s, R, 1, nv, d)
    dst = l * d * R + 16 * m * R + 16 * n + (mv - 1) * R + (nv - 1)
    return oa.max(), ob.max(), dst.max()


def mul_transposed(rows, r):
    mt, nt = (rows + 15) >> 4, (r + 15) >> 4
    m, n = np.meshgrid(np.arange(mt), np.arange(nt), indexing="ij")
    mv, nv = np.minimum(16, rows - 16 * m), np.minimum(16, r - 16 * n)
    oa, ob = ztile_max(16 * m * r, r, 1, mv, 16 * n * r, 1, r, nv, r)
    return oa.max(), ob.max(), (16 * m * r + 16 * n + (mv - 1) * r + nv - 1).max()


def mul(rows, r):
    mt, nt = (rows + 15) >> 4, (r + 15) >> 4
    m, n = np.meshgrid(np.arange(mt), np.arange(nt), indexing="ij")
    mv, nv = np.minimum(16, rows - 16 * m), np.minimum(16, r - 16 * n)
    oa, ob = ztile_max(16 * m * r, r, 1, mv, 16 * n, r, 1, nv, r)
    return oa.max(), ob.max(), (16 * m * r + 16 * n + (mv - 1) * r + nv - 1).max()


def wg_outer(L, d, R, ldr, base=0):
    """largest A/B offsets (relative to the caller's A) and out offset of wg_outer_contract"""
    nwave = 4
    mt = (d + 15) >> 4
    tiles = mt * mt
    split = tiles < nwave
    amax = omax = -1
    for t in range(tiles):
        m, n = t // mt, t % mt
        mv, nv = min(16, d - 16 * m), min(16, d - 16 * n)
        if R == 1 and ldr == 1:
            for wv in (range(nwave) if split else [0]):
                k0, k1 = ((L * wv) // nwave, (L * (wv + 1)) // nwave) if split else (0, L)
                if k1 > k0:
                    oa, ob = ztile_max(base + k0 * d + 16 * m, 1, d, mv, base + k0 * d + 16 * n, d, 1, nv, k1 - k0)
                    amax = max(amax, oa, ob)
        elif not split or L >= nwave:
            for l in range(L):
                a, b = base + (l * d + 16 * m) * ldr, base + (l * d + 16 * n) * ldr
                oa, ob = ztile_max(a, ldr, 1, mv, b, 1, ldr, nv, R)
                amax = max(amax, oa, ob)
        else:
            for wv in range(nwave):
                k0, k1 = (R * wv) // nwave, (R * (wv + 1)) // nwave
                for l in range(L):
                    if k1 > k0:
                        a, b = base + (l * d + 16 * m) * ldr + k0, base + (l * d + 16 * n) * ldr + k0
                        oa, ob = ztile_max(a, ldr, 1, mv, b, 1, ldr, nv, k1 - k0)
                        amax = max(amax, oa, ob)
        omax = max(omax, 16 * m * d + 16 * n + (mv - 1) * d + nv - 1)
    return amax, omax


def t_outer(L, d, R, team):
    """(largest A/B offset, largest out offset, largest offset in a member's tile buffer)"""
    if team == 1:
        a, o = wg_outer(L, d, R, R)
        return a, o, -1
    amax = tmax = -1
    for rank in range(team):
        if L >= team:
            l0, l1 = L * rank // team, L * (rank + 1) // team
            a, o = wg_outer(l1 - l0, d, R, R, base=l0 * d * R) if l1 > l0 else (-1, -1)
        else:
            r0, r1 = R * rank // team, R * (rank + 1) // team
            a, o = wg_outer(L, d, r1 - r0, R, base=r0) if r1 > r0 else (-1, -1)
        amax, tmax = max(amax, a), max(tmax, o)
    return amax, d * d - 1, tmax


def node_views(tree):
    """(L, d, R) of every child mode of every branch node, and (rows, r) of every node"""
    info, kids = tree_children(tree)
    views, mats = [], []
    for i, nd in enumerate(info):
        r = nd["hrank"]
        mats.append((nd["rows"], r))
        if nd["nchild"] == 0:
            continue
        dims = [info[c]["hrank"] for c in kids[i]]
        for c in range(len(dims)):
            L = int(np.prod(dims[:c])) if c else 1
            R = int(np.prod(dims[c + 1:])) * r
            views.append((L, dims[c], R))
    return views, mats


def _check(views, mats, teams):
    for L, d, R in views:
        size = L * d * R
        oa, ob, od = mode_contract(L, d, R)
        assert oa < d * d and ob < size and od < size, (L, d, R)
        for team in teams:
            a, o, t = t_outer(L, d, R, team)
            assert a < size and o < d * d and t < TILE_MAX, (L, d, R, team)
    for rows, r in mats:
        for f in (mul_transposed, mul):
            x, m, dst = f(rows, r)
            assert x < rows * r and m < r * r and dst < rows * r, (rows, r)
        for team in teams:
            a, o, t = t_outer(rows, r, 1, team)
            assert a < rows * r and o < r * r and t < TILE_MAX, (rows, r, team)


def test_tile_addresses_stay_in_bounds_on_the_batched_instances(orc):
    from ttns_lib_amd.workloads import c1_instance, c3_instances

    shapes = set()
    for inst in c3_instances() + [c1_instance()]:
        tree, _ = sb_problem(orc, inst.model)
        v, m = node_views(tree)
        shapes.update(("v",) + x for x in v)
        shapes.update(("m",) + x for x in m)
    views = [s[1:] for s in shapes if s[0] == "v"]
    mats = [s[1:] for s in shapes if s[0] == "m"]
    assert len(views) > 50 and max(r for _, r in mats) == 48
    _check(views, mats, (1, 2, 5, 16, 32))


def test_tile_addresses_stay_in_bounds_on_edge_shapes():
    # the 64^3 contraction node, the device's limits (64-wide modes and bonds), ragged sizes
    views = [(1, 64, 4096), (64, 64, 64), (4096, 64, 1), (1, 19, 19 * 43), (19, 19, 43), (1, 1, 1), (3, 17, 33)]
    mats = [(4096, 64), (361, 43), (1, 1), (17, 3), (900, 8)]
    _check(views, mats, (1, 2, 3, 16, 32))


def test_the_check_catches_an_unclamped_tile():
    # a tile loop that forgot to clamp its last column tile (nv = 16 on R = 33) reads past the
    # node: the same arithmetic with that mistake planted must fail the bound
    L, d, R = 3, 17, 33
    mt, nt = (d + 15) >> 4, (R + 15) >> 4
    worst = max(ztile_max(16 * m * d, d, 1, min(16, d - 16 * m), l * d * R + 16 * n, R, 1, 16, d)[1]
                for m in range(mt) for n in range(nt) for l in range(L))
    assert worst >= L * d * R
    assert mode_contract(L, d, R)[1] < L * d * R
