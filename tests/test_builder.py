"""Builder + flattener: structural invariants of the FastNode arrays and
hand-derived known answers for the reference's SAH split rules
(aabb.cpp:4-206, aabb.h:205-222, node.cpp:254-275, fastNodeManager.h:1495-1688).
"""
import math

import numpy as np
import pytest


def child_count(bvh, i):
    return int(bvh.order[i, 0].max()) + 1


def check_invariants(bvh, tris, N, L, G):
    M, Lm = bvh.node_memory, bvh.leaf_memory
    assert M == math.ceil(N / G) * G and Lm == math.ceil(L / G) * G
    leaf = bvh.is_leaf()
    n = bvh.node_count
    # every input triangle sits in exactly one leaf slot
    ids = bvh.tri_ids[bvh.tri_ids >= 0]
    assert np.array_equal(np.sort(ids), np.arange(len(tris)))
    nleaves = len(bvh.tri_points) // (9 * Lm)
    assert nleaves == leaf.sum()
    per_leaf = (bvh.tri_ids.reshape(nleaves, Lm) >= 0).sum(1)
    assert per_leaf.max() <= L and per_leaf.min() >= 1
    # customTreeOrder: replay the recursion from the arrays and get 0..n-1
    seq = [0]

    def rec(i):
        if leaf[i]:
            return
        c = child_count(bvh, i)
        kids = list(range(bvh.base[i], bvh.base[i] + c))
        seq.extend(kids)
        for k in kids:
            rec(k)

    rec(0)
    assert seq == list(range(n))
    tp = bvh.tri_points
    for i in range(n):
        b = bvh.bounds[i]
        if leaf[i]:
            pib = bvh.base[i]
            ordn = pib // (9 * Lm)
            pc = per_leaf[ordn]
            lrs = math.ceil(pc / G) * G
            blk = tp[pib: pib + 9 * Lm]
            if lrs != Lm:
                assert np.isnan(blk[9 * Lm - 1]) and blk[9 * Lm - 2] == lrs
            # the leaf's triangles, read back at stride lrs, are the input ones
            got = np.stack([blk[v * lrs: v * lrs + pc] for v in range(9)], 1)
            assert np.array_equal(got, tris[bvh.tri_ids[ordn * Lm: ordn * Lm + pc]])
            continue
        c = child_count(bvh, i)
        assert 2 <= c <= N
        nrs = math.ceil(c / G) * G
        if nrs != M:
            assert np.isnan(b[6 * M - 1]) and b[6 * M - 2] == nrs
        mn = np.stack([b[a * nrs: a * nrs + c] for a in range(3)], 1)
        mx = np.stack([b[3 * nrs + a * nrs: 3 * nrs + a * nrs + c] for a in range(3)], 1)
        for a in range(3):
            row = bvh.order[i, a]
            assert sorted(row[:c].tolist()) == list(range(c)) and (row[c:] == 0).all()
            key = (mx[:, a] + mn[:, a]) + np.float32(0.5)
            ks = [(key[j], j) for j in row[:c]]
            assert ks == sorted(ks)
        for j in range(c):
            ch = bvh.base[i] + j
            assert bool((bvh.child_type[i] >> j) & 1) == (not leaf[ch])
            if leaf[ch]:
                pib = bvh.base[ch]
                o = pib // (9 * Lm)
                pts = tris[bvh.tri_ids[o * Lm: o * Lm + per_leaf[o]]].reshape(-1, 3)
            else:
                cc = child_count(bvh, ch)
                cn = math.ceil(cc / G) * G
                cb = bvh.bounds[ch]
                pts = np.concatenate([np.stack([cb[a * cn: a * cn + cc] for a in range(3)], 1),
                                      np.stack([cb[3 * cn + a * cn: 3 * cn + a * cn + cc] for a in range(3)], 1)])
            assert (pts >= mn[j]).all() and (pts <= mx[j]).all()


@pytest.mark.parametrize("N,L,G", [(2, 1, 4), (4, 4, 4), (8, 8, 4), (12, 16, 4), (16, 4, 4),
                                   (6, 3, 4), (8, 8, 8), (16, 16, 16)])
def test_invariants(rt, N, L, G):
    tris = rt.interior_scene(9, 4000, 8)
    bvh = rt.Bvh(tris, N, L, gang=G)
    check_invariants(bvh, tris, N, L, G)


def test_invariants_spheres(rt):
    tris = rt.spheres_scene(0, spheres=6, frequency=5)
    bvh = rt.Bvh(tris, 4, 4)
    check_invariants(bvh, tris, 4, 4, 4)


def test_deterministic_across_threads(rt):
    """300k triangles: the ranges near the root pass 2^17, so the chunked
    centroid scan, sort, SAH sweeps and the threaded flattening all run; the
    arrays must equal the one-thread build byte for byte."""
    tris = rt.interior_scene(4, 300_000)
    a = rt.Bvh(tris, 8, 8, threads=1)
    for th in (3, 8):
        b = rt.Bvh(tris, 8, 8, threads=th)
        for f in ("bounds", "base", "order", "child_type", "tri_points", "tri_ids"):
            x, y = getattr(a, f), getattr(b, f)
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (th, f)
        assert a.average_depth == b.average_depth and a.max_depth == b.max_depth


def small_tri(x, s):
    return [x, 0, 0, x + s, s, 0, x, 0, s]


def test_left_sweep_off_by_one_decides_split(rt):
    """Three tiny triangles A (x=0), B (x=1), C (x=100, larger).

    sweepRight reads *(++primitiveEnd) (aabb.h:205-213), so the left box for
    a split after k prims covers prims 1..k: metric[0] = SA(B) + 2 SA(B+C),
    metric[1] = 2 SA(B+C) + SA(C) -> split {A} | {B, C}.  A correct sweep
    would give {A, B} | {C}.  With N=2, L=1, leafSplitOption 0 the flattened
    tree is root -> [leaf A, inner -> [leaf B, leaf C]]."""
    tris = np.array([small_tri(0, 0.1), small_tri(1, 0.1), small_tri(100, 0.2)], np.float32)
    bvh = rt.Bvh(tris, 2, 1, leaf_split_option=0, sah_factor=1.0)
    leaf = bvh.is_leaf()
    assert bvh.node_count == 5
    assert list(leaf) == [False, True, False, True, True]
    assert bvh.base[0] == 1 and bvh.base[2] == 3
    Lm = bvh.leaf_memory
    first = lambda i: bvh.tri_ids[bvh.base[i] // (9 * Lm) * Lm]  # noqa: E731
    assert [first(1), first(3), first(4)] == [0, 1, 2]


def test_leaf_split_option(rt):
    """A big triangle A (box [0,5]^3, centre x 2.5) and a tiny one B at x=10,
    L=4.  Option 0 returns at once for count <= L (aabb.cpp:80-83): one leaf.
    Option 1 tries a split (aabb.cpp:124-175); with the off-by-one left sweep
    both halves of the only split are B's box, so SAH = 2*4*SA(B)/SA(root) ~
    0.002 and 2 - (0.002 + 0.979) > 0: two leaves.  (Two separated clusters
    would NOT split: both halves of every split contain prim m+1.)"""
    A = [0, 0, 0, 5, 5, 0, 0, 5, 5]
    tris = np.array([A, small_tri(10, 0.1)], np.float32)
    b0 = rt.Bvh(tris, 4, 4, leaf_split_option=0)
    assert b0.node_count == 1 and b0.is_leaf()[0]
    b1 = rt.Bvh(tris, 4, 4, leaf_split_option=1)
    assert b1.node_count == 3 and list(b1.is_leaf()) == [False, True, True]
    pairs = np.array([small_tri(0, 0.1), small_tri(0.2, 0.1), small_tri(50, 0.1), small_tri(50.2, 0.1)],
                     np.float32)
    assert rt.Bvh(pairs, 4, 4, leaf_split_option=1).node_count == 1


def test_sah_factor_table(rt):
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    assert rt.load_sah_factor(4, 4, 4) == f32(0.9790158394920848)
    assert rt.load_sah_factor(12, 16, 4) == f32(0.8735301252614894)
    assert rt.load_sah_factor(5, 4, 4) == f32(0.6019200777645688)  # L=5 -> leaf memory 8
    assert rt.load_sah_factor(8, 8, 8) == 1.0  # the AVX table is not set up
    assert rt.load_sah_factor(20, 4, 4) == 1.0  # not in the table


def test_root_leaf_and_tiny_inputs(rt):
    one = np.array([small_tri(0, 1)], np.float32)
    b = rt.Bvh(one, 4, 4)
    assert b.node_count == 1 and b.is_leaf()[0]
    assert np.isnan(b.tri_points[9 * 4 - 1]) == False  # noqa: E712  (lRS == Lm: no marker)
    b8 = rt.Bvh(one, 8, 8)
    assert np.isnan(b8.tri_points[9 * 8 - 1]) and b8.tri_points[9 * 8 - 2] == 4
    with pytest.raises(rt.RtError):
        rt.Bvh(np.zeros((0, 9), np.float32), 4, 4)
    with pytest.raises(rt.RtError):
        rt.Bvh(one, 4, 4, gang=3)


def test_memory_overrides(rt, orc):
    """Node / leaf memory above the minimum (the node and leaf memory tests,
    rayTracer.cpp:41-52,630-640): the same tree in records of M / Lm slots
    with the loop-count markers.  The closest t is the same; the visit order
    may change, because a node whose child count is below M iterates the
    zero padding of its order row first (the padding quirk), so node counts
    and ties can differ (the GPU parity tests cover these trees)."""
    tris = rt.interior_scene(9, 20_000, 12)
    rays = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 64, 48, 16)
    base = rt.Bvh(tris, 4, 4)
    ref = orc.Oracle.from_bvh(base).trace(rays)
    for nm, lm in ((8, 4), (4, 12), (16, 16)):
        b = rt.Bvh(tris, 4, 4, node_memory=nm, leaf_memory=lm)
        assert (b.node_memory, b.leaf_memory) == (nm, lm)
        assert np.array_equal(b.base[~b.is_leaf()], base.base[~base.is_leaf()])
        assert len(b.tri_points) == len(base.tri_points) // base.leaf_memory * lm
        r = orc.Oracle.from_bvh(b).trace(rays)
        assert np.array_equal(r["hits"]["slot"] >= 0, ref["hits"]["slot"] >= 0)
        assert np.array_equal(r["hits"]["t"], ref["hits"]["t"]), (nm, lm)
        if nm == base.node_memory:
            # same visit order: the same leaves in the same order, so the same
            # hits and counts, at the larger leaf block stride
            assert np.array_equal(r["hits"]["leaf"] // (9 * lm), ref["hits"]["leaf"] // (9 * base.leaf_memory))
            for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
                assert np.array_equal(r["counts"][f], ref["counts"][f]), (nm, lm, f)
    for bad in (dict(node_memory=6), dict(leaf_memory=2), dict(leaf_memory=20)):
        with pytest.raises(rt.RtError):
            rt.Bvh(tris, 4, 4, **bad)
