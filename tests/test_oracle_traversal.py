"""Traversal known answers and cross-checks for the oracle
(FastNodeManager::intersect / intersectSaveDistance / intersectSecondary /
intersectWideAlternative, fastNodeManager.h:835-1493).

Hand-built trees pin the visit order (expected leaf, hit identity and
node/leaf counts worked out by hand in each docstring of tests/handbvh.py);
brute force over all triangles pins the closest t; per-ray vs stream and
N/L invariance pin the scheduling-independent parts.
"""
import numpy as np
import pytest

import oracle
from tests import handbvh


def orc(arr):
    return oracle.Oracle(arr["node_memory"], arr["leaf_memory"], arr["gang"], arr["bounds"],
                         arr["base"], arr["order"], arr["child_type"], arr["tri_points"])


def fray(o, d, tmax=222222.0):
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    with np.errstate(divide="ignore"):
        inv = np.where(d == 0, np.float32(np.finfo(np.float32).max), np.float32(1) / d)
    return np.array([(tmax, *o, *inv.astype(np.float32), *d)], dtype=oracle.RAY_DTYPE)


PLUS_X = fray([0, 0.25, 0.25], [8, 0, 0])
MINUS_X = fray([10, 0.25, 0.25], [-8, 0, 0])


@pytest.mark.parametrize("full,ray,leaf,tri", [
    # c=3 < M=4, dir +x: the zero padding of the order row is iterated first
    # and pushes child 0 first, so child 0 is visited last and wins the tie
    (False, PLUS_X, 0, 10),
    # c=4 == M: plain backward iteration, child 0 is visited first and the
    # later-visited child 1 wins the tie (t <= tMax accepts equal t)
    (True, PLUS_X, 36, 11),
    (False, MINUS_X, 0, 10),
    (True, MINUS_X, 0, 10),
])
def test_padding_quirk_and_last_leaf_wins_ties(full, ray, leaf, tri):
    arr = handbvh.quirk_tree(full)
    r = orc(arr).trace(ray)
    h, c = r["hits"][0], r["counts"][0]
    assert h["slot"] == 0 and h["leaf"] == leaf and h["t"] == np.float32(0.625)
    assert arr["tri_ids"][h["leaf"] // 36 * 4 + h["slot"]] == tri
    assert c["nodes"] == 1 and c["leaves"] == 2
    assert c["node_slots"] == 4 and c["leaf_slots"] == 8


def test_anyhit_pushes_in_slot_order():
    arr = handbvh.anyhit_tree(second_hit=False)
    o = orc(arr)
    # closest: forward over the x row [1, 0, 2]: slot 0 (x=7, miss) is popped
    # first, then slot 1 (x=5) hits at t = 5/8
    r = o.trace(MINUS_X)
    assert r["hits"]["leaf"][0] == 36 and r["hits"]["t"][0] == np.float32(0.625)
    assert r["counts"]["leaves"][0] == 2
    # any-hit: slot order 0, 1 -> slot 1 popped first and it occludes
    r = o.trace(MINUS_X, mode=oracle.ORC_ANY)
    assert r["occluded"][0] == 1
    assert r["counts"]["nodes"][0] == 1 and r["counts"]["leaves"][0] == 1


def test_save_distance_skips_farther_entries():
    arr = handbvh.anyhit_tree(second_hit=True)
    o = orc(arr)
    r = o.trace(MINUS_X)
    assert r["hits"]["leaf"][0] == 0 and r["hits"]["t"][0] == np.float32(0.375)
    assert r["counts"]["leaves"][0] == 2  # intersect never culls on pop
    r = o.trace(MINUS_X, mode=oracle.ORC_SAVE_DIST)
    assert r["hits"]["leaf"][0] == 0 and r["hits"]["t"][0] == np.float32(0.375)
    assert r["counts"]["leaves"][0] == 1  # entry t0 = 0.625 > tMax: skipped uncounted


def test_stale_zero_box_distance_pushes_unhit_child():
    arr = handbvh.stale_tree()
    o = orc(arr)
    origin = fray([0, 0, 0], [8, 0, 0])
    r = o.trace(origin)
    h, c = r["hits"][0], r["counts"][0]
    assert h["slot"] == 0 and h["leaf"] == 72 and h["t"] == np.float32(0.25)
    # root + A; A's slot 3 (never hit) is pushed by the root's stale slot-3
    # distance -> 3 leaves instead of 2
    assert c["nodes"] == 2 and c["leaves"] == 3
    off = fray([0, 0, 0.001], [8, 0, 0])
    c2 = o.trace(off)["counts"][0]
    assert c2["nodes"] == 2 and c2["leaves"] == 2
    # any-hit pushes slot order over all M slots: the stale slot 3 of the root
    # pushes FastNode childIdBegin+3 = A's first leaf directly
    r = o.trace(origin, mode=oracle.ORC_ANY)
    assert r["occluded"][0] == 1 and r["counts"]["nodes"][0] == 1 and r["counts"]["leaves"][0] == 1
    r = o.trace(off, mode=oracle.ORC_ANY)
    assert r["occluded"][0] == 1 and r["counts"]["nodes"][0] == 2 and r["counts"]["leaves"][0] == 1
    # save distance: the stale entry carries t0 = 0 and is visited after the
    # hit at 0.25; the x=3 leaf (0.375) is culled
    r = o.trace(origin, mode=oracle.ORC_SAVE_DIST)
    assert r["hits"]["t"][0] == np.float32(0.25)
    assert r["counts"]["nodes"][0] == 2 and r["counts"]["leaves"][0] == 2


# marker_tree (M = 8, G = 4, Lm = 8): hand-derived results per ray and mode,
# (primIdBegin, slot, t, b1, b2, nodes, node_slots, leaves, leaf_slots).
# +x at y = z = 0.25: root k = 4 (marker) hits its three children; backward
# over [1, 0, 2, pad...] pushes node 1 first (padding), then B, then A, so A
# (k = 4, no hit), B (k = 8, five equal t: gang-4 rule gives slot 3, t =
# 1.1875) and node 1 (k = 8) follow.  Node 1 hits its six children (t0 =
# 0.375 ... 0.9375 <= 1.1875), pushes slot 0 first (padding), then 5..1:
# slot 1 hits at 0.5, slots 2-5 fail t <= tMax, slot 0 (visited last) hits
# at 0.375.  Leaves: A 4 + B 8 + five one-triangle leaves 4 each + the
# six-triangle leaf 8 = 40 slots.
MARKER_PLUS_X = (144, 0, 0.375, 0.25, 0.25, 2, 12, 8, 40)
# +x at y = z = 1.5: node 1's children boxes (y, z <= 1) all miss; B hits
# (y + z = 3 <= 4), b = 1.5 / 4.
MARKER_PLUS_X_HIGH = (72, 3, 1.1875, 0.375, 0.375, 2, 12, 2, 12)
# -x from x = 12: reverse, forward over the row: A, node 1, B pushed; B first
# (t = 0.3125), then node 1 tests its children with tMax 0.3125 (all miss),
# then A.  b = 0.25 / 4 on B's wall.
MARKER_MINUS_X = (72, 3, 0.3125, 0.0625, 0.0625, 2, 12, 2, 12)


@pytest.mark.parametrize("ray,want", [
    (fray([0, 0.25, 0.25], [8, 0, 0]), MARKER_PLUS_X),
    (fray([0, 1.5, 1.5], [8, 0, 0]), MARKER_PLUS_X_HIGH),
    (fray([12, 0.25, 0.25], [-8, 0, 0]), MARKER_MINUS_X),
])
def test_marker_tree_loop_counts_and_order(ray, want):
    """Node and leaf loop counts decoded from the NaN markers when M != G
    and Lm != G (getNodeLoopCount / getLeafLoopCount, fastNodeManager.h:
    1728-1763), and the visit order they feed."""
    arr = handbvh.marker_tree()
    r = orc(arr).trace(ray)
    h, c = r["hits"][0], r["counts"][0]
    got = (int(h["leaf"]), int(h["slot"]), float(h["t"]), float(h["b1"]), float(h["b2"]),
           int(c["nodes"]), int(c["node_slots"]), int(c["leaves"]), int(c["leaf_slots"]))
    assert got == want


def test_marker_tree_anyhit_and_save_distance():
    arr = handbvh.marker_tree()
    o = orc(arr)
    # any hit pushes slots 0, 1, 2 in slot order: B (slot 2) is popped first
    # and occludes: root k = 4, B k = 8
    for ray in (fray([0, 0.25, 0.25], [8, 0, 0]), fray([12, 0.25, 0.25], [-8, 0, 0])):
        r = o.trace(ray, mode=oracle.ORC_ANY)
        c = r["counts"][0]
        assert r["occluded"][0] == 1
        assert (c["nodes"], c["node_slots"], c["leaves"], c["leaf_slots"]) == (1, 4, 1, 8)
    # save distance, +x: A (0.125), B (1.1875, hit), node 1 (0.375); then
    # slot 1 (0.5, hit); slots 2-5 (0.625-0.9375 > 0.5) skipped uncounted;
    # slot 0 (0.375, hit)
    r = o.trace(fray([0, 0.25, 0.25], [8, 0, 0]), mode=oracle.ORC_SAVE_DIST)
    h, c = r["hits"][0], r["counts"][0]
    assert (h["leaf"], h["slot"], h["t"]) == (144, 0, np.float32(0.375))
    assert (c["nodes"], c["node_slots"], c["leaves"], c["leaf_slots"]) == (2, 12, 4, 20)


@pytest.fixture(scope="module")
def small_scene(rt):
    tris = rt.interior_scene(5, 3000, 6)
    rays = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 64, 48, 16)
    g = np.random.default_rng(1)
    o = g.uniform([-14, 0.5, -14], [14, 9.5, 14], size=(2000, 3)).astype(np.float32)
    t = g.uniform([-14, 0.5, -14], [14, 9.5, 14], size=(2000, 3)).astype(np.float32)
    return tris, np.concatenate([rays, rt.make_rays(o, t)])


@pytest.mark.parametrize("N,L", [(2, 1), (4, 4), (8, 8), (12, 4), (16, 16), (5, 3)])
def test_closest_t_matches_bruteforce(rt, small_scene, N, L):
    tris, rays = small_scene
    bvh = rt.Bvh(tris, N, L)
    res = oracle.Oracle.from_bvh(bvh).trace(rays)
    hits = res["hits"]
    assert (res["diag"] & (oracle.DIAG_BOGUS | oracle.DIAG_STACKCAP)).sum() == 0
    mism = 0
    for i in range(len(rays)):
        bi, bt = oracle.bruteforce_closest(tris, rays[i:i + 1].view(np.float32))
        if bi < 0:
            assert hits["slot"][i] == -1
            continue
        assert hits["slot"][i] >= 0, f"ray {i}: brute force hits {bi}, BVH misses"
        assert hits["t"][i] == np.float32(bt), f"ray {i}: t {hits['t'][i]} vs {bt}"
        tid = bvh.tri_ids[hits["leaf"][i] // (9 * bvh.leaf_memory) * bvh.leaf_memory + hits["slot"][i]]
        if tid != bi:
            # only legitimate for an exact tie in t
            ok, t2, _, _ = oracle.mt_f64(rays[i:i + 1].view(np.float32), tris[tid])
            mism += 1
            assert ok
    assert mism <= len(rays) // 100


def test_anyhit_matches_bruteforce(rt, small_scene):
    tris, rays = small_scene
    short = rays.copy()
    short["tmax"] = np.float32(0.5)
    bvh = rt.Bvh(tris, 8, 8)
    occ = oracle.Oracle.from_bvh(bvh).trace(short, mode=oracle.ORC_ANY)["occluded"]
    for i in range(0, len(short), 3):
        assert occ[i] == oracle.bruteforce_any(tris, short[i:i + 1].view(np.float32)), i


def test_t_is_independent_of_n_and_l(rt, small_scene):
    tris, rays = small_scene
    ts = []
    for N, L in ((4, 4), (8, 8), (16, 16), (6, 2)):
        bvh = rt.Bvh(tris, N, L)
        ts.append(oracle.Oracle.from_bvh(bvh).trace(rays)["hits"]["t"])
    for t in ts[1:]:
        assert np.array_equal(ts[0].view(np.uint32), t.view(np.uint32))


@pytest.mark.parametrize("N,L", [(4, 4), (8, 8), (16, 16)])
def test_stream_equals_per_ray(rt, small_scene, N, L):
    tris, rays = small_scene
    bvh = rt.Bvh(tris, N, L)
    o = oracle.Oracle.from_bvh(bvh)
    per = o.trace(rays)
    st = o.trace(rays, stream=True)
    for f in ("slot", "leaf", "t", "b1", "b2"):
        assert np.array_equal(per["hits"][f], st["hits"][f]), f
    for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
        assert np.array_equal(per["counts"][f], st["counts"][f]), f
    pa = o.trace(rays, mode=oracle.ORC_ANY)
    sa = o.trace(rays, mode=oracle.ORC_ANY, stream=True)
    assert np.array_equal(pa["occluded"], sa["occluded"])
    assert np.array_equal(pa["counts"]["leaves"], sa["counts"]["leaves"])


def test_reference_stream_axis_bug_keeps_closest_t(rt, small_scene):
    # intersectWideAlternative reads rayMajorAxis[i] (work-list position)
    # instead of [rayId] (fastNodeManager.h:913,939): the visit order changes,
    # the closest t does not
    # The camera rays of one tile share a major axis, so the bug shows only
    # on the random-direction rays (7 tiles of 256), whose work lists mix
    # axes once finished rays are compacted away.
    tris, rays = small_scene
    rays = rays[3072:3072 + 7 * 256]
    bvh = rt.Bvh(tris, 8, 8)
    o = oracle.Oracle.from_bvh(bvh)
    per = o.trace(rays)
    li, ti, cnt, tmax = o.wide(rays, flags=oracle.ORC_WIDE_AXIS_BUG | oracle.ORC_WIDE_SHARED_DIST)
    hit = per["hits"]["slot"] >= 0
    assert np.array_equal(hit, ti >= 0)
    assert np.array_equal(per["hits"]["t"][hit], tmax[hit])
    # the emulated bug changes some visit orders, hence node counts ...
    assert (cnt["nodes"] != per["counts"]["nodes"]).any()
    # ... and without it the stream traversal equals the per-ray one
    li0, ti0, cnt0, tmax0 = o.wide(rays, flags=0)
    assert np.array_equal(cnt0["nodes"], per["counts"]["nodes"])
    assert np.array_equal(cnt0["leaves"], per["counts"]["leaves"])


def test_stack_depth_within_reference_bound(rt):
    # the reference stack holds 48 entries (fastNodeManager.h:1333); the
    # benchmark-sized trees never need more
    tris = rt.interior_scene(2, 200_000)
    rays = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 128, 128, 16)
    for N, L in ((16, 4), (8, 8)):
        bvh = rt.Bvh(tris, N, L)
        r = oracle.Oracle.from_bvh(bvh).trace(rays)
        assert (r["diag"] & oracle.DIAG_STACK48).sum() == 0
