"""Known answers for the oracle's restatement of the ISPC kernels
(src/ISPC/rayTracerTemplate.ispc): slab test, Moeller-Trumbore, gang-G tie
rule, loop-count decoding.  Every expected value is derived by hand from the
geometry (exact binary fractions), not recomputed with the kernel formula.
"""
import numpy as np
import pytest

import oracle

FLT_MAX = np.float32(np.finfo(np.float32).max)


def ray(o, d, tmax=222222.0):
    """FastRay layout with inv = 1/dir and FLT_MAX for a zero component."""
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    with np.errstate(divide="ignore", over="ignore"):
        inv = np.where(d == 0, FLT_MAX, np.float32(1) / d).astype(np.float32)
    return np.array([tmax, *o, *inv, *d], np.float32)


def boxes(bmin, bmax):
    """SoA planes of k boxes: minX[k] minY[k] minZ[k] maxX[k] maxY[k] maxZ[k]."""
    bmin = np.asarray(bmin, np.float32).reshape(-1, 3)
    bmax = np.asarray(bmax, np.float32).reshape(-1, 3)
    return np.concatenate([bmin[:, 0], bmin[:, 1], bmin[:, 2], bmax[:, 0], bmax[:, 1], bmax[:, 2]]).astype(np.float32)


def slab(b, k, r, fill=7.0):
    dist = np.full(k, fill, np.float32)
    hit = oracle.aabb_intersect(b, k, dist, r)
    return hit, dist


UNIT = ([0, 0, 0], [1, 1, 1])


def pad4(bmin, bmax):
    bmin = list(bmin) + [[0, 0, 0]] * (4 - len(bmin))
    bmax = list(bmax) + [[0, 0, 0]] * (4 - len(bmax))
    return boxes(bmin, bmax)


# ---------------------------------------------------------------- slab test

def test_slab_entry_distance_axis_aligned():
    # ray from x=-4 along +x with |dir| = 8: the unit box is entered at
    # x=0 -> t = 4/8 = 0.5 and left at x=1 -> t = 0.625
    r = ray([-4, 0.5, 0.5], [8, 0, 0])
    hit, dist = slab(pad4([UNIT[0]], [UNIT[1]]), 4, r)
    assert hit == 1
    assert dist[0] == np.float32(0.5)


def test_slab_misses_leave_distance_untouched():
    r = ray([-4, 0.5, 0.5], [8, 0, 0])
    b = pad4([[0, 0, 0], [0, 2, 0], [-3, 0, 0]], [[1, 1, 1], [1, 3, 1], [-2, 1, 1]])
    hit, dist = slab(b, 4, r)
    # box 1 is above the ray -> miss; box 2 spans x in [-3,-2]: t in [1/8, 2/8]
    assert hit == 1
    assert dist[1] == np.float32(7.0)
    assert dist[2] == np.float32(0.125)
    # padded zero box [0,0]^3: the ray passes y=0.5, not the origin -> miss
    assert dist[3] == np.float32(7.0)


def test_slab_box_behind_and_tmax_clip():
    r = ray([4, 0.5, 0.5], [8, 0, 0])  # box is behind the origin
    hit, _ = slab(pad4([UNIT[0]], [UNIT[1]]), 4, r)
    assert hit == 0
    r = ray([-4, 0.5, 0.5], [8, 0, 0], tmax=0.4)  # entry at 0.5 > tMax
    assert slab(pad4([UNIT[0]], [UNIT[1]]), 4, r)[0] == 0
    r = ray([-4, 0.5, 0.5], [8, 0, 0], tmax=0.5)  # entry == tMax: inclusive
    hit, dist = slab(pad4([UNIT[0]], [UNIT[1]]), 4, r)
    assert hit == 1 and dist[0] == np.float32(0.5)


def test_slab_corner_touch_is_inclusive():
    # diagonal ray through the corner (1,1) of box [0,1]x[1,2]: the x slab
    # ends at t = 5/8 exactly where the y slab begins -> t0 == t1 -> hit
    r = ray([-4, -4, 0.5], [8, 8, 0])
    hit, dist = slab(pad4([[0, 1, 0]], [[1, 2, 1]]), 4, r)
    assert hit == 1 and dist[0] == np.float32(0.625)


def test_slab_zero_direction_uses_flt_max():
    # dir.y == 0 -> inv.y = FLT_MAX: a ray grazing the top face y=1 from
    # outside has tFar.y = 0 and misses; from inside the box it hits at 0
    outside = ray([-4, 1, 0.5], [8, 0, 0])
    assert slab(pad4([UNIT[0]], [UNIT[1]]), 4, outside)[0] == 0
    inside = ray([0.5, 1, 0.5], [8, 0, 0])
    hit, dist = slab(pad4([UNIT[0]], [UNIT[1]]), 4, inside)
    assert hit == 1 and dist[0] == 0.0


def test_slab_zero_box_hit_from_origin():
    # the padded zero boxes (fastNodeManager.h:1545-1547) are hit by rays
    # starting at the origin: all slabs give t = 0
    r = ray([0, 0, 0], [1, 2, 3])
    b = pad4([[5, 5, 5]], [[6, 6, 6]])
    hit, dist = slab(b, 4, r)
    assert hit == 1
    assert list(dist[1:]) == [0.0, 0.0, 0.0]


def test_slab_nan_keeps_previous_bounds():
    # a denormal direction component has reciprocal +inf; (b - o) * inf is
    # NaN where b == o.  maxps/minps keep the second operand on NaN, so the
    # y slab is skipped: entry at t = 0.5 as for the x slab alone.
    r = ray([-4, 0.0, 0.5], [8, 1e-40, 0])
    assert np.isinf(r[5])
    b = pad4([[0, 0, 0]], [[1, 1, 1]])
    hit, dist = slab(b, 4, r)
    # y: tN = (0-0)*inf = NaN, tF = (1-0)*inf = inf -> no swap, t0 stays 0.5
    assert hit == 1 and dist[0] == np.float32(0.5)


# ---------------------------------------------------------------- triangles

def leaf(tris, k):
    """SoA leaf block of k slots (stride k): p0.x[k] p0.y[k] ... p2.z[k]."""
    t = np.zeros((k, 9), np.float32)
    t[: len(tris)] = np.asarray(tris, np.float32).reshape(-1, 9)
    return np.ascontiguousarray(t.T.reshape(-1))


def tri_hit(tris, k, r, gang=4):
    r = r.copy()
    res = oracle.tri_intersect(leaf(tris, k), 0, k, r, gang)
    return res, r[0]


RIGHT = [0, 0, 0, 1, 0, 0, 0, 1, 0]  # right triangle in the z=0 plane


def test_mt_interior_point():
    # (0.25, 0.25) on the z=0 triangle from z=1 straight down: t = 1, and the
    # barycentrics of that point are (b1, b2) = (0.25, 0.25)
    r = ray([0.25, 0.25, 1], [0, 0, -1])
    res, t = tri_hit([RIGHT], 4, r)
    assert res == 0 and t == 1.0
    ok, t64, b1, b2 = oracle.mt_f64(r, np.array(RIGHT, np.float32))
    assert ok and t64 == 1.0 and b1 == 0.25 and b2 == 0.25


def test_mt_centroid_barycentrics():
    tri = [0, 0, 0, 3, 0, 0, 0, 3, 0]
    r = ray([1, 1, 2], [0, 0, -1])
    ok, t, b1, b2 = oracle.mt_f64(r, np.array(tri, np.float32))
    assert ok and abs(b1 - 1 / 3) < 1e-15 and abs(b2 - 1 / 3) < 1e-15 and t == 2.0
    res, t32 = tri_hit([tri], 4, r)
    assert res == 0 and t32 == 2.0


@pytest.mark.parametrize("p", [(0, 0), (1, 0), (0, 1), (0.5, 0.5), (0.5, 0), (0, 0.5)])
def test_mt_vertices_and_edges_are_inclusive(p):
    r = ray([p[0], p[1], 1], [0, 0, -1])
    res, t = tri_hit([RIGHT], 4, r)
    assert res == 0 and t == 1.0


@pytest.mark.parametrize("p", [(-0.001, 0.5), (0.5, -0.001), (0.51, 0.5), (1.001, 0)])
def test_mt_outside_points_miss(p):
    r = ray([p[0], p[1], 1], [0, 0, -1])
    assert tri_hit([RIGHT], 4, r)[0] == -1


def test_mt_tmax_inclusive_and_behind():
    assert tri_hit([RIGHT], 4, ray([0.25, 0.25, 1], [0, 0, -1], tmax=1.0))[0] == 0
    assert tri_hit([RIGHT], 4, ray([0.25, 0.25, 1], [0, 0, -1], tmax=0.999))[0] == -1
    assert tri_hit([RIGHT], 4, ray([0.25, 0.25, 1], [0, 0, 1]))[0] == -1  # t = -1


def test_mt_parallel_and_degenerate_rejected():
    assert tri_hit([RIGHT], 4, ray([0.25, 0.25, 1], [1, 0, 0]))[0] == -1  # divisor 0
    line = [0, 0, 0, 1, 1, 0, 2, 2, 0]
    assert tri_hit([line], 4, ray([0.5, 0.5, 1], [0, 0, -1]))[0] == -1
    # padded slots are all-zero triangles: never hit, even through the origin
    assert tri_hit([], 4, ray([0, 0, 1], [0, 0, -1]))[0] == -1


def test_mt_closest_of_several_and_tmax_update():
    near = [0, 0, 0.5, 1, 0, 0.5, 0, 1, 0.5]  # z = 0.5 -> t = 0.5
    far = RIGHT  # z = 0 -> t = 1
    res, t = tri_hit([far, near, far], 4, ray([0.25, 0.25, 1], [0, 0, -1]))
    assert res == 1 and t == 0.5


def shifted(z):
    return [0, 0, z, 1, 0, z, 0, 1, z]


@pytest.mark.parametrize("slots,gang,expect", [
    ((1, 5), 4, 1),     # same program instance (1 % 4): first strict minimum kept
    ((1, 2), 4, 2),     # different instances: reduce_max of the kept slots
    ((2, 5), 4, 5),
    ((0, 4, 1), 4, 1),  # instance 0 keeps slot 0, instance 1 slot 1 -> 1
    ((2, 6), 4, 2),     # gang 4: slots 2 and 6 share instance 2
    ((2, 6), 8, 6),     # gang 8: different instances
    ((3, 7, 11, 15), 4, 3),
    ((3, 7, 11, 15), 8, 7),
])
def test_leaf_tie_rule(slots, gang, expect):
    k = 16
    tris = [shifted(5.0)] * k  # behind: never hit
    for s in slots:
        tris[s] = RIGHT
    res, t = tri_hit(tris, k, ray([0.25, 0.25, 1], [0, 0, -1]), gang=gang)
    assert res == expect and t == 1.0


def test_leaf_hits_beyond_1e6_are_not_recorded():
    # tStore starts at 1e6 (rayTracerTemplate.ispc:29): a hit at t = 2e6 is
    # found by triAnyHit but not by triIntersect
    r = ray([0.25, 0.25, 2e6], [0, 0, -1], tmax=3e6)
    res, t = tri_hit([RIGHT], 4, r)
    assert res == -1 and t == np.float32(3e6)
    assert oracle.tri_anyhit(leaf([RIGHT], 4), 0, 4, r) == 1


def test_anyhit_accepts_nan_tests():
    # a tiny triangle at the ray origin: divisor 1e-40 is denormal, its
    # reciprocal overflows to inf and b1, b2, t are 0*inf = NaN.  The ISPC
    # tests `b1 < 0 || b1 > 1` etc. are false for NaN, so `hit` stays true:
    # triAnyHit reports a hit, triIntersect drops it (tStore > NaN is false).
    tri = [0, 0, 0, 1e-20, 0, 0, 0, 1e-20, 0]
    r = ray([0, 0, 0], [0, 0, -1])
    assert oracle.tri_anyhit(leaf([tri], 4), 0, 4, r) == 1
    assert tri_hit([tri], 4, r)[0] == -1


def test_leaf_offset_and_stride():
    # triId is a float offset into trianglePoints; planes have stride k
    blk = np.concatenate([leaf([shifted(5.0)], 4), leaf([shifted(5.0), RIGHT], 8)])
    r = ray([0.25, 0.25, 1], [0, 0, -1])
    assert oracle.tri_intersect(blk, 36, 8, r.copy(), 4) == 1
    assert oracle.tri_intersect(blk, 0, 4, r.copy(), 4) == -1


def test_dispatch_rule():
    # callIspcTemplateNotConst: 4, 8, 12 pick their variant, anything else 16
    assert [oracle.lib().orc_dispatch_k(k) for k in (4, 8, 12, 16, 0, 5, 20)] == [4, 8, 12, 16, 16, 16, 16]
