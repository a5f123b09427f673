"""The SIMD CPU baseline (oracle/cpu_simd.c: the oracle's traversals with
4-wide SSE node and leaf tests, the shape of the reference's sse4-i32x4 ISPC
kernels) equals the scalar oracle bit for bit: hit identity, t, barycentrics,
occlusion and the four counts, in every mode, per ray and in the stream
traversal, across node and leaf sizes and on the hand-built trees whose
visit order is derived by hand (padding quirk, stale zero-box distance,
NaN-marker loop counts).  No GPU."""
import numpy as np
import pytest

import oracle
from tests import handbvh

MODES = (oracle.ORC_CLOSEST, oracle.ORC_ANY, oracle.ORC_SAVE_DIST)


def _same(a, b):
    for key in ("hits", "occluded", "counts"):
        x, y = a[key], b[key]
        if x is None:
            assert y is None
            continue
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), key


def _pair(arr):
    args = (arr["node_memory"], arr["leaf_memory"], arr["gang"], arr["bounds"], arr["base"], arr["order"],
            arr["child_type"], arr["tri_points"])
    return oracle.Oracle(*args), oracle.Oracle(*args, simd=True)


@pytest.fixture(scope="module")
def scene(rt):
    tris = rt.interior_scene(7, 20_000, 12)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 96, 64, 16)
    g = np.random.default_rng(3)
    o = g.uniform([-14, 0.5, -14], [14, 9.5, 14], size=(3000, 3)).astype(np.float32)
    t = g.uniform([-14, 0.5, -14], [14, 9.5, 14], size=(3000, 3)).astype(np.float32)
    rand = rt.make_rays(o, t)
    # axis-parallel rays (zero direction components: FLT_MAX reciprocals),
    # rays through the origin (the padded zero boxes) and short shadow rays
    axis = rt.make_rays(o[:300], o[:300] + np.float32([[5, 0, 0], [0, -3, 0], [0, 0, 7]] * 100))
    orig = rt.make_rays(np.zeros((200, 3), np.float32), t[:200])
    short = rt.make_rays(o[:300], t[:300], tmax=0.3)
    return tris, np.concatenate([prim, rand, axis, orig, short])


@pytest.mark.parametrize("N,L,G", [(4, 4, 4), (8, 8, 4), (16, 16, 4), (12, 4, 4), (5, 3, 4), (8, 8, 8)])
def test_simd_equals_scalar(rt, scene, N, L, G):
    tris, rays = scene
    bvh = rt.Bvh(tris, N, L, gang=G)
    scalar, simd = oracle.Oracle.from_bvh(bvh), oracle.Oracle.from_bvh(bvh, simd=True)
    for mode in MODES:
        _same(scalar.trace(rays, mode=mode), simd.trace(rays, mode=mode))
    for mode in (oracle.ORC_CLOSEST, oracle.ORC_ANY):
        _same(scalar.trace(rays, mode=mode, stream=True), simd.trace(rays, mode=mode, stream=True))


def test_simd_hand_built_trees():
    rays = np.concatenate([
        np.array([(222222.0, 0, 0.25, 0.25, 0.125, 3.4028235e38, 3.4028235e38, 8, 0, 0)], oracle.RAY_DTYPE),
        np.array([(222222.0, 10, 0.25, 0.25, -0.125, 3.4028235e38, 3.4028235e38, -8, 0, 0)], oracle.RAY_DTYPE),
        np.array([(222222.0, 0, 0, 0, 0.125, 3.4028235e38, 3.4028235e38, 8, 0, 0)], oracle.RAY_DTYPE),
        np.array([(222222.0, 0, 1.5, 1.5, 0.125, 3.4028235e38, 3.4028235e38, 8, 0, 0)], oracle.RAY_DTYPE),
        np.array([(222222.0, 12, 0.25, 0.25, -0.125, 3.4028235e38, 3.4028235e38, -8, 0, 0)], oracle.RAY_DTYPE),
    ])
    for arr in (handbvh.quirk_tree(False), handbvh.quirk_tree(True), handbvh.anyhit_tree(True),
                handbvh.anyhit_tree(False), handbvh.stale_tree(), handbvh.marker_tree()):
        scalar, simd = _pair(arr)
        for mode in MODES:
            _same(scalar.trace(rays, mode=mode), simd.trace(rays, mode=mode))


def test_simd_gang4_tie_rule():
    """Five copies of one triangle in an 8-slot leaf (marker_tree's leaf B):
    every lane meets the same t, lane 0 keeps slot 0 (its slot 4 is not
    strictly less), and reduce_max over the lanes at the minimum gives 3."""
    arr = handbvh.marker_tree()
    scalar, simd = _pair(arr)
    ray = np.array([(222222.0, 0, 1.5, 1.5, 0.125, 3.4028235e38, 3.4028235e38, 8, 0, 0)], oracle.RAY_DTYPE)
    for o in (scalar, simd):
        h = o.trace(ray)["hits"][0]
        assert (h["leaf"], h["slot"], h["t"]) == (72, 3, np.float32(1.1875))
