"""GPU parity: librt.so traversal vs the oracle on the same host-built BVH.

Bar: hit id / leaf / slot / miss flag / node and leaf counts bit-exact; t,
b1 and b2 are bit-exact too (same FP32 operation sequence), which is
stricter than the 1e-5 relative tolerance the path allows.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1(rt):
    tris = rt.spheres_scene(0)
    bvh = rt.Bvh(tris, 4, 4)
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 256, 256, 16)
    return tris, bvh, rays


def test_config1_closest(rt, cfg1):
    tris, bvh, rays = cfg1
    print(f"config1 N4L4 closest: {len(rays)} rays", flush=True)
    scene = rt.Scene(bvh)
    hits, cnt = scene.intersect(rays, counts=True)
    ref = oracle_for(bvh).trace(rays)
    assert (ref["diag"] & oracle.DIAG_BOGUS).sum() == 0
    assert_closest_equal(hits, ref["hits"], bvh, cnt, ref["counts"], "cfg1")
    assert (hits["slot"] >= 0).sum() > 10000


def test_config1_shadow_anyhit(rt, cfg1):
    tris, bvh, rays = cfg1
    scene = rt.Scene(bvh)
    hits = scene.intersect(rays)
    nrm, pos = bvh.surface(rays, hits)
    srays, _ = rt.spawn_shadow(nrm, pos, hits, (0.0, 30.0, 0.0), 1.0)
    print(f"config1 shadow: {len(srays)} rays", flush=True)
    occ, cnt = scene.intersect(srays, mode="any", counts=True)
    ref = oracle_for(bvh).trace(srays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ref["occluded"])
    for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
        assert np.array_equal(cnt[f], ref["counts"][f]), f
    assert 0 < occ.sum() < len(srays)


def test_config1_save_distance(rt, cfg1):
    tris, bvh, rays = cfg1
    scene = rt.Scene(bvh)
    hits, cnt = scene.intersect(rays, mode="save_distance", counts=True)
    ref = oracle_for(bvh).trace(rays, mode=oracle.ORC_SAVE_DIST)
    assert_closest_equal(hits, ref["hits"], bvh, cnt, ref["counts"], "cfg1 save-distance")


@pytest.mark.parametrize("N", [4, 8, 12, 16])
@pytest.mark.parametrize("L", [4, 8, 12, 16])
def test_interior_sweep(rt, N, L):
    tris = rt.interior_scene(7, 60_000, 30)
    bvh = rt.Bvh(tris, N, L)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 128, 96, 16)
    scene = rt.Scene(bvh)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    diff = rt.spawn_diffuse(prim, nrm, pos, hits, 20000, seed=3)
    print(f"interior N{N}L{L}: {len(prim)} primary + {len(diff)} diffuse", flush=True)
    rays = np.concatenate([prim, diff])
    g, gc = scene.intersect(rays, counts=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"N{N}L{L}")
    srays, _ = rt.spawn_shadow(nrm, pos, hits, (0.0, 9.5, 0.0), 1.0)
    occ, oc = scene.intersect(srays, mode="any", counts=True)
    ro = oracle_for(bvh).trace(srays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"])
    assert np.array_equal(oc["leaves"], ro["counts"]["leaves"])
    assert np.array_equal(oc["nodes"], ro["counts"]["nodes"])


@pytest.mark.parametrize("gang,N,L", [(8, 8, 8), (8, 16, 16), (16, 16, 16), (4, 5, 3)])
def test_gang_sizes(rt, gang, N, L):
    tris = rt.interior_scene(11, 30_000, 20)
    bvh = rt.Bvh(tris, N, L, gang=gang)
    rays = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 128, 64, 16)
    print(f"gang {gang} N{N}L{L}", flush=True)
    g, gc = rt.Scene(bvh).intersect(rays, counts=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"G{gang}")


def test_edge_cases(rt):
    # a single triangle (root is a leaf), rays that miss, a ray through the
    # origin (the zero-box padding quirk), t == tMax, zero direction parts
    tri = np.array([[-1, -1, 5, 1, -1, 5, 0, 1, 5]], np.float32)
    bvh = rt.Bvh(tri, 4, 4)
    assert np.isnan(bvh.bounds[0, 0])
    rays = rt.make_rays([[0, 0, 0], [0, 0, 0], [5, 5, 0], [0, 0, 0]],
                        [[0, 0, 1], [0, 0, -1], [5, 5, 1], [0, 0, 5]], tmax=222222.0)
    rays["tmax"][3] = 1.0  # hit exactly at t == tMax (inclusive)
    print("edge cases: root leaf", flush=True)
    scene = rt.Scene(bvh)
    g, gc = scene.intersect(rays, counts=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "root-leaf")
    assert list(g["slot"]) == [0, -1, -1, 0]
    # scene with triangles around the origin: rays from the origin hit the
    # zero boxes of padded slots
    tris = rt.interior_scene(3, 2000, 4)
    tris = tris - np.tile(np.float32([0, 5, 0]), 3)[None, :]
    bvh = rt.Bvh(tris, 6, 4)
    dirs = np.random.default_rng(0).normal(size=(4096, 3)).astype(np.float32)
    rays = rt.make_rays(np.zeros((4096, 3), np.float32), dirs)
    scene = rt.Scene(bvh)
    for mode, om in (("closest", oracle.ORC_CLOSEST), ("save_distance", oracle.ORC_SAVE_DIST)):
        g, gc = scene.intersect(rays, mode=mode, counts=True)
        ref = oracle_for(bvh).trace(rays, mode=om)
        assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"origin rays {mode}")
    occ, oc = scene.intersect(rays, mode="any", counts=True)
    ro = oracle_for(bvh).trace(rays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"])
    assert np.array_equal(oc["nodes"], ro["counts"]["nodes"])


@pytest.mark.parametrize("N,L", [(4, 4), (8, 8), (16, 16), (12, 4)])
def test_stream_scheduling_matches_per_ray(rt, N, L):
    """RT_STREAM (node phase / leaf phase over a block's ray slots) keeps each
    ray's own depth-first order: same hits and counts as the oracle."""
    tris = rt.interior_scene(17, 50_000, 25)
    bvh = rt.Bvh(tris, N, L)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 128, 96, 16)
    scene = rt.Scene(bvh)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    rays = np.concatenate([prim, rt.spawn_diffuse(prim, nrm, pos, hits, 30000, seed=9)])
    print(f"stream N{N}L{L}: {len(rays)} rays", flush=True)
    g, gc = scene.intersect(rays, counts=True, stream=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"stream N{N}L{L}")
    srays, _ = rt.spawn_shadow(nrm, pos, hits, (0.0, 9.5, 0.0), 1.0)
    occ, oc = scene.intersect(srays, mode="any", counts=True, stream=True)
    ro = oracle_for(bvh).trace(srays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"])
    assert np.array_equal(oc["nodes"], ro["counts"]["nodes"])
    assert np.array_equal(oc["leaves"], ro["counts"]["leaves"])


def test_stream_edge_cases(rt):
    tri = np.array([[-1, -1, 5, 1, -1, 5, 0, 1, 5]], np.float32)
    bvh = rt.Bvh(tri, 4, 4)  # root is a leaf
    rays = rt.make_rays([[0, 0, 0], [0, 0, 0], [5, 5, 0]], [[0, 0, 1], [0, 0, -1], [5, 5, 1]])
    g, gc = rt.Scene(bvh).intersect(rays, counts=True, stream=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "stream root-leaf")
    # fewer rays than one block of slots, and origin rays (stale zero boxes)
    tris = rt.interior_scene(3, 2000, 4) - np.tile(np.float32([0, 5, 0]), 3)[None, :]
    bvh = rt.Bvh(tris, 6, 4)
    dirs = np.random.default_rng(0).normal(size=(100, 3)).astype(np.float32)
    rays = rt.make_rays(np.zeros((100, 3), np.float32), dirs)
    scene = rt.Scene(bvh)
    g, gc = scene.intersect(rays, counts=True, stream=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "stream origin rays")
    occ, oc = scene.intersect(rays, mode="any", counts=True, stream=True)
    ro = oracle_for(bvh).trace(rays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"]) and np.array_equal(oc["nodes"], ro["counts"]["nodes"])


@pytest.mark.parametrize("N,L,G", [(4, 4, 4), (8, 8, 4), (16, 16, 4), (4, 16, 4), (12, 8, 4), (8, 8, 8)])
def test_cooperative_matches_oracle(rt, N, L, G):
    """RT_COOP (W lanes per ray, ballot + prefix-popcount pushes, group
    reduction for the gang-G tie rule): same hits and counts as the oracle."""
    tris = rt.interior_scene(19, 50_000, 25)
    bvh = rt.Bvh(tris, N, L, gang=G)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 128, 96, 16)
    scene = rt.Scene(bvh)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    rays = np.concatenate([prim, rt.spawn_diffuse(prim, nrm, pos, hits, 30000, seed=11)])
    print(f"coop N{N}L{L}G{G}: {len(rays)} rays", flush=True)
    g, gc = scene.intersect(rays, counts=True, coop=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"coop N{N}L{L}")
    srays, _ = rt.spawn_shadow(nrm, pos, hits, (0.0, 9.5, 0.0), 1.0)
    occ, oc = scene.intersect(srays, mode="any", counts=True, coop=True)
    ro = oracle_for(bvh).trace(srays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"])
    assert np.array_equal(oc["nodes"], ro["counts"]["nodes"])
    assert np.array_equal(oc["leaves"], ro["counts"]["leaves"])


def test_cooperative_edge_cases(rt):
    tri = np.array([[-1, -1, 5, 1, -1, 5, 0, 1, 5]], np.float32)
    bvh = rt.Bvh(tri, 4, 4)
    rays = rt.make_rays([[0, 0, 0], [0, 0, 0], [5, 5, 0]], [[0, 0, 1], [0, 0, -1], [5, 5, 1]])
    g, gc = rt.Scene(bvh).intersect(rays, counts=True, coop=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "coop root-leaf")
    tris = rt.interior_scene(3, 2000, 4) - np.tile(np.float32([0, 5, 0]), 3)[None, :]
    bvh = rt.Bvh(tris, 6, 4)
    dirs = np.random.default_rng(0).normal(size=(3000, 3)).astype(np.float32)
    rays = rt.make_rays(np.zeros((3000, 3), np.float32), dirs)
    scene = rt.Scene(bvh)
    g, gc = scene.intersect(rays, counts=True, coop=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "coop origin rays")
    occ, oc = scene.intersect(rays, mode="any", counts=True, coop=True)
    ro = oracle_for(bvh).trace(rays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"]) and np.array_equal(oc["nodes"], ro["counts"]["nodes"])
    # exact ties inside a leaf resolved by the gang rule: duplicated triangles
    base = rt.interior_scene(4, 500, 2)
    dup = np.concatenate([base, base, base])
    bvh = rt.Bvh(dup, 8, 8)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 64, 64, 16)
    g, gc = rt.Scene(bvh).intersect(prim, counts=True, coop=True)
    ref = oracle_for(bvh).trace(prim)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "coop ties")


def _special_rays(prim, seed):
    """Ordinary primary rays mixed with rays whose slab operands can produce
    NaNs: infinite inverse direction (denormal or -0 direction part), NaN
    tMax, an origin coordinate beyond 2^100, and an infinite direction part
    (inverse 0), for which the all-zero padding triangles of a leaf have a
    NaN divisor and pass the test."""
    rng = np.random.default_rng(seed)
    sp = prim[rng.choice(len(prim), 4000)].copy()
    k = np.arange(len(sp)) % 6
    sp["dx"][k == 0] = np.float32(1e-40)
    sp["ix"][k == 0] = np.float32(np.inf)
    sp["tmax"][k == 1] = np.float32(np.nan)
    sp["ox"][k == 2] = np.float32(3e30)
    sp["dy"][k == 3] = np.float32(-0.0)
    sp["iy"][k == 3] = np.float32(-np.inf)
    sp["dz"][k == 5] = np.float32(np.inf)
    sp["iz"][k == 5] = np.float32(0.0)
    rays = np.concatenate([prim, sp])
    return rays[rng.permutation(len(rays))]


@pytest.mark.parametrize("scene_kind", ["regular", "far", "inverted"])
def test_slab_special_operands(rt, scene_kind):
    """The min/max slab form is taken only where no NaN can arise; the rays
    above (and, in a scene with a box coordinate beyond 2^100, every ray)
    take the maxps/minps select sequence.  Both kinds share waves.  A box
    with min > max on an axis tests like the ordered box in both forms (the
    reference swaps per axis)."""
    tris = rt.interior_scene(5, 20000, 10)
    if scene_kind == "far":
        tris = np.concatenate([tris, np.float32([[3e30, 0, 0, 3e30, 1, 0, 3e30, 0, 1]])])
    bvh = rt.Bvh(tris, 8, 8)
    if scene_kind == "inverted":
        # swap min and max x of child slot 0 of the root (the reference's
        # slab test swaps per axis, so its interval is unchanged)
        k = 8
        b = bvh.bounds
        assert not np.isnan(b[0, 0])
        b[0, 0], b[0, 3 * k] = b[0, 3 * k].copy(), b[0, 0].copy()
        assert b[0, 0] > b[0, 3 * k]
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 96, 64, 16)
    rays = _special_rays(prim, 5)
    scene = rt.Scene(bvh)
    # a box beyond 2^60 or a child box outside its parent's keeps the scene
    # on the exact-box records; the regular scene runs the quantized ones,
    # where these rays take the path that pushes every child tentatively
    assert (scene.stats()["quantized_record_bytes"] > 0) == (scene_kind == "regular")
    print(f"special slab operands {scene_kind}: {len(rays)} rays", flush=True)
    for kw in ({}, {"stream": True}, {"coop": True}):
        g, gc = scene.intersect(rays, counts=True, **kw)
        ref = oracle_for(bvh).trace(rays)
        assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"special {kw}")
        occ, oc = scene.intersect(rays, mode="any", counts=True, **kw)
        ro = oracle_for(bvh).trace(rays, mode=oracle.ORC_ANY)
        assert np.array_equal(occ, ro["occluded"]), kw
        assert np.array_equal(oc["nodes"], ro["counts"]["nodes"]), kw
    g, gc = scene.intersect(rays, mode="save_distance", counts=True)
    ref = oracle_for(bvh).trace(rays, mode=oracle.ORC_SAVE_DIST)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], "special save-distance")


@pytest.mark.parametrize("sched", [4, 16, 64])
@pytest.mark.parametrize("N,L", [(8, 8), (4, 16)])
def test_wave_schedules(rt, sched, N, L):
    """One phase per iteration (rt_set_schedule): lanes wait, rays keep their
    own step order, so hits and counts equal the oracle's."""
    tris = rt.interior_scene(23, 40_000, 20)
    bvh = rt.Bvh(tris, N, L)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 96, 64, 16)
    scene = rt.Scene(bvh)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    rays = np.concatenate([prim, rt.spawn_diffuse(prim, nrm, pos, hits, 20000, seed=13)])
    scene.set_schedule(sched)
    print(f"schedule {sched} N{N}L{L}: {len(rays)} rays", flush=True)
    g, gc = scene.intersect(rays, counts=True)
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"sched {sched}")
    srays, _ = rt.spawn_shadow(nrm, pos, hits, (0.0, 9.5, 0.0), 1.0)
    occ, oc = scene.intersect(srays, mode="any", counts=True)
    ro = oracle_for(bvh).trace(srays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"])
    assert np.array_equal(oc["nodes"], ro["counts"]["nodes"])
    assert np.array_equal(oc["leaves"], ro["counts"]["leaves"])
    g, gc = scene.intersect(rays, mode="save_distance", counts=True)
    ref = oracle_for(bvh).trace(rays, mode=oracle.ORC_SAVE_DIST)
    assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"sched {sched} save-distance")


@pytest.mark.parametrize("N,L,nm,lm", [(4, 4, 8, 4), (4, 4, 16, 16), (8, 4, 12, 8), (5, 3, 12, 12)])
def test_memory_test_layouts(rt, N, L, nm, lm):
    """Trees flattened with node / leaf memory above the minimum (the memory
    tests of rayTracer.cpp:41-52,630-640): the padded order rows change the
    visit order (padding quirk), and the GPU follows the oracle bit for bit
    in all three traversals."""
    tris = rt.interior_scene(13, 40_000, 25)
    bvh = rt.Bvh(tris, N, L, node_memory=nm, leaf_memory=lm)
    assert (bvh.node_memory, bvh.leaf_memory) == (nm, lm)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 96, 64, 16)
    scene = rt.Scene(bvh)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    rays = np.concatenate([prim, rt.spawn_diffuse(prim, nrm, pos, hits, 8000, seed=4)])
    print(f"memory test N{N}L{L} M{nm} Lm{lm}: {len(rays)} rays", flush=True)
    o = oracle_for(bvh)
    for mode, omode in (("closest", oracle.ORC_CLOSEST), ("save_distance", oracle.ORC_SAVE_DIST)):
        g, gc = scene.intersect(rays, mode=mode, counts=True)
        ref = o.trace(rays, mode=omode)
        assert_closest_equal(g, ref["hits"], bvh, gc, ref["counts"], f"{mode} M{nm} Lm{lm}")
    srays, _ = rt.spawn_shadow(nrm, pos, hits, (0.0, 9.5, 0.0), 1.0)
    occ, oc = scene.intersect(srays, mode="any", counts=True)
    ro = o.trace(srays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"])
    for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
        assert np.array_equal(oc[f], ro["counts"][f]), f
    scene.close()
