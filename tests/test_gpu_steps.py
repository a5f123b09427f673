"""rt_trace_steps (per-step workload of ray tiles) against the oracle's
visit sequences aggregated by oracle.tile_step_records: every record of
every tile identical, and the per-ray visit counts identical to the oracle's
counts (zero for rays not started)."""
import numpy as np
import pytest

import oracle
from tests.helpers import oracle_for

pytestmark = pytest.mark.gpu


def _check(scene, bvh, rays, tile, mode, step_cap=256):
    print(f"steps tile {tile} mode {mode} n {len(rays)} cap {step_cap}", flush=True)
    recs, cnt = scene.workgroup_steps(rays, tile, mode=mode, step_cap=step_cap)
    omode = oracle.ORC_ANY if mode == "any" else oracle.ORC_CLOSEST
    o = oracle_for(bvh)
    ids, lens = o.visit_sequences(rays, omode)
    started = ~np.isnan(rays["ox"]) if mode == "any" else np.ones(len(rays), bool)
    want = oracle.tile_step_records(ids, lens, tile, started)
    assert len(recs) == len(want) == len(rays) // tile
    for i, (g, w) in enumerate(zip(recs, want)):
        assert len(g) == len(w), f"tile {i}: {len(g)} records, oracle {len(w)}"
        assert np.array_equal(g.view(np.uint32).reshape(-1, 5), w.view(np.uint32).reshape(-1, 5)), f"tile {i}"
    v = np.where(started, lens, 0)
    assert np.array_equal(cnt["nodes"] + cnt["leaves"], v)
    ref = o.trace(rays, mode=omode)["counts"]
    assert np.array_equal(cnt["nodes"][started], ref["nodes"][started])
    assert np.array_equal(cnt["leaf_slots"][started], ref["leaf_slots"][started])
    return recs


@pytest.fixture(scope="module")
def spheres(rt):
    tris = rt.spheres_scene(0)
    out = {}
    for N, L in ((4, 4), (8, 8), (16, 16)):
        bvh = rt.Bvh(tris, N, L)
        out[(N, L)] = (rt.Scene(bvh), bvh)
    return out


@pytest.mark.parametrize("NL", [(4, 4), (8, 8), (16, 16)])
@pytest.mark.parametrize("tile", [16, 64, 100, 256, 1024])
def test_primary_tiles(rt, spheres, NL, tile):
    """Tiles of wg x wg rays, including blocks smaller than a wave (4x4) and
    not a multiple of it (10x10)."""
    scene, bvh = spheres[NL]
    wg = int(round(tile ** 0.5))
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 4 * wg, 2 * wg, wg)
    _check(scene, bvh, rays, tile, "closest")


@pytest.mark.parametrize("tile", [1, 4])
def test_tiny_tiles(rt, spheres, tile):
    """Blocks of 1 and 4 threads (workGroupSize 1 and 2): fewer threads than
    the block's five step counters, which must still start at zero."""
    scene, bvh = spheres[(8, 8)]
    wg = int(round(tile ** 0.5))
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 16 * wg, 8 * wg, wg)
    _check(scene, bvh, rays, tile, "closest")


def test_ambient_tiles_skip_nan_origins(rt, spheres):
    """Ambient any-hit rays of a primary frame; pixels without a hit carry a
    NaN origin and are not started (terminations of record 0)."""
    scene, bvh = spheres[(8, 8)]
    prim = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 128, 64, 16)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    g = np.random.default_rng(4)
    d = g.normal(size=(len(prim), 3)).astype(np.float32)
    flip = (d * np.nan_to_num(nrm)).sum(1) < 0
    d[flip] *= -1
    amb = rt.make_rays(np.zeros_like(d), d, tmax=4.0)
    o = pos + nrm * np.float32(0.001)
    amb["ox"], amb["oy"], amb["oz"] = o[:, 0], o[:, 1], o[:, 2]
    miss = hits["slot"] < 0
    assert miss.any() and (~miss).any()
    recs = _check(scene, bvh, amb, 256, "any")
    assert sum(int(r[0]["terminations"]) for r in recs) == int(miss.sum())


def test_step_cap_recall(rt, spheres):
    """A cap below the longest tile: the entry point reports the full record
    counts and the wrapper calls again with the larger cap."""
    scene, bvh = spheres[(4, 4)]
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 64, 32, 16)
    _check(scene, bvh, rays, 256, "closest", step_cap=3)


def test_argument_errors(rt, spheres):
    scene, bvh = spheres[(4, 4)]
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 32, 16, 16)
    with pytest.raises(rt.RtError):
        scene.workgroup_steps(rays[:100], 64)
    with pytest.raises(rt.RtError):
        scene.workgroup_steps(rays, 2048)
    with pytest.raises(ValueError):
        scene.workgroup_steps(rays, 256, mode="save_distance")
