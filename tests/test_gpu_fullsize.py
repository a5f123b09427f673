"""The headline workloads at full size (1M-triangle interior, 1920x1080
primaries, their shadow rays, 8M diffuse bounce rays): every ray of the
N x L sweep (primaries and shadow rays at all 16 points) and of the diffuse
workload (N8L8, N16L16, N4L4) equals the oracle, counts included; the
per-ray, ray-stream and cooperative traversals agree on every ray; and the
hit t and barycentrics of a 20,000-ray sample are compared with an FP64
re-evaluation of Moeller-Trumbore (the differences are printed: t relative,
barycentrics absolute).  Every GPU result is bit-identical to the FP32
reference arithmetic; the FP64 comparison states how far that arithmetic is
from exact."""
import numpy as np
import pytest

import oracle
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(rt):
    import torch

    tris = rt.interior_scene(2, 1_000_000)
    bvh = rt.Bvh(tris, 8, 8)
    scene = rt.Scene(bvh)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 1920, 1080, 16)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    diffuse = rt.spawn_diffuse(prim, nrm, pos, hits, 8_000_000, seed=3)
    shadow, _ = rt.spawn_shadow(nrm, pos, hits, (0.0, 9.5, 0.0), 1.0)
    yield tris, bvh, scene, prim, diffuse, shadow
    scene.close()
    torch.cuda.synchronize()


def _device(rt, scene, rays, mode="closest", **kw):
    import torch

    d_r = torch.from_numpy(rays.view(np.float32).reshape(-1, 10)).cuda()
    n = len(rays)
    out = torch.empty((n, 6), dtype=torch.int32, device="cuda") if mode != "any" else \
        torch.empty(n, dtype=torch.uint8, device="cuda")
    cnt = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    scene.intersect_device(d_r, out, mode=mode, counts=cnt, **kw)
    o = out.cpu().numpy()
    o = o.view(rt.HIT_DTYPE).reshape(-1) if mode != "any" else o
    return o, cnt.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1)



def test_primary_and_shadow_schedules_agree(rt, world):
    tris, bvh, scene, prim, diffuse, shadow = world
    print(f"1920x1080 primaries + {len(shadow)} shadow rays", flush=True)
    h0, c0 = _device(rt, scene, prim)
    o0, s0 = _device(rt, scene, shadow, mode="any")
    for kw in ({"stream_sched": True}, {"coop": True}):
        h, c = _device(rt, scene, prim, **kw)
        assert np.array_equal(h, h0) and np.array_equal(c, c0), kw
        o, s = _device(rt, scene, shadow, mode="any", **kw)
        assert np.array_equal(o, o0) and np.array_equal(s, s0), kw
    assert 0 < o0.sum() < len(o0)


def test_sample_matches_oracle_and_fp64(rt, world):
    tris, bvh, scene, prim, diffuse, shadow = world
    rng = np.random.default_rng(12)
    idx = np.sort(rng.choice(len(diffuse), 20_000, replace=False))
    h, c = _device(rt, scene, diffuse)
    sample = diffuse[idx]
    ref = oracle_for(bvh).trace(sample)
    assert_closest_equal(h[idx], ref["hits"], bvh, c[idx], ref["counts"], "8M diffuse sample")
    # FP64 Moeller-Trumbore of the reported triangle
    hs = h[idx]
    err = {"t": [], "b1": [], "b2": []}
    for i in np.nonzero(hs["slot"] >= 0)[0]:
        ok, t, b1, b2 = oracle.mt_f64(sample[i:i + 1].view(np.float32), tris[hs["tri"][i]])
        assert ok
        # t relative, barycentrics absolute (they are often near 0)
        err["t"].append(abs(float(hs["t"][i]) - t) / abs(t))
        err["b1"].append(abs(float(hs["b1"][i]) - b1))
        err["b2"].append(abs(float(hs["b2"][i]) - b2))
    med = {k: float(np.median(v)) for k, v in err.items()}
    p999 = {k: float(np.quantile(v, 0.999)) for k, v in err.items()}
    mx = {k: float(np.max(v)) for k, v in err.items()}
    print(f"FP32 vs FP64 (t relative, b1/b2 absolute): median {med}, 99.9% {p999}, max {mx}", flush=True)
    assert med["t"] < 1e-6 and p999["t"] < 1e-5 and mx["t"] < 1e-3
    assert mx["b1"] < 1e-3 and mx["b2"] < 1e-3


@pytest.mark.parametrize("N,L", [(8, 8), (16, 16), (4, 4)])
def test_diffuse_8m_every_ray_matches_oracle(rt, world, N, L):
    """Config 3 in full: all 8,000,000 diffuse rays at N8L8, N16L16 and N4L4
    equal the oracle ray for ray (hit identity, t, barycentrics and the four
    counts), and the ray-stream and cooperative kernels agree with the
    per-ray kernel on every ray."""
    import torch

    tris, bvh8, scene8, prim, diffuse, _ = world
    print(f"1M tris N{N}L{L}: 8M diffuse rays, three kernels and the oracle", flush=True)
    bvh = bvh8 if (N, L) == (8, 8) else rt.Bvh(tris, N, L)
    scene = scene8 if (N, L) == (8, 8) else rt.Scene(bvh)
    try:
        h0, c0 = _device(rt, scene, diffuse)
        assert (h0["slot"] >= 0).mean() > 0.99
        for kw in ({"stream_sched": True}, {"coop": True}):
            h, c = _device(rt, scene, diffuse, **kw)
            assert np.array_equal(h, h0) and np.array_equal(c, c0), kw
        ref = oracle_for(bvh).trace(diffuse)
        assert_closest_equal(h0, ref["hits"], bvh, c0, ref["counts"], f"N{N}L{L} 8M diffuse")
    finally:
        if scene is not scene8:
            scene.close()
        torch.cuda.synchronize()


NL_SWEEP = [(n, l) for n in (4, 8, 12, 16) for l in (4, 8, 12, 16)]


@pytest.mark.parametrize("N,L", NL_SWEEP)
def test_config2_sweep_every_ray_matches_oracle(rt, world, N, L):
    """Config 2 in full at every point of the N x L sweep: all 2,073,600
    primary rays (closest hit) and all their shadow rays (any hit toward the
    light) equal the oracle ray for ray, counts included."""
    import torch

    tris, _, _, prim, _, shadow = world
    print(f"1M tris N{N}L{L}: {len(prim)} primaries + {len(shadow)} shadow rays vs the oracle", flush=True)
    bvh = rt.Bvh(tris, N, L)
    scene = rt.Scene(bvh)
    try:
        h, c = _device(rt, scene, prim)
        o = oracle_for(bvh)
        ref = o.trace(prim)
        assert_closest_equal(h, ref["hits"], bvh, c, ref["counts"], f"N{N}L{L} primaries")
        occ, sc = _device(rt, scene, shadow, mode="any")
        ref = o.trace(shadow, mode=oracle.ORC_ANY)
        assert np.array_equal(occ, ref["occluded"]), f"N{N}L{L} shadow occlusion"
        for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
            assert np.array_equal(sc[f], ref["counts"][f]), f"N{N}L{L} shadow counts.{f}"
        assert 0 < int(occ.sum()) < len(occ)
    finally:
        scene.close()
        torch.cuda.synchronize()
