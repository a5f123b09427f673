"""The batching configuration at its stated scale: 64 interior scenes
(seeds 100..163, triangles clamp(LogNormal(ln 2e5, 1), 1e4, 2e6), rays
clamp(LogNormal(ln 6.4e4, 1.2), 4096, 2^20) per scene, half primary and half
diffuse) traced in ONE persistent hipGraph launch.  Every ray of every scene
equals the plain per-scene launch (hits and counts), and a 2,000-ray sample
per scene equals the oracle (bit-exact, counts included).  Two batches: all
scenes at N8L8 (the kernel specialised on node memory 8), and N/L varying
per scene over sixteen combinations (the generic kernel that reads the node
memory per scene, as an autotuned batch runs).  A third batch gives every
scene its own node format, launch shape and schedule (RT_SCENE_SHAPES: one
persistent launch per distinct shape in the graph)."""
import numpy as np
import pytest

from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu

MIXED_NL = [(n, l) for n in (4, 8, 12, 16) for l in (4, 8, 12, 16)]


def _world(rt, nl_of):
    from bvhraytracer_amd import workloads as wl

    ntris, nrays = wl.batch_sizes()
    out = []
    for i in range(wl.BATCH_SCENES):
        N, L = nl_of(i)
        out.append(wl.batch_scene(i, N, L, ntris[i], nrays[i]))
    print(f"64 scenes: {int(ntris.sum())} triangles, {sum(len(w[3]) for w in out)} rays", flush=True)
    return out


def _check_batch(rt, world, label, scene_shapes=False):
    import torch

    dev = []
    for tris, bvh, sc, rays in world:
        d_r = torch.from_numpy(rays.view(np.float32).reshape(-1, 10)).cuda()
        n = len(rays)
        dev.append((d_r, torch.empty((n, 6), dtype=torch.int32, device="cuda"),
                    torch.empty((n, 4), dtype=torch.int32, device="cuda")))
    batch = rt.Batch(64)
    try:
        items = [(w[2], d_r, o, "closest", c) for w, (d_r, o, c) in zip(world, dev)]
        print(f"{label}: one graph launch over {len(items)} items", flush=True)
        batch.trace(items, graph=True, scene_shapes=scene_shapes)
        torch.cuda.synchronize()
        assert batch.stats()["error_rays"] == 0
        g = np.random.default_rng(44)
        for i, ((tris, bvh, sc, rays), (d_r, o, c)) in enumerate(zip(world, dev)):
            ref = torch.empty_like(o)
            ref_c = torch.empty_like(c)
            sc.intersect_device(d_r, ref, counts=ref_c)
            assert torch.equal(o, ref), f"{label} scene {i}: hits differ from the plain launch"
            assert torch.equal(c, ref_c), f"{label} scene {i}: counts differ from the plain launch"
            idx = np.sort(g.choice(len(rays), min(2000, len(rays)), replace=False))
            hits = o.cpu().numpy().view(rt.HIT_DTYPE).reshape(-1)[idx]
            cnt = c.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1)[idx]
            want = oracle_for(bvh).trace(rays[idx])
            assert_closest_equal(hits, want["hits"], bvh, cnt, want["counts"], f"{label} scene {i}")
    finally:
        batch.close()


def _close(world):
    import torch

    for w in world:
        w[2].close()
    torch.cuda.synchronize()


def test_config4_uniform_n8l8(rt):
    world = _world(rt, lambda i: (8, 8))
    try:
        _check_batch(rt, world, "N8L8 batch")
    finally:
        _close(world)


def test_config4_mixed_node_and_leaf_sizes(rt):
    world = _world(rt, lambda i: MIXED_NL[i % len(MIXED_NL)])
    try:
        _check_batch(rt, world, "mixed N/L batch")
    finally:
        _close(world)


SHAPES = [(False, 128, 8, 16), (True, 256, 16, 16), (False, 256, 16, 0), (True, 128, 32, 8), (False, 64, 8, 32),
          (True, 192, 16, 16)]


def test_config4_per_scene_shapes(rt):
    world = _world(rt, lambda i: MIXED_NL[(3 * i) % len(MIXED_NL)])
    try:
        for i, (tris, bvh, sc, rays) in enumerate(world):
            ex, bd, S, sched = SHAPES[i % len(SHAPES)]
            sc.exact_boxes = ex
            sc.set_launch(bd, S)
            sc.set_schedule(sched)
        _check_batch(rt, world, "per-scene shape batch", scene_shapes=True)
    finally:
        _close(world)
