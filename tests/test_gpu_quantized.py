"""Quantized node records: the per-ray closest and any-hit kernels test the
children on conservative 8-bit planes and the exact box at the child's pop.
Every result and every count must equal the exact-box kernels' (the
RT_EXACT_BOXES flag) and the oracle's, also for rays aimed exactly at box
corners and rays starting on box faces, where the conservative margin is
all that separates a hit from a miss."""
import numpy as np
import pytest

import oracle
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu


def child_boxes(bvh):
    """(lo, hi) float32 arrays of every real child box of the inner nodes,
    read from the FastNode arrays (plane stride = the node's loop count)."""
    M, G = bvh.node_memory, bvh.gang
    los, his = [], []
    for i in np.nonzero(~bvh.is_leaf())[0]:
        b = bvh.bounds[i]
        k = M
        if M != G and np.isnan(b[6 * M - 1]):
            k = int(b[6 * M - 2])
        c = int(bvh.order[i, 0].max()) + 1
        for s in range(c):
            los.append([b[p * k + s] for p in range(3)])
            his.append([b[(3 + p) * k + s] for p in range(3)])
    return np.array(los, np.float32), np.array(his, np.float32)


def make_rays(rt, o, d, tmax):
    r = np.zeros(len(o), rt.RAY_DTYPE)
    r["tmax"] = tmax
    for j, a in enumerate("xyz"):
        r["o" + a] = o[:, j]
        r["d" + a] = d[:, j]
        with np.errstate(divide="ignore"):
            r["i" + a] = np.float32(1.0) / d[:, j]
    return r


def corner_rays(rt, bvh, seed, count=6000):
    """Rays from random room points to exact child-box corners (tmax 1: the
    corner is the end of the segment, or 222222), and rays starting on a
    corner in a random direction."""
    rng = np.random.default_rng(seed)
    lo, hi = child_boxes(bvh)
    pick = rng.integers(0, len(lo), count)
    sel = rng.integers(0, 2, (count, 3)).astype(bool)
    corner = np.where(sel, hi[pick], lo[pick]).astype(np.float32)
    org = (rng.random((count, 3)) * np.float32([30, 10, 30]) - np.float32([15, 0, 15])).astype(np.float32)
    d = (corner - org).astype(np.float32)
    tmax = np.where(np.arange(count) % 2 == 0, np.float32(1.0), np.float32(222222.0)).astype(np.float32)
    to_corner = make_rays(rt, org, d, tmax)
    dirs = rng.normal(size=(count, 3)).astype(np.float32)
    from_corner = make_rays(rt, corner, dirs, np.full(count, 222222.0, np.float32))
    return np.concatenate([to_corner, from_corner])


@pytest.mark.parametrize("N,L", [(4, 4), (8, 8), (8, 4), (12, 4), (16, 16), (5, 3)])
def test_quantized_equals_exact_and_oracle(rt, N, L):
    tris = rt.interior_scene(31, 60_000)
    bvh = rt.Bvh(tris, N, L)
    scene = rt.Scene(bvh)
    assert scene.stats()["quantized_record_bytes"] == (128 if bvh.node_memory <= 8 else 192), \
        "a builder scene should qualify for the quantized records"
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 128, 96, 16)
    h = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, h)
    diff = rt.spawn_diffuse(prim, nrm, pos, h, 20_000, seed=N * 17 + L)
    rays = np.concatenate([prim, diff, corner_rays(rt, bvh, N * 31 + L)])
    print(f"quantized N{N}L{L}: {len(rays)} rays", flush=True)
    q, qc = scene.intersect(rays, counts=True)
    x, xc = scene.intersect(rays, counts=True, exact_boxes=True)
    assert np.array_equal(q.view(np.uint8), x.view(np.uint8)), "quantized and exact-box hits differ"
    assert np.array_equal(qc.view(np.uint8), xc.view(np.uint8)), "quantized and exact-box counts differ"
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(q, ref["hits"], bvh, qc, ref["counts"], f"quantized N{N}L{L}")
    shadow, _ = rt.spawn_shadow(nrm, pos, h, (0.0, 9.5, 0.0), 1.0)
    srays = np.concatenate([shadow, corner_rays(rt, bvh, N * 37 + L, 3000)])
    occ, oc = scene.intersect(srays, mode="any", counts=True)
    xo, xoc = scene.intersect(srays, mode="any", counts=True, exact_boxes=True)
    assert np.array_equal(occ, xo) and np.array_equal(oc.view(np.uint8), xoc.view(np.uint8))
    ro = oracle_for(bvh).trace(srays, mode=oracle.ORC_ANY)
    assert np.array_equal(occ, ro["occluded"])
    for f in ("nodes", "leaves", "node_slots", "leaf_slots"):
        assert np.array_equal(oc[f], ro["counts"][f]), f


def test_quantized_batch_equals_exact(rt):
    """The batched launch over quantized records (grouped by node memory)
    equals the plain exact-box launches."""
    import torch

    batch = rt.Batch(8)
    items, refs = [], []
    for seed, (N, L) in ((41, (4, 4)), (42, (8, 8)), (43, (16, 8))):
        tris = rt.interior_scene(seed, 30_000)
        bvh = rt.Bvh(tris, N, L)
        sc = rt.Scene(bvh)
        prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 96, 64, 16)
        h = sc.intersect(prim)
        nrm, pos = bvh.surface(prim, h)
        rays = np.concatenate([rt.spawn_diffuse(prim, nrm, pos, h, 8000, seed=seed), corner_rays(rt, bvh, seed, 2000)])
        d_r = torch.from_numpy(rays.view(np.float32).reshape(-1, 10)).cuda()
        o = torch.empty((len(rays), 6), dtype=torch.int32, device="cuda")
        c = torch.empty((len(rays), 4), dtype=torch.int32, device="cuda")
        items.append((sc, d_r, o, "closest", c))
        ro = torch.empty_like(o)
        rc = torch.empty_like(c)
        sc.intersect_device(d_r, ro, counts=rc, exact_boxes=True)
        refs.append((ro, rc, bvh, sc))
    batch.trace(items, graph=True)
    torch.cuda.synchronize()
    for (sc, d_r, o, _, c), (ro, rc, bvh, _) in zip(items, refs):
        assert torch.equal(o, ro) and torch.equal(c, rc), f"batch N{bvh.branching}"
    batch.close()
