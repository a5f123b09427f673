"""Batched launch parity: one persistent launch over several scenes (mixed
N/L, mixed closest/any items, uneven sizes) must give, for every ray, the
same result as the plain per-scene launch and the oracle."""
import numpy as np
import pytest

import oracle
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(rt):
    out = []
    for seed, (N, L), tri in ((21, (4, 4), 20_000), (22, (8, 8), 50_000), (23, (16, 8), 30_000),
                              (24, (12, 16), 40_000), (25, (8, 4), 5_000)):
        tris = rt.interior_scene(seed, tri)
        bvh = rt.Bvh(tris, N, L)
        sc = rt.Scene(bvh)
        prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 64, 48, 16)
        h = sc.intersect(prim)
        nrm, pos = bvh.surface(prim, h)
        diff = rt.spawn_diffuse(prim, nrm, pos, h, 5000 + 1000 * seed % 7000, seed=seed)
        shadow, _ = rt.spawn_shadow(nrm, pos, h, (0.0, 9.5, 0.0), 1.0)
        out.append((bvh, sc, np.concatenate([prim, diff]), shadow))
    return out


def dev(rt, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.float32).reshape(-1, 10)).cuda()


def run_batch(rt, scenes, order, graph, batch):
    import torch
    items, outs = [], []
    for i, kind in order:
        bvh, sc, rays, shadow = scenes[i]
        if kind == "closest":
            r = dev(rt, rays)
            o = torch.empty((len(rays), 6), dtype=torch.int32, device="cuda")
            c = torch.empty((len(rays), 4), dtype=torch.int32, device="cuda")
            items.append((sc, r, o, "closest", c))
        else:
            r = dev(rt, shadow)
            o = torch.empty(len(shadow), dtype=torch.uint8, device="cuda")
            c = torch.empty((len(shadow), 4), dtype=torch.int32, device="cuda")
            items.append((sc, r, o, "any", c))
        outs.append((i, kind, o, c))
    batch.trace(items, graph=graph)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_oracle(rt, scenes, graph):
    batch = rt.Batch(64)
    order = [(0, "closest"), (1, "any"), (2, "closest"), (3, "closest"), (1, "closest"),
             (4, "any"), (0, "any"), (4, "closest")]
    # two calls with different item sets: the graph is re-armed, not re-captured
    for rep, ordr in enumerate((order, order[::-1][:5])):
        print(f"batch graph={graph} call {rep}: {len(ordr)} items", flush=True)
        outs = run_batch(rt, scenes, ordr, graph, batch)
        for i, kind, o, c in outs:
            bvh, sc, rays, shadow = scenes[i]
            cnt = c.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1)
            if kind == "closest":
                hits = o.cpu().numpy().view(rt.HIT_DTYPE).reshape(-1)
                ref = oracle_for(bvh).trace(rays)
                assert_closest_equal(hits, ref["hits"], bvh, cnt, ref["counts"], f"batch scene {i}")
            else:
                occ = o.cpu().numpy()
                ref = oracle_for(bvh).trace(shadow, mode=oracle.ORC_ANY)
                assert np.array_equal(occ, ref["occluded"]), f"batch any scene {i}"
                assert np.array_equal(cnt["leaves"], ref["counts"]["leaves"])
    st = batch.stats()
    assert st["error_rays"] == 0 and st["rays_traced"] > 0
    batch.close()


def test_batch_equals_plain(rt, scenes):
    import torch
    batch = rt.Batch(16)
    order = [(i, "closest") for i in range(len(scenes))]
    outs = run_batch(rt, scenes, order, True, batch)
    for i, kind, o, c in outs:
        bvh, sc, rays, _ = scenes[i]
        r = dev(rt, rays)
        ref = torch.empty((len(rays), 6), dtype=torch.int32, device="cuda")
        sc.intersect_device(r, ref)
        assert torch.equal(o, ref), f"scene {i}"


def test_batch_uniform_node_memory(rt, scenes):
    """Batches whose scenes share one node memory run the kernel specialised
    on it; a graph captured for one node memory is re-captured when the next
    batch mixes them (the grouped kernel, in its form for node memories up to
    8 when no scene has more).  Every result equals the plain launch's."""
    import torch
    batch = rt.Batch(16)
    same_m = [(1, "closest"), (4, "any"), (4, "closest"), (1, "any")]  # N8L8 and N8L4: M = 8
    mixed = [(0, "closest"), (1, "closest"), (2, "any"), (3, "closest")]
    small_m = [(0, "closest"), (1, "any"), (4, "closest"), (0, "any")]  # M = 4 and 8: the small grouped kernel
    for rep, order in enumerate((same_m, mixed, same_m, small_m, mixed, small_m)):
        print(f"uniform-M batch call {rep}: {len(order)} items", flush=True)
        outs = run_batch(rt, scenes, order, True, batch)
        for i, kind, o, c in outs:
            bvh, sc, rays, shadow = scenes[i]
            rr = rays if kind == "closest" else shadow
            mode = kind
            ref, rc = sc.intersect(rr, mode=mode, counts=True)
            got = o.cpu().numpy()
            if kind == "closest":
                assert np.array_equal(np.ascontiguousarray(got).view(rt.HIT_DTYPE).reshape(-1), ref), f"scene {i}"
            else:
                assert np.array_equal(got, ref), f"scene {i}"
            cc = c.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1)
            assert np.array_equal(cc["nodes"], rc["nodes"]) and np.array_equal(cc["leaves"], rc["leaves"])
    batch.close()


def test_batch_many_small_items(rt, scenes):
    """9000 items of a few rays over five scenes (the kernel stages the item
    table in LDS, 144 KB of it, many rows per thread): every ray equals the
    plain launch, direct and replayed, and a later batch of 50 items replays
    the graph captured with the larger table.  A batch whose tables cannot fit the
    kernel's LDS is refused with RT_ERR_ARG and leaves the batch usable."""
    import torch
    refs, items, spans = [], [], []
    for i, (bvh, sc, rays, _) in enumerate(scenes):
        r = dev(rt, rays)
        ref = torch.empty((len(rays), 6), dtype=torch.int32, device="cuda")
        sc.intersect_device(r, ref)
        out = torch.full_like(ref, -7)
        refs.append((ref, out))
        cuts = np.linspace(0, len(rays), 1801).astype(int)
        items += [(sc, r[a:b], out[a:b], "closest") for a, b in zip(cuts[:-1], cuts[1:])]
        spans += [(i, a, b) for a, b in zip(cuts[:-1], cuts[1:])]
    few = [(it, sp) for it, sp in zip(items, spans) if sp[0] == 2][:50]
    rng = np.random.default_rng(5)
    items = [items[j] for j in rng.permutation(len(items))]
    batch = rt.Batch(30_000)
    for graph in (False, True):
        for _, out in refs:
            out.fill_(-7)
        print(f"many-item batch graph={graph}: {len(items)} items", flush=True)
        batch.trace(items, graph=graph)
        torch.cuda.synchronize()
        for k, (ref, out) in enumerate(refs):
            assert torch.equal(out, ref), f"scene {k} graph={graph}"
    refs[2][1].fill_(-7)
    batch.trace([it for it, _ in few], graph=True)
    torch.cuda.synchronize()
    last = few[-1][1][2]
    assert torch.equal(refs[2][1][:last], refs[2][0][:last]) and bool((refs[2][1][last:] == -7).all())
    # 25000 one-ray items: 400 KB of item table, more than a CU's LDS
    sc, r = scenes[1][1], dev(rt, scenes[1][2])
    one = torch.empty((1, 6), dtype=torch.int32, device="cuda")
    huge = [(sc, r[j % len(r):j % len(r) + 1], one, "closest") for j in range(25_000)]
    with pytest.raises(rt.RtError) as e:
        batch.trace(huge, graph=True)
    assert e.value.code == -1 and "LDS" in str(e.value)
    for _, out in refs:
        out.fill_(-7)
    batch.trace(items, graph=True)
    torch.cuda.synchronize()
    assert all(torch.equal(out, ref) for ref, out in refs)
    batch.close()


def test_batch_scene_shapes(rt, scenes):
    """RT_SCENE_SHAPES: every scene's items run with that scene's own launch
    shape, schedule and node format, one persistent launch per distinct shape
    in one graph.  Every result and count equals the plain launch's, direct
    and replayed, and after a call with other segments."""
    import torch
    shapes = [(64, 8, 16, True), (256, 16, 16, False), (128, 32, 0, False), (128, 8, 16, False), (192, 16, 8, True)]
    for (bvh, sc, _, _), (bd, S, sched, ex) in zip(scenes, shapes):
        sc.set_launch(bd, S)
        sc.set_schedule(sched)
        sc.exact_boxes = ex
    batch = rt.Batch(64)
    order = [(0, "closest"), (1, "any"), (2, "closest"), (3, "closest"), (1, "closest"), (4, "any"), (0, "any"),
             (4, "closest")]
    try:
        for rep, (ordr, graph) in enumerate(((order, False), (order, True), (order, True), (order[:3], True),
                                            (order, True))):
            print(f"scene-shape batch call {rep}: {len(ordr)} items, graph={graph}", flush=True)
            items, outs = [], []
            for i, kind in ordr:
                bvh, sc, rays, shadow = scenes[i]
                rr = rays if kind == "closest" else shadow
                r = dev(rt, rr)
                o = (torch.empty((len(rr), 6), dtype=torch.int32, device="cuda") if kind == "closest"
                     else torch.empty(len(rr), dtype=torch.uint8, device="cuda"))
                c = torch.empty((len(rr), 4), dtype=torch.int32, device="cuda")
                items.append((sc, r, o, kind, c))
                outs.append((i, kind, rr, o, c))
            batch.trace(items, graph=graph, scene_shapes=True)
            torch.cuda.synchronize()
            for i, kind, rr, o, c in outs:
                sc = scenes[i][1]
                ref, rc = sc.intersect(rr, mode=kind, counts=True)
                got = o.cpu().numpy()
                if kind == "closest":
                    got = np.ascontiguousarray(got).view(rt.HIT_DTYPE).reshape(-1)
                assert np.array_equal(got, ref), f"scene {i} {kind} call {rep}"
                cc = c.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1)
                assert np.array_equal(cc, rc), f"scene {i} {kind} counts call {rep}"
        st = batch.stats()
        assert st["error_rays"] == 0 and st["max_stack"] > 0
    finally:
        for _, sc, _, _ in scenes:
            sc.set_launch(256, 16)
            sc.set_schedule(16)
            sc.exact_boxes = False
        batch.close()
