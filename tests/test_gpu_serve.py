"""Serving: every request's result in the mix equals the same request traced
alone (and the oracle); admission by footprint, cancel, bad ids."""
import numpy as np
import pytest

import oracle
from tests.helpers import oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(rt):
    scenes = []
    for seed, (N, L), tri in ((31, (8, 8), 40_000), (32, (4, 4), 15_000), (33, (16, 16), 60_000)):
        tris = rt.interior_scene(seed, tri)
        bvh = rt.Bvh(tris, N, L)
        sc = rt.Scene(bvh)
        prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 128, 96, 16)
        h = sc.intersect(prim)
        nrm, pos = bvh.surface(prim, h)
        diff = rt.spawn_diffuse(prim, nrm, pos, h, 60_000, seed=seed)
        shadow, _ = rt.spawn_shadow(nrm, pos, h, (0.0, 9.5, 0.0), 1.0)
        scenes.append((bvh, sc, np.concatenate([prim, diff]), shadow))
    return scenes


def test_mix_equals_alone(rt, world):
    srv = rt.Server(hbm_cap_bytes=64 << 20, slots=2)
    ids = [srv.add_scene(sc) for _, sc, _, _ in world]
    g = np.random.default_rng(5)
    reqs = []
    for i in range(60):
        s = int(g.integers(0, len(world)))
        mode = "any" if g.random() < 0.5 else "closest"
        src = world[s][3] if mode == "any" else world[s][2]
        n = int(min(len(src), g.integers(500, 40_000)))
        off = int(g.integers(0, len(src) - n + 1))
        rays = src[off:off + n]
        rid, out = srv.submit(ids[s], rays, mode)
        reqs.append((rid, out, s, mode, rays))
    print(f"submitted {len(reqs)} requests", flush=True)
    for rid, out, s, mode, rays in reqs:
        srv.wait(rid, timeout=60)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["launches"] < len(reqs)  # requests were coalesced
    for rid, out, s, mode, rays in reqs:
        bvh, sc, _, _ = world[s]
        alone = sc.intersect(rays, mode=mode)
        if mode == "any":
            assert np.array_equal(out, alone)
            assert np.array_equal(out, oracle_for(bvh).trace(rays, mode=oracle.ORC_ANY)["occluded"])
        else:
            assert np.array_equal(out.view(np.uint8), alone.view(np.uint8))
    srv.close()


def test_device_requests_mixed_with_host_requests(rt, world):
    """Requests whose rays and results stay in device memory (RT_DEVICE_PTRS)
    join the same launches as host requests; each result equals the request
    traced alone.  A small admission cap holds host requests back (only they
    take pool bytes) while device requests pass."""
    import torch
    srv = rt.Server(hbm_cap_bytes=4 << 20, slots=2)
    ids = [srv.add_scene(sc) for _, sc, _, _ in world]
    g = np.random.default_rng(6)
    reqs = []
    for i in range(80):
        s = int(g.integers(0, len(world)))
        mode = "any" if g.random() < 0.5 else "closest"
        src = world[s][3] if mode == "any" else world[s][2]
        n = int(min(len(src), g.integers(500, 30_000)))
        off = int(g.integers(0, len(src) - n + 1))
        rays = src[off:off + n]
        if i % 2:
            d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.float32).reshape(-1, 10)).cuda()
            d_o = (torch.full((n,), 7, dtype=torch.uint8, device="cuda") if mode == "any"
                   else torch.full((n, 6), -7, dtype=torch.int32, device="cuda"))
            rid, out = srv.submit(ids[s], d_r, mode, out=d_o)
        else:
            rid, out = srv.submit(ids[s], rays, mode)
        reqs.append((rid, out, s, mode, rays))
    print(f"submitted {len(reqs)} requests, half of them device-resident", flush=True)
    for rid, *_ in reqs:
        srv.wait(rid, timeout=60)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for rid, out, s, mode, rays in reqs:
        sc = world[s][1]
        alone = sc.intersect(rays, mode=mode)
        got = out.cpu().numpy() if hasattr(out, "is_cuda") else out
        if mode == "any":
            assert np.array_equal(np.asarray(got).reshape(-1), alone)
        else:
            assert np.array_equal(np.ascontiguousarray(got).view(np.uint8).reshape(-1), alone.view(np.uint8))
    with pytest.raises(ValueError):
        srv.submit(ids[0], torch.zeros((4, 10), dtype=torch.float32, device="cuda"), "closest")
    srv.close()


def test_admission_cancel_and_errors(rt, world):
    bvh, sc, rays, _ = world[0]
    # 1 MiB cap: a 100k-ray request (6.4 MB) can never be admitted
    srv = rt.Server(hbm_cap_bytes=1 << 20, slots=1)
    sid = srv.add_scene(sc)
    with pytest.raises(rt.RtError) as e:
        srv.submit(sid, rays[:100_000])
    assert e.value.code == -6
    with pytest.raises(rt.RtError):
        srv.submit(sid + 5, rays[:10])
    # requests that only fit one at a time are serialized, all complete
    subs = [srv.submit(sid, rays[i * 9000:(i + 1) * 9000]) for i in range(8)]
    # cancel the last one if it has not been launched yet
    cancelled = False
    try:
        srv.cancel(subs[-1][0])
        cancelled = True
    except rt.RtError:
        pass
    for rid, out in subs[:-1]:
        srv.wait(rid, timeout=60)
    if cancelled:
        with pytest.raises(rt.RtError):
            srv.wait(subs[-1][0])
    else:
        srv.wait(subs[-1][0], timeout=60)
    st = srv.stats()
    assert st["rejected"] == 1
    assert st["peak_admitted_bytes"] <= 1 << 20
    ref = sc.intersect(rays[:9000])
    assert np.array_equal(subs[0][1].view(np.uint8), ref.view(np.uint8))
    srv.close()


def test_wait_timeout_and_close_in_flight(rt, world):
    """A zero-timeout wait right after submitting a 512k-ray request (20 MB of
    rays to copy in) times out with RT_ERR_BUSY and leaves the request alive;
    a later wait collects it (results equal rt_trace alone) and a third wait
    finds the id gone.  Closing the server with requests queued and in flight
    returns: launched work is waited for, queued requests are dropped."""
    bvh, sc, rays, _ = world[0]
    big = np.resize(rays, 1 << 19)
    srv = rt.Server(hbm_cap_bytes=1 << 30, slots=2)
    sid = srv.add_scene(sc)
    rid, out = srv.submit(sid, big)
    with pytest.raises(rt.RtError) as e:
        srv.wait(rid, timeout=0.0)
    assert e.value.code == -6
    srv.wait(rid, timeout=60)
    alone = sc.intersect(big)
    assert np.array_equal(out.view(np.uint8), alone.view(np.uint8))
    with pytest.raises(rt.RtError) as e:
        srv.wait(rid, timeout=0.0)
    assert e.value.code == -1
    for _ in range(6):
        srv.submit(sid, big)
    srv.close()
    # the scene outlives the server and still traces as before
    assert np.array_equal(sc.intersect(big[:5000]).view(np.uint8), alone[:5000].view(np.uint8))
