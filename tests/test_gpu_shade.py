"""Ray generation and hit shading on the GPU are bit-identical to the host
entry points (which test_raygen.py pins against float64 models), and an
all-device bounce pipeline (camera -> trace -> surface -> shadow / diffuse
-> trace) gives the host pipeline's results."""
import numpy as np
import pytest

import oracle
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu

CAM_POS, CAM_TGT = (0.0, 4.0, 32.0), (0.0, 0.0, 0.0)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _host(t, dtype):
    return t.cpu().numpy().view(dtype).reshape(-1)


@pytest.mark.parametrize("W,H,wg", [(1920, 1080, 16), (256, 256, 16), (64, 40, 8)])
def test_camera_rays_device(rt, W, H, wg):
    import torch

    print(f"camera {W}x{H} wg {wg}", flush=True)
    host = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), W, H, wg)
    dev = rt.camera_rays_device((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), W, H, wg)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(_host(dev, rt.RAY_DTYPE)), _bits(host))


@pytest.fixture(scope="module")
def frame(rt):
    """A traced 320x240 frame of the spheres scene (N8L8; rays hit spheres
    and the ground or miss) with host-made surfaces, and the same buffers on
    the device."""
    import torch

    tris = rt.spheres_scene(0)
    bvh = rt.Bvh(tris, 8, 8)
    scene = rt.Scene(bvh)
    prim = rt.camera_rays(CAM_POS, CAM_TGT, 320, 240, 16)
    hits = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, hits)
    d_prim = torch.from_numpy(prim.view(np.float32).reshape(-1, 10)).cuda()
    d_hits = torch.from_numpy(hits.view(np.int32).reshape(-1, 6)).cuda()
    return tris, bvh, scene, prim, hits, nrm, pos, d_prim, d_hits


def test_surface_device(rt, frame):
    import torch

    tris, bvh, scene, prim, hits, nrm, pos, d_prim, d_hits = frame
    assert 0 < (hits["slot"] >= 0).sum() < len(hits)
    dn, dp = scene.surface_device(d_prim, d_hits)
    torch.cuda.synchronize()
    gn, gp = dn.cpu().numpy(), dp.cpu().numpy()
    hit = hits["slot"] >= 0
    assert np.array_equal(_bits(gn[hit]), _bits(nrm[hit]))
    assert np.array_equal(_bits(gp[hit]), _bits(pos[hit]))
    assert np.isnan(gn[~hit]).all() and np.isnan(gp[~hit]).all()


def test_shadow_device(rt, frame):
    import torch

    tris, bvh, scene, prim, hits, nrm, pos, d_prim, d_hits = frame
    light = (0.0, 9.5, 0.0)
    host, src = rt.spawn_shadow(nrm, pos, hits, light, 1.0)
    d_n = torch.from_numpy(nrm).cuda()
    d_p = torch.from_numpy(pos).cuda()
    dev = _host(rt.spawn_shadow_device(d_n, d_p, d_hits, light, 1.0), rt.RAY_DTYPE)
    assert len(dev) == len(hits)
    assert np.array_equal(_bits(dev[src]), _bits(host))
    miss = hits["slot"] < 0
    assert miss.any() and np.all(dev["tmax"][miss] == -1.0)
    # traced: the dense device rays give the host rays' answers; a miss's ray
    # is never occluded and stops at the root
    occ_h = scene.intersect(host, mode="any")
    occ_d, cnt = scene.intersect(dev, mode="any", counts=True)
    assert np.array_equal(occ_d[src], occ_h)
    assert not occ_d[miss].any()
    assert np.all(cnt["nodes"][miss] == 1) and np.all(cnt["leaves"][miss] == 0)
    ro = oracle_for(bvh).trace(dev, mode=oracle.ORC_ANY)
    assert np.array_equal(occ_d, ro["occluded"]) and np.array_equal(cnt["nodes"], ro["counts"]["nodes"])


@pytest.mark.parametrize("count", [1, 50_000])
def test_diffuse_device(rt, frame, count):
    import torch

    tris, bvh, scene, prim, hits, nrm, pos, d_prim, d_hits = frame
    host = rt.spawn_diffuse(prim, nrm, pos, hits, count, seed=5)
    d_n = torch.from_numpy(nrm).cuda()
    d_p = torch.from_numpy(pos).cuda()
    dev = _host(rt.spawn_diffuse_device(d_prim, d_n, d_p, d_hits, count, seed=5), rt.RAY_DTYPE)
    assert np.array_equal(_bits(dev), _bits(host))


def test_diffuse_device_no_hits(rt, frame):
    import torch

    tris, bvh, scene, prim, hits, nrm, pos, d_prim, d_hits = frame
    none = torch.full_like(d_hits, -1)
    d_n = torch.from_numpy(nrm).cuda()
    d_p = torch.from_numpy(pos).cuda()
    dev = _host(rt.spawn_diffuse_device(d_prim, d_n, d_p, none, 100, seed=5), rt.RAY_DTYPE)
    assert np.all(dev["tmax"] == -1.0)
    g = scene.intersect(dev)
    assert np.all(g["slot"] == -1)


def test_device_bounce_pipeline(rt, frame):
    """camera -> trace -> surface -> diffuse -> trace, all on the device and
    on one stream, against the oracle on the host-made bounce rays."""
    import torch

    tris, bvh, scene, prim, hits, nrm, pos, d_prim, d_hits = frame
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        d_rays = rt.camera_rays_device(CAM_POS, CAM_TGT, 320, 240, 16, stream=stream)
        d_h = torch.empty((d_rays.shape[0], 6), dtype=torch.int32, device="cuda")
        scene.intersect_device(d_rays, d_h, stream=stream, sync=False)
        d_n, d_p = scene.surface_device(d_rays, d_h, stream=stream)
        d_b = rt.spawn_diffuse_device(d_rays, d_n, d_p, d_h, 40_000, seed=9, stream=stream)
        d_bh = torch.empty((40_000, 6), dtype=torch.int32, device="cuda")
        d_bc = torch.empty((40_000, 4), dtype=torch.int32, device="cuda")
        scene.intersect_device(d_b, d_bh, counts=d_bc, stream=stream, sync=False)
    stream.synchronize()
    bounce = rt.spawn_diffuse(prim, nrm, pos, hits, 40_000, seed=9)
    assert np.array_equal(_bits(_host(d_b, rt.RAY_DTYPE)), _bits(bounce))
    ref = oracle_for(bvh).trace(bounce)
    assert_closest_equal(_host(d_bh, rt.HIT_DTYPE), ref["hits"], bvh, _host(d_bc, rt.COUNTS_DTYPE),
                         ref["counts"], "device bounce")
    assert (ref["diag"] & oracle.DIAG_BOGUS).sum() == 0


def test_null_stream_chain(rt, frame):
    """Every device entry point called with no stream, without waiting in
    between: camera rays -> trace -> surfaces -> diffuse rays -> trace on the
    null stream's order.  The library traces on a stream of its own, which
    must wait for the null stream's work and make it wait in turn."""
    import torch

    tris, bvh, scene, prim, hits, nrm, pos, d_prim, d_hits = frame
    for rep in range(3):
        d_rays = rt.camera_rays_device(CAM_POS, CAM_TGT, 320, 240, 16)
        d_h = torch.empty((d_rays.shape[0], 6), dtype=torch.int32, device="cuda")
        scene.intersect_device(d_rays, d_h, sync=False)
        d_n, d_p = scene.surface_device(d_rays, d_h)
        d_b = rt.spawn_diffuse_device(d_rays, d_n, d_p, d_h, 40_000, seed=11)
        d_bh = torch.empty((40_000, 6), dtype=torch.int32, device="cuda")
        scene.intersect_device(d_b, d_bh, sync=False)
        got = _host(d_bh, rt.HIT_DTYPE)  # .cpu() runs on the null stream
        bounce = rt.spawn_diffuse(prim, nrm, pos, hits, 40_000, seed=11)
        assert np.array_equal(_bits(_host(d_b, rt.RAY_DTYPE)), _bits(bounce)), rep
        assert np.array_equal(got, scene.intersect(bounce)), rep
