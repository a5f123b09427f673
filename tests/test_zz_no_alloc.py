"""rt_trace_noalloc (scratch in a fixed device array, no allocation) against
rt_trace on the same rays, through the C ABI: two calls in a row, the second
with a smaller input, in every mode with counts; results bit-identical."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(rt):
    tris = rt.interior_scene(41, 60_000)
    bvh = rt.Bvh(tris, 8, 8)
    scene = rt.Scene(bvh)
    prim = rt.camera_rays((0.0, 5.0, -14.0), (0.0, 3.0, 0.0), 256, 128, 16)
    h = scene.intersect(prim)
    nrm, pos = bvh.surface(prim, h)
    diff = rt.spawn_diffuse(prim, nrm, pos, h, 100_000, seed=41)
    shadow, _ = rt.spawn_shadow(nrm, pos, h, (0.0, 9.5, 0.0), 1.0)
    yield bvh, scene, np.concatenate([prim, diff]), shadow
    scene.close()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.float32).reshape(-1, 10)).cuda()


def _out(n, mode):
    import torch
    if mode == "any":
        return torch.empty(n, dtype=torch.uint8, device="cuda")
    return torch.empty((n, 6), dtype=torch.int32, device="cuda")


@pytest.mark.parametrize("mode", ["closest", "any", "save_distance"])
def test_noalloc_equals_rt_trace(rt, world, mode):
    import torch

    bvh, scene, rays, shadow = world
    src = shadow if mode == "any" else rays
    lib = rt.lib()
    flags = rt.host.MODES[mode] | rt._lib.RT_COUNTS | rt._lib.RT_DEVICE_PTRS
    for n in (len(src), len(src) // 3 + 17):
        print(f"rt_trace_noalloc {mode} n={n}", flush=True)
        d_r = _dev(src[:n])
        ref, got = _out(n, mode), _out(n, mode)
        ref_c = torch.empty((n, 4), dtype=torch.int32, device="cuda")
        got_c = torch.full((n, 4), -1, dtype=torch.int32, device="cuda")
        scene.intersect_device(d_r, ref, mode=mode, counts=ref_c)
        s = torch.cuda.current_stream()
        rc = lib.rt_trace_noalloc(scene._h, d_r.data_ptr(), n, flags, got.data_ptr(), got_c.data_ptr(),
                                  ctypes.c_void_p(s.cuda_stream))
        assert rc == 0, lib.rt_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (mode, n)
        assert torch.equal(got_c, ref_c), (mode, n)


def test_noalloc_refuses_stream_and_cooperative(rt, world):
    bvh, scene, rays, _ = world
    lib = rt.lib()
    for f in (rt._lib.RT_STREAM, rt._lib.RT_COOP):
        assert lib.rt_trace_noalloc(scene._h, 1, 1, f | rt._lib.RT_DEVICE_PTRS, 1, None, None) == -4
