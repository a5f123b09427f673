"""Argument handling of rt_trace and rt_batch_trace through the C ABI: zero
rays read no buffer, bad flag combinations and missing buffers come back as
status codes with a message in rt_last_error(), and a scene or batch that
refused a call still traces correctly afterwards (results equal the
oracle's).  The reference's ISPC entry points check nothing
(fastNodeManager.h:271-275 maps any unknown loop count to the 16 variant);
these are the boundary's own rules."""
import ctypes

import numpy as np
import pytest

from bvhraytracer_amd import _lib
from tests.helpers import assert_closest_equal, oracle_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(rt):
    tris = rt.spheres_scene(1, spheres=6, frequency=4)
    bvh = rt.Bvh(tris, 4, 4)
    scene = rt.Scene(bvh)
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 64, 32, 16)
    yield bvh, scene, rays
    scene.close()


def _call(rt, scene, rays, n, flags, out, counts):
    lib = rt.lib()
    p = lambda a: None if a is None else a.ctypes.data
    rc = lib.rt_trace(scene._h, p(rays), n, flags, p(out), p(counts), None)
    return rc, lib.rt_last_error()


def test_trace_argument_errors(rt, small):
    bvh, scene, rays = small
    n = len(rays)
    hits = np.zeros(n, rt.HIT_DTYPE)
    cnt = np.zeros(n, rt.COUNTS_DTYPE)
    print(f"rt_trace argument errors, {n} rays", flush=True)
    # zero rays: no buffer is read, even null ones
    assert _call(rt, scene, None, 0, _lib.RT_CLOSEST, None, None)[0] == _lib.RT_OK
    assert _call(rt, scene, None, 0, _lib.RT_COUNTS | _lib.RT_DEVICE_PTRS, None, None)[0] == _lib.RT_OK
    cases = [
        ((None, n, _lib.RT_CLOSEST, hits, None), -1, b"null rays or output"),
        ((rays, n, _lib.RT_CLOSEST, None, None), -1, b"null rays or output"),
        ((rays, n, _lib.RT_COUNTS, hits, None), -1, b"RT_COUNTS without a counts buffer"),
        ((rays, n, _lib.RT_ANY | _lib.RT_SAVE_DIST, hits, None), -1, b"exclusive"),
        ((rays, n, _lib.RT_COOP | _lib.RT_SAVE_DIST, hits, None), -4, b"cooperative"),
        ((rays, n, _lib.RT_COOP | _lib.RT_STREAM, hits, None), -4, b"cooperative"),
        ((rays, n, _lib.RT_STREAM | _lib.RT_SAVE_DIST, hits, None), -4, b"save-distance"),
        ((rays, n, 512, hits, None), -1, b"unknown flag bits"),
        ((rays, n, _lib.RT_GRAPH, hits, None), -1, b"unknown flag bits"),
        ((rays, 0x80000000, _lib.RT_CLOSEST, hits, None), -1, b"2^31-1"),
    ]
    for args, code, msg in cases:
        rc, err = _call(rt, scene, *args)
        assert rc == code and msg in err, (args[1:3], rc, err)
    # refused calls leave nothing behind: a full call still equals the oracle
    rc, _ = _call(rt, scene, rays, n, _lib.RT_COUNTS, hits, cnt)
    assert rc == _lib.RT_OK
    ref = oracle_for(bvh).trace(rays)
    assert_closest_equal(hits, ref["hits"], bvh, cnt, ref["counts"], "after refused calls")


def test_histogram_and_steps_flag_errors(rt, small):
    bvh, scene, rays = small
    lib = rt.lib()
    h = _lib.RtTraceHist()
    out = np.zeros(len(rays), rt.HIT_DTYPE)
    for f in (_lib.RT_STREAM, _lib.RT_COOP, _lib.RT_COUNTS, _lib.RT_NO_SYNC):
        rc = lib.rt_trace_histograms(scene._h, rays.ctypes.data, len(rays), f, out.ctypes.data,
                                     ctypes.byref(h), None)
        assert rc == -1 and b"only RT_CLOSEST" in lib.rt_last_error()
    rc = lib.rt_trace_histograms(scene._h, rays.ctypes.data, len(rays), 0, out.ctypes.data, None, None)
    assert rc == -1 and b"null histogram" in lib.rt_last_error()
    with pytest.raises(rt.RtError, match="multiple of tile"):
        scene.workgroup_steps(rays[:300], 256)


def test_batch_argument_errors(rt, small):
    import torch

    bvh, scene, rays = small
    n = len(rays)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(n, 10)).cuda()
    o1 = torch.zeros((n, 6), dtype=torch.int32, device="cuda")
    o0 = torch.zeros((0, 6), dtype=torch.int32, device="cuda")
    b = rt.Batch(max_items=2)
    print("rt_batch_trace argument errors", flush=True)
    try:
        with pytest.raises(rt.RtError, match="more items"):
            b.trace([(scene, d_rays, o1, "closest")] * 3)
        with pytest.raises(rt.RtError, match="RT_CLOSEST and RT_ANY only"):
            b.trace([(scene, d_rays, o1, "save_distance")])
        lib = rt.lib()
        item = (_lib.RtBatchItem * 1)(_lib.RtBatchItem(scene._h, d_rays.data_ptr(), o1.data_ptr(), None, n, 0))
        assert lib.rt_batch_trace(b._h, item, 1, _lib.RT_COOP, None) == -1
        assert b"unknown flag bits" in lib.rt_last_error()
        assert lib.rt_batch_trace(b._h, None, 0, 0, None) == _lib.RT_OK
        # an empty item next to a full one: the empty one is skipped
        b.trace([(scene, d_rays[:0], o0, "closest"), (scene, d_rays, o1, "closest")])
        got = o1.cpu().numpy().view(rt.HIT_DTYPE).reshape(-1)
        ref = oracle_for(bvh).trace(rays)
        assert_closest_equal(got, ref["hits"], bvh, None, None, "batch after refused calls")
    finally:
        b.close()


def test_batch_scene_shape_limits(rt, small):
    """RT_SCENE_SHAPES runs one launch per distinct scene shape, at most 16:
    17 scenes with 17 schedules are refused with RT_ERR_ARG, and the same
    batch with 16 of them traces every item as the plain launch does.
    rt_set_node_format takes RT_EXACT_BOXES only."""
    import torch

    bvh, _, rays = small
    n = len(rays)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(n, 10)).cuda()
    scenes = [rt.Scene(bvh) for _ in range(17)]
    outs = [torch.zeros((n, 6), dtype=torch.int32, device="cuda") for _ in scenes]
    b = rt.Batch(max_items=32)
    try:
        for i, sc in enumerate(scenes):
            sc.set_schedule(i + 1)
        items = [(sc, d_rays, o, "closest") for sc, o in zip(scenes, outs)]
        print("17 distinct scene shapes", flush=True)
        with pytest.raises(rt.RtError, match="distinct scene launch shapes"):
            b.trace(items, scene_shapes=True)
        b.trace(items[:16], scene_shapes=True)
        ref = scenes[0].intersect(rays)
        for o in outs[:16]:
            got = o.cpu().numpy().view(rt.HIT_DTYPE).reshape(-1)
            assert np.array_equal(got, ref)
        lib = rt.lib()
        assert lib.rt_set_node_format(scenes[0]._h, _lib.RT_COUNTS) == -1
        assert b"RT_EXACT_BOXES only" in lib.rt_last_error()
    finally:
        b.close()
        for sc in scenes:
            sc.close()
