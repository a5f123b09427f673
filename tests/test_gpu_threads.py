"""Concurrent use of one scene from many host threads, as the reference
calls its kernels from OpenMP workers (cameraFast.h:320,465): every call
returns what the same call returns alone, with host buffers (the library's
own stream and staging) or with device buffers on each thread's own stream,
and rt_last_error() keeps each thread's own message (ctypes releases the
GIL during the foreign call, so the calls overlap)."""
import ctypes
import threading

import numpy as np
import pytest

from bvhraytracer_amd import _lib

pytestmark = pytest.mark.gpu

THREADS = 16
MODES = ("closest", "any", "save_distance")


@pytest.fixture(scope="module")
def world(rt):
    bvh = rt.Bvh(rt.spheres_scene(0), 8, 8)
    scene = rt.Scene(bvh)
    rays = rt.camera_rays((0.0, 4.0, 32.0), (0.0, 0.0, 0.0), 256, 256, 16)
    alone = {m: scene.intersect(rays, mode=m, counts=True) for m in MODES}
    yield scene, rays, alone
    scene.close()


def _run(work):
    errors = []

    def wrap(i):
        try:
            work(i)
        except BaseException as e:  # reported after join
            errors.append((i, repr(e)))

    ts = [threading.Thread(target=wrap, args=(i,)) for i in range(THREADS)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors


def test_host_buffers_from_many_threads(rt, world):
    scene, rays, alone = world
    parts = np.array_split(np.arange(len(rays)), THREADS)
    print(f"{THREADS} threads, host buffers, {len(rays)} rays", flush=True)

    def work(i):
        idx = parts[i]
        for rep in range(4):
            mode = MODES[(i + rep) % 3]
            out, cnt = scene.intersect(rays[idx], mode=mode, counts=True)
            ref_out, ref_cnt = alone[mode]
            assert np.array_equal(out, ref_out[idx]), (i, mode)
            assert np.array_equal(cnt, ref_cnt[idx]), (i, mode)

    _run(work)


def test_device_buffers_on_own_streams(rt, world):
    import torch

    scene, rays, alone = world
    parts = np.array_split(np.arange(len(rays)), THREADS)
    print(f"{THREADS} threads, device buffers on their own streams", flush=True)

    def work(i):
        idx = parts[i]
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            d_rays = torch.from_numpy(rays[idx].view(np.float32).reshape(-1, 10)).cuda()
            d_hits = torch.empty((len(idx), 6), dtype=torch.int32, device="cuda")
            d_cnt = torch.empty((len(idx), 4), dtype=torch.int32, device="cuda")
        for rep in range(4):
            scene.intersect_device(d_rays, d_hits, counts=d_cnt, stream=st)
            st.synchronize()
            got = d_hits.cpu().numpy().view(rt.HIT_DTYPE).reshape(-1)
            cnt = d_cnt.cpu().numpy().view(rt.COUNTS_DTYPE).reshape(-1)
            assert np.array_equal(got, alone["closest"][0][idx]), (i, rep)
            assert np.array_equal(cnt, alone["closest"][1][idx]), (i, rep)

    _run(work)


def test_last_error_is_per_thread(rt, world):
    scene, rays, alone = world
    lib = rt.lib()
    n = len(rays)
    barrier = threading.Barrier(THREADS, timeout=60)

    def work(i):
        hits = np.zeros(n, rt.HIT_DTYPE)
        barrier.wait()
        if i % 2:
            rc = lib.rt_trace(scene._h, rays.ctypes.data, n, _lib.RT_ANY | _lib.RT_SAVE_DIST,
                              hits.ctypes.data, None, None)
            assert rc == -1
            mine = b"exclusive"
        else:
            rc = lib.rt_trace(scene._h, rays.ctypes.data, n, _lib.RT_COUNTS, hits.ctypes.data, None, None)
            assert rc == -1
            mine = b"without a counts buffer"
        barrier.wait()  # every thread has failed once: the messages must not mix
        assert mine in lib.rt_last_error(), (i, lib.rt_last_error())
        assert lib.rt_trace(scene._h, rays.ctypes.data, n, 0, hits.ctypes.data, None, None) == 0
        assert np.array_equal(hits, alone["closest"][0]), i

    _run(work)


def test_many_streams_keep_bounded_scratch(rt, world):
    """48 calls on fresh torch streams (torch hands them out from a pool of
    32, so at most 32 distinct handles), each call returning before its
    kernel ends (RT_NO_SYNC): the scene keeps the scratch of 16 streams and
    hands the least recently used one to the next new stream after its last
    launch, so every result is intact and the library's device memory stays
    bounded."""
    import torch

    scene, rays, alone = world
    n = len(rays)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1, 10)).cuda()
    ref = torch.from_numpy(alone["closest"][0].view(np.int32).reshape(n, 6)).cuda()

    def library_bytes():
        torch.cuda.synchronize()
        free, _ = torch.cuda.mem_get_info()
        return -free - torch.cuda.memory_reserved()

    base = library_bytes()
    first = torch.cuda.Stream()
    out0 = torch.empty((n, 6), dtype=torch.int32, device="cuda")
    scene.intersect_device(d_rays, out0, stream=first)
    one = library_bytes() - base  # the scratch of one new stream
    outs = []
    print(f"48 streams, {one / 1e6:.1f} MB of scratch per stream", flush=True)
    for i in range(48):
        st = torch.cuda.Stream()
        st.wait_stream(torch.cuda.current_stream())
        out = torch.empty((n, 6), dtype=torch.int32, device="cuda")
        scene.intersect_device(d_rays, out, stream=st, sync=False)
        outs.append(out)
    torch.cuda.synchronize()
    grown = library_bytes() - base
    assert torch.equal(out0, ref)
    for i, out in enumerate(outs):
        assert torch.equal(out, ref), i
    assert grown <= 17 * max(one, 1 << 20), (grown, one)
