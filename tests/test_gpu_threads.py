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


def test_waiting_call_does_not_block_other_streams(rt, world):
    """rt_trace waits for its launch without holding the scene's lock.  While
    thread A makes calls that each trace 1M rays and wait for them, thread B
    enqueues small calls on its own stream without waiting (RT_NO_SYNC): B's
    calls return in far less than one of A's waits (holding the lock across
    the wait would make each of B's calls wait for A's current call).  Every
    result equals the call made alone."""
    import time

    import torch

    scene, rays, alone = world
    big = np.resize(rays, 1 << 20)
    d_big = torch.from_numpy(big.view(np.float32).reshape(-1, 10)).cuda()
    out_big = torch.empty((len(big), 6), dtype=torch.int32, device="cuda")
    n = 4096
    d_small = torch.from_numpy(rays[:n].view(np.float32).reshape(-1, 10)).cuda()
    out_small = torch.empty((n, 6), dtype=torch.int32, device="cuda")
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    scene.intersect_device(d_big, out_big, stream=sa)  # scratch of both streams exists
    scene.intersect_device(d_small, out_small, stream=sb)
    t = time.perf_counter()
    scene.intersect_device(d_big, out_big, stream=sa)
    one_wait = time.perf_counter() - t
    started = threading.Event()
    a_done = []

    def a_work():
        started.set()
        t0 = time.perf_counter()
        for _ in range(20):
            scene.intersect_device(d_big, out_big, stream=sa)
        a_done.append(time.perf_counter() - t0)

    th = threading.Thread(target=a_work)
    th.start()
    started.wait()
    time.sleep(one_wait / 2)  # A is inside a call, waiting
    lat = []
    for _ in range(100):
        t0 = time.perf_counter()
        scene.intersect_device(d_small, out_small, stream=sb, sync=False)
        lat.append(time.perf_counter() - t0)
    b_total = sum(lat)
    th.join()
    sb.synchronize()
    print(f"A: 20 waiting calls of 1M rays in {a_done[0] * 1e3:.1f} ms ({one_wait * 1e3:.2f} ms alone); "
          f"B: 100 queued calls in {b_total * 1e3:.2f} ms, median {np.median(lat) * 1e6:.0f} us", flush=True)
    ref = torch.from_numpy(np.resize(alone["closest"][0], 1 << 20).view(np.int32).reshape(-1, 6)).cuda()
    assert torch.equal(out_big, ref)
    assert torch.equal(out_small, ref[:n])
    assert np.median(lat) < 0.25 * one_wait, (np.median(lat), one_wait)
