"""Host-side mirror of the reference's interface for the traversal path.

Reference counterparts:
  Bvh(prims, N, L, sortEachSplit, leafSplitOption) + recursiveBvh(sahFactor)
      src/accelerationStructure/bvh.h:28-54, bvh.cpp:28-40
  FastNodeManager<G, M, WG>(bvh, leafMemory, cache)   (the flattened arrays)
      src/accelerationStructure/fastNodeManager.h:158-215, 1495-1673
  FastNodeManager::intersect / intersectSaveDistance / intersectSecondary
      fastNodeManager.h:1232-1493  -> Scene.intersect(..., mode=...)
  intersectWideAlternative / intersectSecondaryWideAlternative
      fastNodeManager.h:835-1230   -> Scene.intersect(..., stream=True)
  getSurfaceNormalPosition      fastNodeManager.h:1699-1726 -> FastNodeManager.surface()
  Camera::getRayTargetPosition  src/camera.h:91-98 -> camera_rays()
  RayTracer::loadSahFactor      src/rayTracer.cpp:880-925 -> load_sah_factor()

The Scene methods run on the GPU through librt.so; nothing here computes a
traversal result on the CPU.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import (COUNTS_DTYPE, HIT_DTYPE, RAY_DTYPE, RtBatchItem, RtBuildParams, RtBvhDesc,
                   RtCamera, RtServerConfig, RtServerStats, RtStats, check, lib, ptr)


def _stream(stream):
    return None if stream is None else ctypes.c_void_p(stream.cuda_stream)


def roundup(x, g):
    return int(math.ceil(x / float(g)) * g)


def load_sah_factor(leaf_size, branching, gang=4, table_path=None):
    """RayTracer::loadSahFactor (rayTracer.cpp:880-925)."""
    p = table_path.encode() if table_path else None
    return float(lib().rt_sah_factor(int(leaf_size), int(branching), int(gang), p))


def generate_scene(kind, seed, count, param=0):
    """Procedural triangles as float32 (n, 9): p0, p1, p2."""
    L = lib()
    out = ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_uint32()
    k = {"spheres": _lib.RT_SCENE_SPHERES, "interior": _lib.RT_SCENE_INTERIOR}.get(kind, kind)
    check(L.rt_scene_generate(int(k), int(seed), int(count), int(param), ctypes.byref(out),
                              ctypes.byref(n)))
    try:
        arr = np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(n.value, 9)
    finally:
        L.rt_free(ctypes.cast(out, ctypes.c_void_p))
    return arr


def spheres_scene(seed=0, spheres=50, frequency=10):
    """100,002 triangles for the defaults: 50 frequency-10 icospheres + ground."""
    return generate_scene("spheres", seed, spheres, frequency)


def interior_scene(seed, triangles=1_000_000, clusters=0):
    return generate_scene("interior", seed, triangles, clusters)


class Bvh:
    """Bvh + recursiveBvh + FastNodeManager flattening, done by librt's host code.

    Holds the reference FastNode arrays (numpy views) that both the device
    upload and the oracle consume.
    """

    def __init__(self, tris, branching, leaf_size, gang=4, sort_each_split=True,
                 leaf_split_option=1, sah_factor=None, threads=0, node_memory=0, leaf_memory=0):
        tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        if sah_factor is None:
            sah_factor = load_sah_factor(leaf_size, branching, gang)
        self.branching, self.leaf_size, self.gang = int(branching), int(leaf_size), int(gang)
        self.sah_factor = float(sah_factor)
        self.tris = tris
        p = RtBuildParams(self.branching, self.leaf_size, self.gang, int(bool(sort_each_split)),
                          int(leaf_split_option), ctypes.c_float(self.sah_factor).value, int(threads),
                          int(node_memory), int(leaf_memory))
        h = ctypes.c_void_p()
        check(lib().rt_build(ptr(tris), tris.shape[0], ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        d = RtBvhDesc()
        check(lib().rt_host_bvh_desc(h, ctypes.byref(d)))
        self.desc = d
        M, Lm, n = d.node_memory, d.leaf_memory, d.node_count
        self.node_memory, self.leaf_memory, self.node_count = M, Lm, n

        def view(addr, ctype, count, dtype):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            buf = (ctype * count).from_address(addr)
            return np.frombuffer(buf, dtype=dtype, count=count)

        self.bounds = view(d.bounds, ctypes.c_float, n * 6 * M, np.float32).reshape(n, 6 * M)
        self.base = view(d.base, ctypes.c_uint32, n, np.uint32)
        self.order = view(d.order, ctypes.c_int8, n * 3 * M, np.int8).reshape(n, 3, M)
        self.child_type = view(d.child_type, ctypes.c_uint16, n, np.uint16)
        self.tri_points = view(d.tri_points, ctypes.c_float, d.tri_float_count, np.float32)
        nl = d.tri_float_count // (9 * Lm)
        self.tri_ids = view(d.tri_ids, ctypes.c_int32, nl * Lm, np.int32)
        self.average_depth = d.average_depth
        self.triangle_count = d.triangle_count
        self.max_depth = d.max_depth

    def is_leaf(self):
        return np.isnan(self.bounds[:, 0])

    def surface(self, rays, hits):
        """getSurfaceNormalPosition for every hit: (normal (n,3), position (n,3))."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
        n = len(rays)
        nrm = np.empty((n, 3), np.float32)
        pos = np.empty((n, 3), np.float32)
        check(lib().rt_hit_surface(self._h, ptr(rays), ptr(hits), n, ptr(nrm), ptr(pos)))
        return nrm, pos

    def footprint(self):
        b = ctypes.c_uint64()
        check(lib().rt_scene_footprint(ctypes.byref(self.desc), ctypes.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "_h", None):
            lib().rt_host_bvh_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FastNodeArrays:
    """Caller-provided FastNode arrays (the reference layout, rt_bvh_desc).

    Lets a host that already flattened its BVH (or a test with a hand-built
    tree) hand the arrays to Scene without going through the builder."""

    def __init__(self, node_memory, leaf_memory, gang, bounds, base, order, child_type,
                 tri_points, tri_ids, branching=None, leaf_size=None):
        self.node_memory, self.leaf_memory, self.gang = int(node_memory), int(leaf_memory), int(gang)
        self.bounds = np.ascontiguousarray(bounds, np.float32).reshape(-1, 6 * self.node_memory)
        self.base = np.ascontiguousarray(base, np.uint32)
        self.order = np.ascontiguousarray(order, np.int8).reshape(-1, 3, self.node_memory)
        self.child_type = np.ascontiguousarray(child_type, np.uint16)
        self.tri_points = np.ascontiguousarray(tri_points, np.float32).reshape(-1)
        self.tri_ids = np.ascontiguousarray(tri_ids, np.int32)
        self.node_count = len(self.base)
        self.desc = RtBvhDesc(branching or self.node_memory, leaf_size or self.leaf_memory, self.gang,
                              self.node_memory, self.leaf_memory, self.node_count,
                              ptr(self.bounds), ptr(self.base), ptr(self.order), ptr(self.child_type),
                              self.tri_points.size, ptr(self.tri_points), ptr(self.tri_ids),
                              0.0, 0, 0)

    def is_leaf(self):
        return np.isnan(self.bounds[:, 0])


MODES = {"closest": _lib.RT_CLOSEST, "any": _lib.RT_ANY, "save_distance": _lib.RT_SAVE_DIST}


class Scene:
    """A BVH resident in HBM (rt_upload_bvh) with the traversal entry points."""

    def __init__(self, bvh):
        self.bvh = bvh
        h = ctypes.c_void_p()
        check(lib().rt_upload_bvh(ctypes.byref(bvh.desc), ctypes.byref(h)))
        self._h = h
        self._exact_boxes = False

    @property
    def exact_boxes(self):
        """Node format of the per-ray closest / any-hit calls that do not name
        one and of this scene's items in batches traced with scene_shapes:
        the quantized records (False) or the exact-box records (True)."""
        return self._exact_boxes

    @exact_boxes.setter
    def exact_boxes(self, value):
        check(lib().rt_set_node_format(self._h, _lib.RT_EXACT_BOXES if value else 0))
        self._exact_boxes = bool(value)

    def set_launch(self, block_size=0, stack_window=0):
        check(lib().rt_set_launch(self._h, int(block_size), int(stack_window)))

    def set_schedule(self, leaf_ratio16):
        """Wave schedule of the per-ray kernel (rt_set_schedule)."""
        check(lib().rt_set_schedule(self._h, int(leaf_ratio16)))

    def intersect(self, rays, mode="closest", counts=False, stream=False, coop=False, exact_boxes=None):
        """Host-array form: rays (n,) RAY_DTYPE -> hits (n,) HIT_DTYPE, or
        occluded (n,) uint8 for mode='any'; plus counts when asked.
        exact_boxes: the exact-box node records (RT_EXACT_BOXES); None = the
        scene's choice (Scene.exact_boxes)."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = len(rays)
        flags = MODES[mode] | (_lib.RT_STREAM if stream else 0) | (_lib.RT_COOP if coop else 0)
        if self.exact_boxes if exact_boxes is None else exact_boxes:
            flags |= _lib.RT_EXACT_BOXES
        out = np.empty(n, np.uint8) if mode == "any" else np.empty(n, HIT_DTYPE)
        cnt = np.zeros(n, COUNTS_DTYPE) if counts else None
        if counts:
            flags |= _lib.RT_COUNTS
        check(lib().rt_trace(self._h, ptr(rays), n, flags, ptr(out), ptr(cnt), None))
        return (out, cnt) if counts else out

    def intersect_device(self, rays, out, mode="closest", counts=None, stream=None, sync=True,
                         stream_sched=False, coop=False, exact_boxes=None):
        """Device-buffer form: rays/out/counts are torch CUDA tensors (or
        device addresses); stream is a torch.cuda.Stream or None."""
        n = rays.shape[0] if hasattr(rays, "shape") else int(rays[1])
        flags = MODES[mode] | _lib.RT_DEVICE_PTRS
        if self.exact_boxes if exact_boxes is None else exact_boxes:
            flags |= _lib.RT_EXACT_BOXES
        if counts is not None:
            flags |= _lib.RT_COUNTS
        if not sync:
            flags |= _lib.RT_NO_SYNC
        if stream_sched:
            flags |= _lib.RT_STREAM
        if coop:
            flags |= _lib.RT_COOP
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        check(lib().rt_trace(self._h, ptr(rays), n, flags, ptr(out),
                             ptr(counts) if counts is not None else None, s))

    def histograms(self, rays, mode="closest"):
        """rt_trace_histograms on host rays: returns (out, stats) where out is
        what intersect() returns and stats a dict of the visit counters
        (numpy arrays for the histograms)."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = len(rays)
        out = np.empty(n, np.uint8) if mode == "any" else np.empty(n, HIT_DTYPE)
        h = _lib.RtTraceHist()
        check(lib().rt_trace_histograms(self._h, ptr(rays), n, MODES[mode], ptr(out), ctypes.byref(h), None))
        stats = {}
        for f, _ in _lib.RtTraceHist._fields_:
            v = getattr(h, f)
            stats[f] = np.array(v[:], np.int64) if hasattr(v, "_length_") else int(v)
        return out, stats

    def workgroup_steps(self, rays, tile, mode="closest", step_cap=256):
        """rt_trace_steps on host rays: the per-step workload of tiles of
        `tile` consecutive rays.  Returns (tiles, counts): tiles[i] is a
        STEP_DTYPE array of tile i's records (record 0 = before the first
        step), counts the visits of every ray (COUNTS_DTYPE).  Calls again
        with a larger cap when a tile takes more than step_cap records."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = len(rays)
        if mode not in ("closest", "any"):
            raise ValueError("workgroup_steps: mode must be 'closest' or 'any'")
        nt = n // tile if tile else 0
        cnt = np.zeros(n, COUNTS_DTYPE)
        while True:
            out = np.zeros(max(1, nt * step_cap), _lib.STEP_DTYPE)
            lens = np.zeros(max(1, nt), np.uint32)
            check(lib().rt_trace_steps(self._h, ptr(rays), n, MODES[mode], int(tile), int(step_cap), ptr(out),
                                       ptr(lens), ptr(cnt), None))
            need = int(lens[:nt].max()) if nt else 0
            if need <= step_cap:
                break
            step_cap = need
        out = out[:nt * step_cap].reshape(nt, step_cap) if nt else out[:0]
        return [out[i, :lens[i]].copy() for i in range(nt)], cnt

    def render_ambient(self, pos, target, width, height, wg=16, samples=1, dist=1.0, mode="closest", wide=False,
                       up=(0.0, 1.0, 0.0), focal=0.866):
        """rt_render_ambient: one performance-renderer frame on the GPU.
        Returns (gray values per pixel in tile order, (primary ms, frame ms))."""
        cam = RtCamera((ctypes.c_float * 3)(*pos), (ctypes.c_float * 3)(*target), (ctypes.c_float * 3)(*up),
                       float(focal))
        if mode not in ("closest", "save_distance"):
            raise ValueError("render_ambient: mode must be 'closest' or 'save_distance'")
        flags = MODES[mode] | (_lib.RT_STREAM if wide else 0)
        gray = np.empty(width * height, np.uint8)
        ms = (ctypes.c_float * 2)()
        check(lib().rt_render_ambient(self._h, ctypes.byref(cam), width, height, wg, flags, int(samples), float(dist),
                                      ptr(gray), ms, None))
        return gray, (ms[0], ms[1])

    def surface_device(self, rays, hits, normal=None, position=None, stream=None):
        """rt_hit_surface_device: hit reconstruction from the resident
        triangles.  rays (n,10) float32, hits (n,6) int32 CUDA tensors;
        returns (normal (n,3), position (n,3)) CUDA tensors."""
        import torch

        n = rays.shape[0]
        if normal is None:
            normal = torch.empty((n, 3), dtype=torch.float32, device=rays.device)
        if position is None:
            position = torch.empty((n, 3), dtype=torch.float32, device=rays.device)
        check(lib().rt_hit_surface_device(self._h, ptr(rays), ptr(hits), n, ptr(normal), ptr(position),
                                          _stream(stream)))
        return normal, position

    def stats(self):
        s = RtStats()
        check(lib().rt_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in RtStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().rt_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """One persistent launch over many (scene, rays) items (rt_batch_trace).

    items: iterable of (scene, rays, out, mode[, counts]) with device buffers
    (torch CUDA tensors); mode 'closest' writes rt_hit records, 'any' uint8."""

    def __init__(self, max_items=1024):
        h = ctypes.c_void_p()
        check(lib().rt_batch_create(int(max_items), ctypes.byref(h)))
        self._h = h
        self.max_items = max_items

    def set_launch(self, block_size=0, stack_window=0):
        check(lib().rt_batch_set_launch(self._h, int(block_size), int(stack_window)))

    def trace(self, items, graph=True, stream=None, sync=True, exact_boxes=False, scene_shapes=False):
        """exact_boxes: every item on the exact-box records.  scene_shapes:
        every scene's items with that scene's own launch shape, schedule and
        node format (RT_SCENE_SHAPES), one persistent launch per distinct
        shape."""
        arr = (RtBatchItem * len(items))()
        for i, it in enumerate(items):
            scene, rays, out, mode = it[:4]
            counts = it[4] if len(it) > 4 else None
            arr[i] = RtBatchItem(scene._h, ptr(rays), ptr(out), ptr(counts) if counts is not None else None,
                                 int(rays.shape[0]), MODES[mode])
        flags = (_lib.RT_GRAPH if graph else 0) | (0 if sync else _lib.RT_NO_SYNC)
        if exact_boxes:
            flags |= _lib.RT_EXACT_BOXES
        if scene_shapes:
            flags |= _lib.RT_SCENE_SHAPES
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        check(lib().rt_batch_trace(self._h, arr, len(items), flags, s))

    def stats(self):
        s = RtStats()
        check(lib().rt_batch_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in RtStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().rt_batch_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over resident scenes (rt_server_*): admission by device
    footprint, continuous batching into batched persistent launches."""

    def __init__(self, hbm_cap_bytes=0, slots=0, max_items=0, block_size=0, stack_window=0,
                 use_graph=True, copy_threads=0):
        cfg = RtServerConfig(int(hbm_cap_bytes), int(slots), int(max_items), int(block_size),
                             int(stack_window), int(bool(use_graph)), int(copy_threads))
        h = ctypes.c_void_p()
        check(lib().rt_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._scenes = []
        self._live = {}  # request id -> (rays, out) kept alive until waited

    def add_scene(self, scene):
        sid = ctypes.c_uint32()
        check(lib().rt_server_add_scene(self._h, scene._h, ctypes.byref(sid)))
        self._scenes.append(scene)
        return sid.value

    def submit(self, scene_id, rays, mode="closest", out=None):
        """Returns (request id, output array filled when the request completes).
        out: optional caller-owned result array (HIT_DTYPE or uint8, len(rays)).
        Device requests: rays a CUDA tensor of (n, 10) float32 ray records and
        out a CUDA tensor of (n, 6) int32 (closest) or n uint8 (any); the
        server reads and writes them in place (RT_DEVICE_PTRS)."""
        if hasattr(rays, "is_cuda") and rays.is_cuda:
            n = rays.shape[0]
            if out is None or not getattr(out, "is_cuda", False) or out.shape[0] != n or not out.is_contiguous():
                raise ValueError("submit: a device request needs a contiguous CUDA `out` of len(rays) results")
            rid = ctypes.c_uint64()
            check(lib().rt_server_submit(self._h, int(scene_id), ptr(rays), n, MODES[mode] | _lib.RT_DEVICE_PTRS,
                                         ptr(out), ctypes.byref(rid)))
            self._live[rid.value] = (rays, out)
            return rid.value, out
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        want = np.dtype(np.uint8) if mode == "any" else HIT_DTYPE
        if out is None:
            out = np.empty(len(rays), want)
        elif out.dtype != want or len(out) != len(rays) or not out.flags.c_contiguous:
            raise ValueError("submit: out must be a contiguous array of len(rays) results of the mode's type")
        rid = ctypes.c_uint64()
        check(lib().rt_server_submit(self._h, int(scene_id), ptr(rays), len(rays), MODES[mode], ptr(out),
                                     ctypes.byref(rid)))
        self._live[rid.value] = (rays, out)
        return rid.value, out

    def wait(self, rid, timeout=-1.0):
        rc = lib().rt_server_wait(self._h, int(rid), float(timeout))
        if rc != _lib.RT_ERR_BUSY:
            self._live.pop(rid, None)
        check(rc)

    def cancel(self, rid):
        check(lib().rt_server_cancel(self._h, int(rid)))

    def drain(self):
        check(lib().rt_server_drain(self._h))

    def stats(self, reset=False):
        s = RtServerStats()
        check(lib().rt_server_get_stats(self._h, ctypes.byref(s), int(bool(reset))))
        return {f: getattr(s, f) for f, _ in RtServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().rt_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def camera_rays(pos, target, width, height, wg=16, up=(0.0, 1.0, 0.0), focal=0.866):
    """Primary rays in CameraFast tile order (cameraFast.h:318-342)."""
    cam = RtCamera((ctypes.c_float * 3)(*pos), (ctypes.c_float * 3)(*target),
                   (ctypes.c_float * 3)(*up), float(focal))
    out = np.empty(width * height, RAY_DTYPE)
    check(lib().rt_camera_rays(ctypes.byref(cam), width, height, wg, ptr(out)))
    return out


def spawn_shadow(normal, position, hits, light, tmax=1.0):
    hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
    n = len(hits)
    out = np.empty(n, RAY_DTYPE)
    src = np.empty(n, np.uint32)
    k = ctypes.c_uint32()
    lt = (ctypes.c_float * 3)(*light)
    check(lib().rt_spawn_shadow(ptr(np.ascontiguousarray(normal, np.float32)),
                                ptr(np.ascontiguousarray(position, np.float32)), ptr(hits), n,
                                lt, float(tmax), ptr(out), ptr(src), ctypes.byref(k)))
    return out[:k.value].copy(), src[:k.value].copy()


def spawn_ambient(normal, position, hits, sample, dist):
    """rt_spawn_ambient: ambient-occlusion rays of one sample for every hit
    pixel (compacted); returns (rays, source pixel index)."""
    hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
    n = len(hits)
    out = np.empty(n, RAY_DTYPE)
    src = np.empty(n, np.uint32)
    k = ctypes.c_uint32()
    check(lib().rt_spawn_ambient(ptr(np.ascontiguousarray(normal, np.float32)),
                                 ptr(np.ascontiguousarray(position, np.float32)), ptr(hits), n, int(sample),
                                 float(dist), ptr(out), ptr(src), ctypes.byref(k)))
    return out[:k.value].copy(), src[:k.value].copy()


def spawn_diffuse(primary, normal, position, hits, count, seed=3, tmax=222222.0):
    primary = np.ascontiguousarray(primary, dtype=RAY_DTYPE)
    hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
    out = np.empty(count, RAY_DTYPE)
    check(lib().rt_spawn_diffuse(ptr(primary), ptr(np.ascontiguousarray(normal, np.float32)),
                                 ptr(np.ascontiguousarray(position, np.float32)), ptr(hits),
                                 len(hits), int(seed), int(count), float(tmax), ptr(out)))
    return out


def camera_rays_device(pos, target, width, height, wg=16, up=(0.0, 1.0, 0.0), focal=0.866, out=None,
                       stream=None):
    """rt_camera_rays_device: primary rays in CameraFast tile order, made on
    the GPU; returns a (width*height, 10) float32 CUDA tensor."""
    import torch

    cam = RtCamera((ctypes.c_float * 3)(*pos), (ctypes.c_float * 3)(*target),
                   (ctypes.c_float * 3)(*up), float(focal))
    if out is None:
        out = torch.empty((width * height, 10), dtype=torch.float32, device="cuda")
    check(lib().rt_camera_rays_device(ctypes.byref(cam), width, height, wg, ptr(out), _stream(stream)))
    return out


def spawn_shadow_device(normal, position, hits, light, tmax=1.0, out=None, stream=None):
    """rt_spawn_shadow_device: one shadow ray per hit record (NaN origin for
    a miss); returns an (n, 10) float32 CUDA tensor."""
    import torch

    n = hits.shape[0]
    if out is None:
        out = torch.empty((n, 10), dtype=torch.float32, device=hits.device)
    lt = (ctypes.c_float * 3)(*light)
    check(lib().rt_spawn_shadow_device(ptr(normal), ptr(position), ptr(hits), n, lt, float(tmax), ptr(out),
                                       _stream(stream)))
    return out


def spawn_diffuse_device(primary, normal, position, hits, count, seed=3, tmax=222222.0, out=None, stream=None):
    """rt_spawn_diffuse_device: the rays of spawn_diffuse, made on the GPU;
    returns a (count, 10) float32 CUDA tensor."""
    import torch

    if out is None:
        out = torch.empty((count, 10), dtype=torch.float32, device=hits.device)
    check(lib().rt_spawn_diffuse_device(ptr(primary), ptr(normal), ptr(position), ptr(hits), hits.shape[0],
                                        int(seed), int(count), float(tmax), ptr(out), _stream(stream)))
    return out


def make_rays(origins, targets, tmax=222222.0):
    """FastRay(pos, target, maxDistance) (fastRay.cpp:3-20) for arrays of
    points; direction = target - pos, inverse with FLT_MAX for 0."""
    o = np.asarray(origins, np.float32).reshape(-1, 3)
    t = np.asarray(targets, np.float32).reshape(-1, 3)
    d = (t - o).astype(np.float32)
    with np.errstate(divide="ignore"):
        inv = np.where(d == 0, np.float32(np.finfo(np.float32).max),
                       np.float32(1.0) / d).astype(np.float32)
    r = np.empty(len(o), RAY_DTYPE)
    r["tmax"] = np.float32(tmax)
    r["ox"], r["oy"], r["oz"] = o[:, 0], o[:, 1], o[:, 2]
    r["ix"], r["iy"], r["iz"] = inv[:, 0], inv[:, 1], inv[:, 2]
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    return r


def measure_gather(active_lanes=64, buffer_bytes=16384):
    """Wave-level 16-byte loads per second on the current GPU when each of the
    first `active_lanes` lanes of every wave loads from its own 128-byte line
    (rt_measure_gather): the ceiling of the traversal kernels' record loads."""
    out = ctypes.c_double()
    check(lib().rt_measure_gather(int(active_lanes), int(buffer_bytes), ctypes.byref(out)))
    return out.value
