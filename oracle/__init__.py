"""ORACLE — test infrastructure, not product code.

ctypes binding of liboracle.so (cpu_ref.c), the CPU restatement of the
reference traversal, and of liboracle_simd.so (cpu_simd.c), the same
traversals with 4-wide SSE node and leaf tests (the shape of the reference's
ISPC kernels; bit-identical results).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg use it, and only as the checker or the timed
CPU baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SIMD_LIB_PATH = os.path.join(HERE, "liboracle_simd.so")

ORC_CLOSEST, ORC_ANY, ORC_SAVE_DIST = 0, 1, 2
ORC_WIDE_AXIS_BUG, ORC_WIDE_SHARED_DIST, ORC_WIDE_ROOT_AS_NODE = 1, 2, 4
DIAG_STACK48, DIAG_BOGUS, DIAG_STACKCAP = 1, 2, 4

RAY_DTYPE = np.dtype([("tmax", "<f4"), ("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"),
                      ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4")])
HIT_DTYPE = np.dtype([("slot", "<i4"), ("leaf", "<u4"), ("t", "<f4"), ("b1", "<f4"), ("b2", "<f4")])
COUNTS_DTYPE = np.dtype([("nodes", "<u4"), ("leaves", "<u4"),
                         ("node_slots", "<u4"), ("leaf_slots", "<u4")])


class OrcBvh(ctypes.Structure):
    _fields_ = [("M", ctypes.c_uint32), ("Lm", ctypes.c_uint32), ("G", ctypes.c_uint32),
                ("node_count", ctypes.c_uint32), ("bounds", ctypes.c_void_p),
                ("base", ctypes.c_void_p), ("order", ctypes.c_void_p),
                ("child_type", ctypes.c_void_p), ("tri_float_count", ctypes.c_uint64),
                ("tri", ctypes.c_void_p)]


_P = ctypes.c_void_p
_libs = {}


def lib(simd=False):
    """liboracle.so, or with simd=True liboracle_simd.so (same entry points)."""
    if simd not in _libs:
        path = SIMD_LIB_PATH if simd else LIB_PATH
        if not os.path.exists(path):
            raise ImportError(f"{path} missing: run make -C oracle")
        L = ctypes.CDLL(path)
        sig = {
            "orc_aabb_intersect": (ctypes.c_int, [_P, ctypes.c_int, _P, _P]),
            "orc_tri_intersect": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int, _P, ctypes.c_int]),
            "orc_tri_anyhit": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int, _P]),
            "orc_node_loop_count": (ctypes.c_int, [_P, ctypes.c_uint32]),
            "orc_leaf_loop_count": (ctypes.c_int, [_P, ctypes.c_uint32]),
            "orc_dispatch_k": (ctypes.c_int, [ctypes.c_int]),
            "orc_intersect": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
            "orc_intersect_save_distance": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
            "orc_intersect_secondary": (ctypes.c_int, [_P, _P, _P, _P]),
            "orc_intersect_wide": (None, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P, _P]),
            "orc_intersect_secondary_wide": (None, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P]),
            "orc_trace": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint32, ctypes.c_int, _P, _P, _P, _P]),
            "orc_mt_f64": (ctypes.c_int, [_P, _P, _P, _P, _P]),
            "orc_bruteforce_closest": (ctypes.c_int32, [_P, ctypes.c_uint32, _P, _P]),
            "orc_bruteforce_any": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
            "orc_visit_sequences": (None, [_P, _P, ctypes.c_uint32, ctypes.c_int, _P, ctypes.c_uint32, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[simd] = L
    return _libs[simd]


def _p(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """Oracle view of a set of FastNode arrays (keeps the arrays alive).
    simd=True runs the traversals with the 4-wide SSE node and leaf tests."""

    def __init__(self, M, Lm, G, bounds, base, order, child_type, tri_points, simd=False):
        self.simd = bool(simd)
        self.arrays = [np.ascontiguousarray(bounds, np.float32), np.ascontiguousarray(base, np.uint32),
                       np.ascontiguousarray(order, np.int8), np.ascontiguousarray(child_type, np.uint16),
                       np.ascontiguousarray(tri_points, np.float32)]
        b, ba, o, ct, t = self.arrays
        self.M, self.Lm, self.G = int(M), int(Lm), int(G)
        self.node_count = len(ba)
        self.s = OrcBvh(self.M, self.Lm, self.G, self.node_count, _p(b), _p(ba), _p(o), _p(ct),
                        t.size, _p(t))

    @classmethod
    def from_bvh(cls, bvh, simd=False):
        return cls(bvh.node_memory, bvh.leaf_memory, bvh.gang, bvh.bounds, bvh.base, bvh.order,
                   bvh.child_type, bvh.tri_points, simd=simd)

    def trace(self, rays, mode=ORC_CLOSEST, stream=False, tile=256, threads=0, counts=True):
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
        n = len(rays)
        hits = np.zeros(n, HIT_DTYPE) if mode != ORC_ANY else None
        occ = np.zeros(n, np.uint8) if mode == ORC_ANY else None
        cnt = np.zeros(n, COUNTS_DTYPE) if counts else None
        diag = np.zeros(n, np.uint32)
        used = lib(self.simd).orc_trace(ctypes.byref(self.s), _p(rays), n, mode, int(stream), tile, threads,
                               _p(hits), _p(occ), _p(cnt), _p(diag))
        return {"hits": hits, "occluded": occ, "counts": cnt, "diag": diag, "threads": used}

    def visit_sequences(self, rays, mode=ORC_CLOSEST, cap=4096):
        """(ids, lens): ids[i, :lens[i]] are the FastNode ids ray i visits in
        order (leaves | 1 << 31).  Raises if a ray visits more than cap."""
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
        n = len(rays)
        ids = np.zeros((n, cap), np.uint32)
        lens = np.zeros(n, np.uint32)
        lib(self.simd).orc_visit_sequences(ctypes.byref(self.s), _p(rays), n, mode, _p(ids), cap, _p(lens))
        if n and int(lens.max()) > cap:
            raise ValueError(f"a ray takes {int(lens.max())} visits, more than cap {cap}")
        return ids, lens

    def wide(self, rays, flags=0):
        rays = np.array(rays, copy=True).view(RAY_DTYPE)
        n = len(rays)
        li = np.zeros(n, np.uint32)
        ti = np.zeros(n, np.int32)
        cnt = np.zeros(n, COUNTS_DTYPE)
        diag = np.zeros(n, np.uint32)
        lib(self.simd).orc_intersect_wide(ctypes.byref(self.s), _p(rays), n, flags, _p(li), _p(ti), _p(cnt), _p(diag))
        return li, ti, cnt, rays["tmax"].copy()

    def secondary_wide(self, rays, flags=0):
        rays = np.array(rays, copy=True).view(RAY_DTYPE)
        n = len(rays)
        res = np.zeros(n, np.uint16)
        cnt = np.zeros(n, COUNTS_DTYPE)
        diag = np.zeros(n, np.uint32)
        lib(self.simd).orc_intersect_secondary_wide(ctypes.byref(self.s), _p(rays), n, flags, _p(res), _p(cnt), _p(diag))
        return res, cnt


def aabb_intersect(bounds, k, dist, ray):
    return lib().orc_aabb_intersect(_p(bounds), k, _p(dist), _p(ray))


def tri_intersect(tri, tri_id, k, ray, gang=4):
    return lib().orc_tri_intersect(_p(tri), tri_id, k, _p(ray), gang)


def tri_anyhit(tri, tri_id, k, ray):
    return lib().orc_tri_anyhit(_p(tri), tri_id, k, _p(ray))


def mt_f64(ray, tri9):
    t, b1, b2 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    ok = lib().orc_mt_f64(_p(ray), _p(np.ascontiguousarray(tri9, np.float32)), ctypes.byref(t),
                          ctypes.byref(b1), ctypes.byref(b2))
    return ok, t.value, b1.value, b2.value


def bruteforce_closest(tris, ray):
    tris = np.ascontiguousarray(tris, np.float32)
    t = ctypes.c_float()
    i = lib().orc_bruteforce_closest(_p(tris), tris.shape[0], _p(ray), ctypes.byref(t))
    return i, t.value


def bruteforce_any(tris, ray):
    tris = np.ascontiguousarray(tris, np.float32)
    return lib().orc_bruteforce_any(_p(tris), tris.shape[0], _p(ray))


STEP_DTYPE = np.dtype([("node_work", "<u4"), ("node_unique", "<u4"), ("leaf_work", "<u4"),
                       ("leaf_unique", "<u4"), ("terminations", "<u4")])


def tile_step_records(ids, lens, tile, started):
    """Per-step records of tiles of `tile` consecutive rays, as the counting
    of CompactNodeManager::intersectWideAlternative (compactNode.cpp:358-601)
    produces them when every unfinished ray takes one visit per step in its
    own per-ray order (ids, lens from Oracle.visit_sequences).  Record 0 is
    the state before the loop (:419-424): rays started, 1, 0, 0, rays not
    started; record t counts the rays whose t-th visit is a node / a leaf,
    the distinct ids among them (:429-435,509-519 std::set) and the rays
    whose last visit is the t-th (:474-477,573-584).  A ray not started
    (started[i] false) takes no visit."""
    n = len(lens)
    out = []
    for i in range(n // tile):
        sl = slice(i * tile, (i + 1) * tile)
        st = np.asarray(started[sl], bool)
        v = np.where(st, lens[sl].astype(np.int64), 0)
        steps = int(v.max()) if len(v) else 0
        rec = np.zeros(steps + 1, STEP_DTYPE)
        rec[0] = (int(st.sum()), 1, 0, 0, int((~st).sum()))
        tid = ids[sl]
        for t in range(1, steps + 1):
            act = v >= t
            col = tid[act, t - 1]
            leaf = (col >> np.uint32(31)) != 0
            rec[t] = (int((~leaf).sum()), len(np.unique(col[~leaf])), int(leaf.sum()),
                      len(np.unique(col[leaf])), int((v == t).sum()))
        out.append(rec)
    return out
