"""ctypes binding of librt.so (the C ABI declared in include/rt.h).

The library is loaded from the package directory; importing the package
raises when it is missing, so no caller can silently fall back to anything
else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librt.so")

RT_OK = 0
RT_ERR_BUSY = -6
RT_CLOSEST = 0
RT_ANY = 1
RT_SAVE_DIST = 2
RT_STREAM = 4
RT_COUNTS = 8
RT_DEVICE_PTRS = 16
RT_NO_SYNC = 32
RT_COOP = 64
RT_EXACT_BOXES = 256

RT_GRAPH = 128
RT_SCENE_SHAPES = 512

RT_SCENE_SPHERES = 1
RT_SCENE_INTERIOR = 2

# numpy views of the C records
RAY_DTYPE = np.dtype([("tmax", "<f4"), ("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"),
                      ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4")])
HIT_DTYPE = np.dtype([("tri", "<i4"), ("leaf", "<u4"), ("slot", "<i4"),
                      ("t", "<f4"), ("b1", "<f4"), ("b2", "<f4")])
COUNTS_DTYPE = np.dtype([("nodes", "<u4"), ("leaves", "<u4"),
                         ("node_slots", "<u4"), ("leaf_slots", "<u4")])


class RtBvhDesc(ctypes.Structure):
    _fields_ = [
        ("branching", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32), ("gang", ctypes.c_uint32),
        ("node_memory", ctypes.c_uint32), ("leaf_memory", ctypes.c_uint32),
        ("node_count", ctypes.c_uint32),
        ("bounds", ctypes.c_void_p), ("base", ctypes.c_void_p), ("order", ctypes.c_void_p),
        ("child_type", ctypes.c_void_p),
        ("tri_float_count", ctypes.c_uint64), ("tri_points", ctypes.c_void_p),
        ("tri_ids", ctypes.c_void_p),
        ("average_depth", ctypes.c_float), ("triangle_count", ctypes.c_uint32),
        ("max_depth", ctypes.c_uint32),
    ]


class RtBuildParams(ctypes.Structure):
    _fields_ = [
        ("branching", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32), ("gang", ctypes.c_uint32),
        ("sort_each_split", ctypes.c_uint32), ("leaf_split_option", ctypes.c_int32),
        ("sah_factor", ctypes.c_float), ("threads", ctypes.c_uint32),
        ("node_memory", ctypes.c_uint32), ("leaf_memory", ctypes.c_uint32),
    ]


class RtCamera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("target", ctypes.c_float * 3),
                ("up", ctypes.c_float * 3), ("focal", ctypes.c_float)]


class RtStats(ctypes.Structure):
    _fields_ = [
        ("bytes_resident", ctypes.c_uint64), ("inner_nodes", ctypes.c_uint32),
        ("leaves", ctypes.c_uint32), ("node_record_bytes", ctypes.c_uint32),
        ("leaf_block_bytes", ctypes.c_uint32), ("last_kernel_ms", ctypes.c_float),
        ("rays_traced", ctypes.c_uint64), ("block_size", ctypes.c_uint32),
        ("stack_window", ctypes.c_uint32), ("max_stack", ctypes.c_uint32),
        ("error_rays", ctypes.c_uint32), ("quantized_record_bytes", ctypes.c_uint32),
    ]


RT_HIST_DEPTHS = 64
RT_HIST_FULL = 17


class RtTraceHist(ctypes.Structure):
    _fields_ = [
        ("rays", ctypes.c_uint64), ("node_visits", ctypes.c_uint64), ("leaf_visits", ctypes.c_uint64),
        ("aabb_tests", ctypes.c_uint64), ("aabb_hits", ctypes.c_uint64),
        ("tri_tests", ctypes.c_uint64), ("tri_hits", ctypes.c_uint64),
        ("node_depth", ctypes.c_uint64 * RT_HIST_DEPTHS), ("leaf_depth", ctypes.c_uint64 * RT_HIST_DEPTHS),
        ("child_fullness", ctypes.c_uint64 * RT_HIST_FULL), ("leaf_fullness", ctypes.c_uint64 * RT_HIST_FULL),
    ]


class RtBatchItem(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("rays", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("counts", ctypes.c_void_p), ("n", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class RtServerConfig(ctypes.Structure):
    _fields_ = [("hbm_cap_bytes", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("max_items", ctypes.c_uint32), ("block_size", ctypes.c_uint32),
                ("stack_window", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("copy_threads", ctypes.c_uint32)]


class RtServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("rays_completed", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("wall_s", ctypes.c_double),
                ("mrays_per_s", ctypes.c_double), ("peak_admitted_bytes", ctypes.c_uint64),
                ("mean_requests_per_launch", ctypes.c_double)]


# rt_step_stat
STEP_DTYPE = np.dtype([("node_work", "<u4"), ("node_unique", "<u4"), ("leaf_work", "<u4"),
                       ("leaf_unique", "<u4"), ("terminations", "<u4")])


# (name, restype, argtypes) of every exported entry point in include/rt.h and
# include/rt_report.h
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
SIGNATURES = [
    ("rt_upload_bvh", ctypes.c_int, [ctypes.POINTER(RtBvhDesc), ctypes.POINTER(_P)]),
    ("rt_trace", ctypes.c_int, [_P, _P, _U32, _U32, _P, _P, _P]),
    ("rt_trace_noalloc", ctypes.c_int, [_P, _P, _U32, _U32, _P, _P, _P]),
    ("rt_stats", ctypes.c_int, [_P, ctypes.POINTER(RtStats)]),
    ("rt_trace_histograms", ctypes.c_int, [_P, _P, _U32, _U32, _P, ctypes.POINTER(RtTraceHist), _P]),
    ("rt_trace_steps", ctypes.c_int, [_P, _P, _U32, _U32, _U32, _U32, _P, _P, _P, _P]),
    ("rt_check_recip", ctypes.c_int, [_U32, ctypes.c_uint64, _P, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64), _P]),
    ("rt_measure_gather", ctypes.c_int, [_U32, _U32, ctypes.POINTER(ctypes.c_double)]),
    ("rt_set_launch", ctypes.c_int, [_P, _U32, _U32]),
    ("rt_set_node_format", ctypes.c_int, [_P, _U32]),
    ("rt_set_schedule", ctypes.c_int, [_P, _U32]),
    ("rt_release", ctypes.c_int, [_P]),
    ("rt_last_error", ctypes.c_char_p, []),
    ("rt_scene_footprint", ctypes.c_int, [ctypes.POINTER(RtBvhDesc), ctypes.POINTER(ctypes.c_uint64)]),
    ("rt_batch_create", ctypes.c_int, [_U32, ctypes.POINTER(_P)]),
    ("rt_batch_trace", ctypes.c_int, [_P, ctypes.POINTER(RtBatchItem), _U32, _U32, _P]),
    ("rt_batch_set_launch", ctypes.c_int, [_P, _U32, _U32]),
    ("rt_batch_stats", ctypes.c_int, [_P, ctypes.POINTER(RtStats)]),
    ("rt_batch_release", ctypes.c_int, [_P]),
    ("rt_server_create", ctypes.c_int, [ctypes.POINTER(RtServerConfig), ctypes.POINTER(_P)]),
    ("rt_server_add_scene", ctypes.c_int, [_P, _P, ctypes.POINTER(_U32)]),
    ("rt_server_submit", ctypes.c_int, [_P, _U32, _P, _U32, _U32, _P, ctypes.POINTER(ctypes.c_uint64)]),
    ("rt_server_wait", ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_double]),
    ("rt_server_cancel", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("rt_server_drain", ctypes.c_int, [_P]),
    ("rt_server_get_stats", ctypes.c_int, [_P, ctypes.POINTER(RtServerStats), ctypes.c_int]),
    ("rt_server_destroy", ctypes.c_int, [_P]),
    ("rt_scene_generate", ctypes.c_int, [_U32, ctypes.c_uint64, _U32, _U32,
                                         ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
                                         ctypes.POINTER(_U32)]),
    ("rt_free", None, [_P]),
    ("rt_build", ctypes.c_int, [_P, _U32, ctypes.POINTER(RtBuildParams), ctypes.POINTER(_P)]),
    ("rt_host_bvh_desc", ctypes.c_int, [_P, ctypes.POINTER(RtBvhDesc)]),
    ("rt_host_bvh_free", ctypes.c_int, [_P]),
    ("rt_camera_rays", ctypes.c_int, [ctypes.POINTER(RtCamera), _U32, _U32, _U32, _P]),
    ("rt_hit_surface", ctypes.c_int, [_P, _P, _P, _U32, _P, _P]),
    ("rt_render_ambient", ctypes.c_int, [_P, ctypes.POINTER(RtCamera), _U32, _U32, _U32, _U32, _U32, ctypes.c_float,
                                         _P, _P, _P]),
    ("rt_spawn_ambient", ctypes.c_int, [_P, _P, _P, _U32, _U32, ctypes.c_float, _P, _P, _P]),
    ("rt_spawn_shadow", ctypes.c_int, [_P, _P, _P, _U32, _P, ctypes.c_float, _P, _P,
                                       ctypes.POINTER(_U32)]),
    ("rt_spawn_diffuse", ctypes.c_int, [_P, _P, _P, _P, _U32, ctypes.c_uint64, _U32,
                                        ctypes.c_float, _P]),
    ("rt_camera_rays_device", ctypes.c_int, [ctypes.POINTER(RtCamera), _U32, _U32, _U32, _P, _P]),
    ("rt_hit_surface_device", ctypes.c_int, [_P, _P, _P, _U32, _P, _P, _P]),
    ("rt_spawn_shadow_device", ctypes.c_int, [_P, _P, _P, _U32, _P, ctypes.c_float, _P, _P]),
    ("rt_spawn_diffuse_device", ctypes.c_int, [_P, _P, _P, _P, _U32, ctypes.c_uint64, _U32,
                                               ctypes.c_float, _P, _P]),
    ("rt_sah_factor", ctypes.c_float, [_U32, _U32, _U32, ctypes.c_char_p]),
    # include/rt_report.h
    ("rt_report_read_config", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_size_t)]),
    ("rt_report_write_perf", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _U32, _U32, _U32, _U32, _U32,
                                            ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_uint64]),
    ("rt_report_write_info", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _U32, _U32, _U32,
                                            ctypes.POINTER(RtTraceHist), ctypes.POINTER(RtTraceHist)]),
    ("rt_report_write_workgroup_analysis", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                                          ctypes.c_int32, _U32, _U32, _U32, _P, _P, _P, _P, _P, _P,
                                                          _P, _P, ctypes.c_int32, ctypes.c_char_p,
                                                          ctypes.c_size_t]),
    ("rt_report_tile_order_to_rows", ctypes.c_int, [_P, _U32, _U32, _U32, _U32, _P]),
    ("rt_report_write_png", ctypes.c_int, [ctypes.c_char_p, _P, _U32, _U32]),
    ("rt_report_ambient_gray", ctypes.c_int, [_P, _P, _U32, _U32, _P]),
    ("rt_report_ambient_distance", ctypes.c_int, [_P, _U32, ctypes.POINTER(ctypes.c_float)]),
]

_lib = None


class RtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rt error {code}: {msg}")
        self.code = code


def lib():
    """Returns the loaded librt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run bvhraytracer_amd/build.py "
                              "(or __graft_entry__.build()) first")
        # PyTorch-ROCm ships its own libamdhip64.so (soname libamdhip64.so.7).
        # Loading torch first makes librt bind to that same runtime by soname,
        # so device memory, streams and kernels share one HIP runtime; loading
        # librt first would bring a second HIP/HSA runtime into the process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != RT_OK:
        raise RtError(rc, lib().rt_last_error().decode(errors="replace"))
    return rc


def ptr(a):
    """Address of a numpy array or a torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()
