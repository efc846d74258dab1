"""ctypes binding of the C ABI in include/tsdf_hip.h.

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _build

TSDF_OK = 0
TSDF_ERR_INVALID_ARG = -1
TSDF_ERR_HIP = -2
TSDF_ERR_NO_MEMORY = -3
TSDF_ERR_CAPACITY = -4
TSDF_ERR_BUDGET = -5
TSDF_ERR_DEADLINE = -6
TSDF_ERR_DROPPED = -7

TSDF_STATE_F32 = 0
TSDF_STATE_F16 = 1
TSDF_STATE_F16_QW = 2

TSDF_FIELD_FLOAT32 = 7

STAGES = ["traverse", "assign", "scan", "scatter", "fold", "publish", "gc", "total"]


class DroppedUpdatesWarning(RuntimeWarning):
    """A scan ran to the end but its map could not store every update (TSDF_ERR_DROPPED)."""


class TsdfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("tsdf error %d: %s" % (code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("truncation_distance", C.c_float), ("max_weight", C.c_float),
                ("publish_distance_squared", C.c_float), ("garbage_collect_distance_squared", C.c_float)]


class EngineConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("state_format", C.c_int32), ("heap_blocks", C.c_int64),
                ("total_buckets", C.c_int64), ("max_batch_points", C.c_int64)]


class MapConfig(C.Structure):
    _fields_ = [("buckets_log2", C.c_int32), ("overflow_buckets", C.c_int32), ("max_chain", C.c_int32),
                ("reserved", C.c_int32), ("block_quota", C.c_int64)]


class ScanStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "points", "valid", "skipped", "block_cells", "voxel_updates", "unique_voxels", "new_blocks",
        "trunc_vox", "trunc_blk", "index_dropped", "insert_failures", "heap_failures", "updates_dropped",
        "published", "gc_removed", "quota_failures")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class ScanDesc(C.Structure):
    _fields_ = [("map_index", C.c_int32), ("n_points", C.c_int32), ("point_offset", C.c_int64),
                ("sensor", C.c_float * 3), ("reserved", C.c_int32)]


class LaunchConfig(C.Structure):
    _fields_ = [("traverse_block", C.c_int32), ("traverse_grid_cap", C.c_int32), ("fold_grid", C.c_int32),
                ("rays_per_lane", C.c_int32), ("traverse_queue", C.c_int32), ("reserved", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("map_index", C.c_int32), ("n_points", C.c_int32), ("point_offset", C.c_int64),
                ("sensor", C.c_float * 3), ("arrival_s", C.c_float)]


class ServeConfig(C.Structure):
    _fields_ = [("max_batch_scans", C.c_int32), ("max_inflight", C.c_int32), ("max_batch_points", C.c_int64),
                ("deadline_ms", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("requests", C.c_int64), ("batches", C.c_int64), ("deadline_misses", C.c_int64),
                ("wall_s", C.c_double), ("busy_s", C.c_double), ("points", C.c_int64)]


class Cloud2Layout(C.Structure):
    _fields_ = [("width", C.c_int64), ("height", C.c_int64), ("row_step", C.c_int64), ("data_bytes", C.c_int64),
                ("point_step", C.c_int32), ("x_offset", C.c_int32), ("y_offset", C.c_int32), ("z_offset", C.c_int32),
                ("datatype", C.c_int32), ("is_bigendian", C.c_int32)]


vp = C.c_void_p
fp = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)

# name -> (argtypes); every entry point returns int except where noted
SIGNATURES = {
    "tsdf_engine_create": [C.POINTER(EngineConfig), C.POINTER(vp)],
    "tsdf_engine_destroy": [vp],
    "tsdf_engine_synchronize": [vp],
    "tsdf_engine_free_blocks": [vp, i64p],
    "tsdf_map_create": [vp, C.POINTER(Params), C.POINTER(MapConfig), C.POINTER(vp)],
    "tsdf_map_admit": [vp, C.POINTER(Params), C.POINTER(MapConfig), C.POINTER(vp)],
    "tsdf_map_destroy": [vp],
    "tsdf_map_num_blocks": [vp, i64p],
    "tsdf_map_export": [vp, i32p, fp, C.c_int64, i64p],
    "tsdf_map_import": [vp, i32p, fp, C.c_int64, i64p, i64p],
    "tsdf_integrate": [vp, vp, C.c_int64, fp],
    "tsdf_integrate_noalloc": [vp, vp, C.c_int64, fp],
    "tsdf_publish": [vp, fp, fp, fp, C.c_int64, i64p],
    "tsdf_publish_msg": [vp, fp, fp, C.c_int64, i64p],
    "tsdf_gc": [vp, fp, i64p],
    "tsdf_process_scan": [vp, fp, C.c_int64, fp, fp, fp, C.c_int64, i64p],
    "tsdf_last_stats": [vp, C.POINTER(ScanStats)],
    "tsdf_integrate_batch": [vp, C.POINTER(vp), C.c_int32, C.POINTER(ScanDesc), C.c_int32, vp],
    "tsdf_batch_graph_create": [vp, C.POINTER(vp), C.c_int32, C.c_int64, C.c_int32, C.POINTER(vp)],
    "tsdf_batch_graph_launch": [vp, C.POINTER(ScanDesc), C.c_int32, vp],
    "tsdf_batch_graph_destroy": [vp],
    "tsdf_engine_set_launch": [vp, C.POINTER(LaunchConfig)],
    "tsdf_engine_get_launch": [vp, C.POINTER(LaunchConfig)],
    "tsdf_timing_enable": [vp, C.c_int],
    "tsdf_timing_read": [vp, C.POINTER(C.c_double), i64p, C.c_int],
    "tsdf_serve_run": [vp, C.POINTER(vp), C.c_int32, C.POINTER(Request), C.c_int32, vp, C.POINTER(ServeConfig),
                       C.POINTER(C.c_double), i32p, C.POINTER(ServeStats)],
    "tsdf_device_alloc": [vp, C.c_int64, C.POINTER(vp)],
    "tsdf_device_free": [vp, vp],
    "tsdf_host_alloc": [vp, C.c_int64, C.POINTER(vp)],
    "tsdf_host_free": [vp],
    "tsdf_copy_to_device": [vp, vp, vp, C.c_int64],
    "tsdf_copy_to_host": [vp, vp, vp, C.c_int64],
    "tsdf_device_copy": [vp, vp, vp, C.c_int64],
    "tsdf_derived_params": [C.POINTER(Params), fp, i32p],
    "tsdf_cloud2_to_points": [vp, C.POINTER(Cloud2Layout), vp, C.c_int, vp, C.c_int64, i64p],
    "tsdf_msg_markers": [vp, vp, C.c_int64, C.c_int, fp, fp, i64p, i64p],
}
SPECIAL_RESTYPE = {"tsdf_last_error": C.c_char_p, "tsdf_engine_stream": vp}
SPECIAL_ARGS = {"tsdf_last_error": [], "tsdf_engine_stream": [vp]}

_lib = None


def library_path() -> str:
    """The built engine library; TSDF_HIP_LIBRARY overrides it (for comparing builds)."""
    return os.environ.get("TSDF_HIP_LIBRARY", _build.ENGINE_SO)


def load():
    """Load the built HIP library; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError("HIP engine library %s is missing: build it first "
                           "(python -m ros_tsdf_amd._build)" % path)
    L = C.CDLL(path)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    for name, rt in SPECIAL_RESTYPE.items():
        f = getattr(L, name)
        f.restype = rt
        f.argtypes = SPECIAL_ARGS[name]
    _lib = L
    return L


def check(code):
    if code != TSDF_OK:
        msg = load().tsdf_last_error()
        raise TsdfError(code, msg.decode() if msg else "")
    return code
