"""ORACLE — test infrastructure only.

ctypes binding of the CPU restatement in oracle.cpp.  Only tests/, the smoke
check and the benchmark's CPU-baseline leg import this module, and only to
check or time the reference algorithm; the HIP engine never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class ScanStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "n_points", "n_valid", "n_skipped", "n_block_cells", "n_voxel_updates", "n_unique_voxels",
        "n_new_blocks", "n_trunc_vox", "n_trunc_blk", "n_index_dropped",
        "n_vox_blocks_outside_blockdda", "n_blocks_after")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.or_map_create.restype = C.c_void_p
        L.or_map_create.argtypes = [fp]
        L.or_map_destroy.argtypes = [C.c_void_p]
        L.or_derived.argtypes = [fp, fp, ip]
        L.or_segment.argtypes = [fp, fp, C.c_float, fp, fp]
        L.or_cell_center.restype = C.c_float
        L.or_cell_center.argtypes = [C.c_float, C.c_float, ip]
        L.or_point_valid.argtypes = [fp, fp, C.c_float]
        L.or_traverse.argtypes = [fp, fp, C.c_float, C.c_int, ip, fp, ip]
        L.or_voxel_to_block.argtypes = [fp, fp, ip, ip]
        L.or_signed_distance.restype = C.c_double
        L.or_signed_distance.argtypes = [fp, fp, fp]
        L.or_fold.argtypes = [dp, dp, C.c_double, C.c_double]
        L.or_integrate.argtypes = [C.c_void_p, fp, C.c_int64, fp, C.POINTER(ScanStats)]
        L.or_integrate_mt.argtypes = [C.c_void_p, fp, C.c_int64, fp, C.c_int, C.POINTER(ScanStats)]
        L.or_publish.restype = C.c_int64
        L.or_publish.argtypes = [C.c_void_p, fp, fp, dp, dp, C.c_int64]
        L.or_gc.restype = C.c_int64
        L.or_gc.argtypes = [C.c_void_p, fp]
        L.or_num_blocks.restype = C.c_int64
        L.or_num_blocks.argtypes = [C.c_void_p]
        L.or_export.restype = C.c_int64
        L.or_export.argtypes = [C.c_void_p, ip, dp, dp, C.c_int64]
        L.or_stats_size.restype = C.c_int
        assert L.or_stats_size() == C.sizeof(ScanStats)
        _lib = L
    return _lib


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def derived(params):
    p, pp = _f(params)
    of = (C.c_float * 5)()
    oi = (C.c_int32 * 2)()
    lib().or_derived(pp, of, oi)
    return dict(half_vs=of[0], BL=of[1], half_BL=of[2], blk_eps=of[3], vox_eps=of[4],
                max_vox=oi[0], max_blk=oi[1])


def segment(p, s, T):
    p, pp = _f(p)
    s, sp = _f(s)
    a = np.zeros(3, np.float32)
    b = np.zeros(3, np.float32)
    lib().or_segment(pp, sp, C.c_float(T), a.ctypes.data_as(C.POINTER(C.c_float)),
                     b.ctypes.data_as(C.POINTER(C.c_float)))
    return a, b


def cell_center(x, L):
    k = C.c_int32()
    c = lib().or_cell_center(C.c_float(x), C.c_float(L), C.byref(k))
    return np.float32(c), int(k.value)


def point_valid(p, s, vs):
    p, pp = _f(p)
    s, sp = _f(s)
    return bool(lib().or_point_valid(pp, sp, C.c_float(vs)))


def traverse(a, b, L, max_cells):
    a, ap = _f(a)
    b, bp = _f(b)
    keys = np.zeros((max(max_cells, 1), 3), np.int32)
    cen = np.zeros((max(max_cells, 1), 3), np.float32)
    tr = C.c_int32()
    n = lib().or_traverse(ap, bp, C.c_float(L), max_cells, keys.ctypes.data_as(C.POINTER(C.c_int32)),
                          cen.ctypes.data_as(C.POINTER(C.c_float)), C.byref(tr))
    return keys[:n].copy(), cen[:n].copy(), bool(tr.value)


def voxel_to_block(params, vc):
    p, pp = _f(params)
    v, vp = _f(vc)
    kb = np.zeros(3, np.int32)
    loc = C.c_int32()
    ok = lib().or_voxel_to_block(pp, vp, kb.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(loc))
    return bool(ok), kb, int(loc.value)


def signed_distance(p, s, x):
    p, pp = _f(p)
    s, sp = _f(s)
    x, xp = _f(x)
    return lib().or_signed_distance(pp, sp, xp)


def fold(S, W, d, maxW):
    s = C.c_double(S)
    w = C.c_double(W)
    lib().or_fold(C.byref(s), C.byref(w), C.c_double(d), C.c_double(maxW))
    return s.value, w.value


class OracleMap:
    """One map of the FP64 oracle (params = vs, T, maxW, pubD2, gcD2)."""

    def __init__(self, params):
        self.params = np.asarray(params, dtype=np.float32)
        _, pp = _f(self.params)
        self.h = lib().or_map_create(pp)

    def close(self):
        if self.h:
            lib().or_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def integrate(self, points, sensor, threads: int = 0):
        pts, pp = _f(points)
        s, sp = _f(sensor)
        st = ScanStats()
        if threads and threads > 1:
            lib().or_integrate_mt(self.h, pp, pts.shape[0], sp, threads, C.byref(st))
        else:
            lib().or_integrate(self.h, pp, pts.shape[0], sp, C.byref(st))
        return st.as_dict()

    def publish(self, sensor, cap=None):
        s, sp = _f(sensor)
        n = lib().or_publish(self.h, sp, None, None, None, 0)
        pos = np.zeros((max(n, 1), 3), np.float32)
        sdf = np.zeros(max(n, 1), np.float64)
        w = np.zeros(max(n, 1), np.float64)
        lib().or_publish(self.h, sp, pos.ctypes.data_as(C.POINTER(C.c_float)),
                         sdf.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), n)
        return pos[:n], sdf[:n], w[:n]

    def gc(self, sensor):
        s, sp = _f(sensor)
        return int(lib().or_gc(self.h, sp))

    def num_blocks(self):
        return int(lib().or_num_blocks(self.h))

    def export(self):
        n = self.num_blocks()
        keys = np.zeros((max(n, 1), 3), np.int32)
        sdf = np.zeros((max(n, 1), 512), np.float64)
        w = np.zeros((max(n, 1), 512), np.float64)
        lib().or_export(self.h, keys.ctypes.data_as(C.POINTER(C.c_int32)),
                        sdf.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), n)
        return keys[:n], sdf[:n], w[:n]
