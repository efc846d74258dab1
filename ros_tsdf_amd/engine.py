"""Python host interface over the C ABI: engines, maps, device buffers."""
from __future__ import annotations

import ctypes as C
import warnings
import weakref

import numpy as np

from . import abi
from .abi import check


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(abi.fp)


_F3 = C.c_float * 3


def _out_rows(a, cols):
    """A caller's output array: float32, C-contiguous, cols floats per row."""
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous or a.ndim != 2 \
            or a.shape[1] != cols:
        raise ValueError("output array must be a C-contiguous float32 array of shape (rows, %d)" % cols)
    return a


def _out_pair(out, capacity):
    """(pos [c,3], data [c,2]) output arrays and the capacity they hold (c unless a smaller one is asked)."""
    pos, data = _out_rows(out[0], 3), _out_rows(out[1], 2)
    rows = min(pos.shape[0], data.shape[0])
    capacity = rows if capacity is None else int(capacity)
    if capacity > rows:
        raise ValueError("capacity %d exceeds the output arrays (%d rows)" % (capacity, rows))
    return pos, data, capacity


def _vec3(v):
    """A sensor origin as a float[3] argument (a copy, so any float-convertible 3-vector works)."""
    a = np.ascontiguousarray(v, np.float32)
    if a.size != 3:
        raise ValueError("sensor origin must have 3 components, got %d" % a.size)
    return _F3.from_buffer_copy(a)


def params_struct(voxel_size, truncation_distance, max_weight, publish_distance_squared,
                  garbage_collect_distance_squared) -> abi.Params:
    return abi.Params(voxel_size, truncation_distance, max_weight, publish_distance_squared,
                      garbage_collect_distance_squared)


class DeviceBuffer:
    """A block of HBM owned by an engine (stream-ordered copies)."""

    def __init__(self, engine: "Engine", nbytes: int):
        self.engine = engine
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(abi.load().tsdf_device_alloc(engine.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, engine: "Engine", a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(engine, a.nbytes)
        b.upload(a)
        return b

    def _span(self, offset: int, nbytes: int):
        if self.ptr is None:
            raise ValueError("the buffer was freed")
        if offset < 0 or offset + nbytes > self.nbytes:
            raise ValueError("bytes [%d, %d) outside the %d-byte buffer" % (offset, offset + nbytes, self.nbytes))

    def check_points(self, offset_points: int, n: int):
        """Raises unless float4 points [offset_points, offset_points + n) lie inside the buffer."""
        if n < 0:
            raise ValueError("negative point count")
        self._span(16 * int(offset_points), 16 * int(n))

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        self._span(offset, a.nbytes)
        check(abi.load().tsdf_copy_to_device(self.engine.h, C.c_void_p(self.ptr + offset), a.ctypes.data, a.nbytes))

    def download(self, out: np.ndarray, offset: int = 0):
        if not out.flags.c_contiguous:
            raise ValueError("download target must be C-contiguous")
        self._span(offset, out.nbytes)
        check(abi.load().tsdf_copy_to_host(self.engine.h, out.ctypes.data, C.c_void_p(self.ptr + offset), out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.engine.h:
            check(abi.load().tsdf_device_free(self.engine.h, C.c_void_p(self.ptr)))
        self.ptr = None


class Engine:
    """Device block heap + hash-entry pool shared by its maps; one HIP stream."""

    def __init__(self, heap_blocks=1 << 18, total_buckets=1 << 20, state_format=abi.TSDF_STATE_F32,
                 max_batch_points=1 << 17, device=0):
        cfg = abi.EngineConfig(device, state_format, heap_blocks, total_buckets, max_batch_points)
        h = C.c_void_p()
        check(abi.load().tsdf_engine_create(C.byref(cfg), C.byref(h)))
        self.h = h.value
        self.state_format = state_format
        self.heap_blocks = heap_blocks

    def close(self):
        if self.h:
            abi.load().tsdf_engine_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        check(abi.load().tsdf_engine_synchronize(self.h))

    def free_blocks(self) -> int:
        v = C.c_int64()
        check(abi.load().tsdf_engine_free_blocks(self.h, C.byref(v)))
        return v.value

    def map(self, params, buckets_log2=16, overflow_buckets=0, max_chain=64, block_quota=0, admit=False) -> "Map":
        return Map(self, params, buckets_log2, overflow_buckets, max_chain, block_quota, admit)

    def stats(self) -> dict:
        """Counters of the last integration pass (and of the last publish / gc).  A pass that dropped
        updates shows them in insert_failures / heap_failures / updates_dropped (the C ABI's
        TSDF_ERR_DROPPED status, which is not an error of this query)."""
        s = abi.ScanStats()
        r = abi.load().tsdf_last_stats(self.h, C.byref(s))
        if r != abi.TSDF_ERR_DROPPED:
            check(r)
        return s.as_dict()

    def integrate_batch(self, maps, scans, points: DeviceBuffer):
        """scans: list of (map_index, n_points, point_offset, sensor)."""
        for _, n, off, _ in scans:
            points.check_points(off, n)
        arr = (abi.ScanDesc * len(scans))()
        for i, (mi, n, off, s) in enumerate(scans):
            arr[i].map_index = mi
            arr[i].n_points = n
            arr[i].point_offset = off
            arr[i].sensor[:] = [float(x) for x in s]
        mp = (C.c_void_p * len(maps))(*[m.h for m in maps])
        check(abi.load().tsdf_integrate_batch(self.h, mp, len(maps), arr, len(scans), C.c_void_p(points.ptr)))

    def set_launch(self, traverse_block=256, traverse_grid_cap=65536, fold_grid=4096, rays_per_lane=1,
                   traverse_queue=0):
        """Launch shapes of the pass (tsdf_launch_config).  Passes captured in a batch graph keep the shapes of
        their capture."""
        cfg = abi.LaunchConfig(traverse_block, traverse_grid_cap, fold_grid, rays_per_lane, traverse_queue, 0)
        check(abi.load().tsdf_engine_set_launch(self.h, C.byref(cfg)))

    def get_launch(self) -> dict:
        cfg = abi.LaunchConfig()
        check(abi.load().tsdf_engine_get_launch(self.h, C.byref(cfg)))
        return dict(traverse_block=cfg.traverse_block, traverse_grid_cap=cfg.traverse_grid_cap, fold_grid=cfg.fold_grid,
                    rays_per_lane=cfg.rays_per_lane, traverse_queue=cfg.traverse_queue)

    def timing(self, on: bool):
        check(abi.load().tsdf_timing_enable(self.h, 1 if on else 0))

    def timing_read(self, reset=True) -> dict:
        ms = (C.c_double * 8)()
        n = (C.c_int64 * 8)()
        check(abi.load().tsdf_timing_read(self.h, ms, n, 1 if reset else 0))
        return {abi.STAGES[k]: (ms[k], n[k]) for k in range(8)}

    def buffer(self, nbytes) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def host_array(self, shape, dtype=np.float32) -> np.ndarray:
        """A zeroed numpy array in page-locked host memory (tsdf_host_alloc): publish outputs and host points
        kept there cross PCIe without staging copies.  The memory is released with the last reference to the
        array or its views, also after the engine is closed."""
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        p = C.c_void_p()
        check(abi.load().tsdf_host_alloc(self.h, n, C.byref(p)))
        raw = (C.c_ubyte * max(n, 1)).from_address(p.value)
        fin = weakref.finalize(raw, abi.load().tsdf_host_free, p.value)
        fin.atexit = False  # at process exit the memory goes with the process; no HIP call during shutdown
        a = np.frombuffer(raw, dtype=np.uint8, count=n).view(dtype).reshape(shape)
        a[...] = 0
        return a

    def cloud2_to_points(self, data, width, height, point_step, row_step, offsets, out: DeviceBuffer,
                         is_bigendian=False, datatype=abi.TSDF_FIELD_FLOAT32, data_bytes=None) -> int:
        """A sensor_msgs/PointCloud2 data block (host uint8 array or DeviceBuffer) -> float4 points in out.

        offsets: the byte offsets of the x, y, z fields.  Returns the number of points written
        (width * height).  Replaces Transformer::convertPointCloud2ToPointCloudXYZ
        (tsdf_package/src/transformer.cpp:63-67)."""
        on_dev = isinstance(data, DeviceBuffer)
        if on_dev:
            ptr, nbytes = data.ptr, data.nbytes
        else:
            data = np.ascontiguousarray(data, dtype=np.uint8)
            ptr, nbytes = data.ctypes.data, data.nbytes
        lay = abi.Cloud2Layout(width, height, row_step, nbytes if data_bytes is None else data_bytes, point_step,
                               offsets[0], offsets[1], offsets[2], datatype, 1 if is_bigendian else 0)
        n = C.c_int64()
        check(abi.load().tsdf_cloud2_to_points(self.h, C.byref(lay), C.c_void_p(ptr), 1 if on_dev else 0,
                                               C.c_void_p(out.ptr), out.nbytes // 16, C.byref(n)))
        return int(n.value)

    def msg_markers(self, msg):
        """The Tsdf message's debug markers (the consumer of README.md:80-85, debug_viz.cpp's fetchTSDF): msg is
        a float32 [n, 5] array of Voxel.msg records {sdf, weight, x, y, z} on the host, or a DeviceBuffer of n x 20
        bytes.  Returns (positions of the sdf >= 0 voxels, positions of the sdf < 0 voxels), float32 [k, 3] each,
        in message order."""
        on_dev = isinstance(msg, DeviceBuffer)
        if on_dev:
            ptr, n = msg.ptr, msg.nbytes // 20
        else:
            msg = np.ascontiguousarray(msg, dtype=np.float32).reshape(-1, 5)
            ptr, n = msg.ctypes.data, msg.shape[0]
        nonneg = np.empty((max(n, 1), 3), np.float32)
        neg = np.empty((max(n, 1), 3), np.float32)
        a, b = C.c_int64(), C.c_int64()
        check(abi.load().tsdf_msg_markers(self.h, C.c_void_p(ptr), n, 1 if on_dev else 0, _fptr(nonneg), _fptr(neg),
                                          C.byref(a), C.byref(b)))
        return nonneg[:a.value], neg[:b.value]


class Map:
    def __init__(self, engine: Engine, params, buckets_log2=16, overflow_buckets=0, max_chain=64, block_quota=0,
                 admit=False):
        self.engine = engine
        self.params = tuple(float(x) for x in params)
        p = params_struct(*self.params)
        cfg = abi.MapConfig(buckets_log2, overflow_buckets, max_chain, 0, block_quota)
        h = C.c_void_p()
        f = abi.load().tsdf_map_admit if admit else abi.load().tsdf_map_create
        check(f(engine.h, C.byref(p), C.byref(cfg), C.byref(h)))
        self.h = h.value

    def close(self):
        if self.h and self.engine.h:
            check(abi.load().tsdf_map_destroy(self.h))
        self.h = None

    def integrate(self, points: DeviceBuffer, n: int, sensor, offset_points: int = 0):
        if offset_points < 0 or n < 0 or 16 * (offset_points + n) > points.nbytes or points.ptr is None:
            points.check_points(offset_points, n)  # raises with the reason
        check(abi.load().tsdf_integrate(self.h, points.ptr + 16 * offset_points, n, _vec3(sensor)))

    def integrate_noalloc(self, points: DeviceBuffer, n: int, sensor, offset_points: int = 0):
        """integrate() through tsdf_integrate_noalloc: the pass scratch comes from the library's fixed
        per-GPU array, so the call allocates nothing (TsdfError TSDF_ERR_CAPACITY beyond its size)."""
        points.check_points(offset_points, n)
        check(abi.load().tsdf_integrate_noalloc(self.h, points.ptr + 16 * offset_points, n, _vec3(sensor)))

    def publish(self, sensor, capacity=None, out=None):
        """Voxels to publish: (pos float32[k,3], data float32[k,2] {sdf, weight}, total).  k = min(total,
        capacity).  out: (pos, data) float32 arrays of c rows to fill instead of new ones (capacity defaults
        to c), which a caller keeps between scans as the reference's node keeps its publish arrays."""
        s = _vec3(sensor)
        n = C.c_int64()
        if out is not None:
            pos, data, capacity = _out_pair(out, capacity)
        else:
            if capacity is None:
                check(abi.load().tsdf_publish(self.h, s, None, None, 0, C.byref(n)))
                capacity = n.value
            pos = np.zeros((max(capacity, 1), 3), np.float32)
            data = np.zeros((max(capacity, 1), 2), np.float32)
        check(abi.load().tsdf_publish(self.h, s, _fptr(pos), _fptr(data), capacity, C.byref(n)))
        k = min(n.value, capacity)
        return pos[:k], data[:k], n.value

    def publish_msg(self, sensor, capacity=None, out=None):
        """Voxels to publish in the Voxel.msg layout {sdf, weight, x, y, z}: (msg float32[k,5], total).
        out: a float32 array of c x 5 to fill instead of a new one."""
        s = _vec3(sensor)
        n = C.c_int64()
        if out is not None:
            msg = _out_rows(out, 5)
            capacity = msg.shape[0] if capacity is None else int(capacity)
            if capacity > msg.shape[0]:
                raise ValueError("capacity %d exceeds the output array" % capacity)
        else:
            if capacity is None:
                check(abi.load().tsdf_publish_msg(self.h, s, None, 0, C.byref(n)))
                capacity = n.value
            msg = np.zeros((max(capacity, 1), 5), np.float32)
        check(abi.load().tsdf_publish_msg(self.h, s, _fptr(msg), capacity, C.byref(n)))
        return msg[:min(n.value, capacity)], n.value

    def gc(self, sensor) -> int:
        s = _vec3(sensor)
        r = C.c_int64()
        check(abi.load().tsdf_gc(self.h, s, C.byref(r)))
        return r.value

    def num_blocks(self) -> int:
        v = C.c_int64()
        check(abi.load().tsdf_map_num_blocks(self.h, C.byref(v)))
        return v.value

    def export(self):
        """(keys int32[n,3], sdf float32[n,512], weight float32[n,512]) sorted by key."""
        n = C.c_int64()
        check(abi.load().tsdf_map_export(self.h, None, None, 0, C.byref(n)))
        cap = n.value
        keys = np.zeros((max(cap, 1), 3), np.int32)
        vox = np.zeros((max(cap, 1), 512, 2), np.float32)
        check(abi.load().tsdf_map_export(self.h, keys.ctypes.data_as(abi.i32p), _fptr(vox), cap, C.byref(n)))
        keys, vox = keys[:cap], vox[:cap]
        order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
        keys, vox = keys[order], vox[order]
        return keys, vox[:, :, 0].copy(), vox[:, :, 1].copy()

    def import_blocks(self, keys, sdf, weight):
        """Restore blocks in the export layout (keys [n,3], sdf/weight [n,512]); returns (stored, refused)."""
        keys = np.ascontiguousarray(keys, np.int32).reshape(-1, 3)
        vox = np.ascontiguousarray(np.stack([np.asarray(sdf, np.float32), np.asarray(weight, np.float32)], axis=-1))
        assert vox.shape == (keys.shape[0], 512, 2)
        a, b = C.c_int64(), C.c_int64()
        check(abi.load().tsdf_map_import(self.h, keys.ctypes.data_as(abi.i32p), _fptr(vox), keys.shape[0],
                                         C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def process_scan(self, points: np.ndarray, sensor, capacity=None, out=None):
        """The reference's processPointCloudAndUpdateVoxels sequence with host arrays: integrate the
        float4 points, publish into (pos, data), gc.  out: (pos, data) arrays kept by the caller, as in
        publish; otherwise new arrays of capacity rows (default 2^20)."""
        pts = np.ascontiguousarray(points, np.float32)
        if pts.ndim != 2 or pts.shape[1] != 4:
            raise ValueError("points must be float4 rows (x, y, z, pad) of shape (n, 4), got %s" % (pts.shape,))
        s = _vec3(sensor)
        n = C.c_int64()
        if out is not None:
            pos, data, capacity = _out_pair(out, capacity)
        else:
            if capacity is None:
                capacity = 1 << 20
            pos = np.zeros((capacity, 3), np.float32)
            data = np.zeros((capacity, 2), np.float32)
        r = abi.load().tsdf_process_scan(self.h, _fptr(pts), pts.shape[0], s, _fptr(pos), _fptr(data),
                                         capacity, C.byref(n))
        if r == abi.TSDF_ERR_DROPPED:
            # the scan ran to the end and the outputs are complete; the map lost the refused blocks' updates
            warnings.warn(abi.load().tsdf_last_error().decode(), abi.DroppedUpdatesWarning, stacklevel=2)
        else:
            check(r)
        k = min(n.value, capacity)
        return pos[:k], data[:k], n.value


def derived_params(params):
    p = params_struct(*params)
    of = (C.c_float * 5)()
    oi = (C.c_int32 * 2)()
    check(abi.load().tsdf_derived_params(C.byref(p), of, oi))
    return dict(half_vs=of[0], BL=of[1], half_BL=of[2], blk_eps=of[3], vox_eps=of[4], max_vox=oi[0], max_blk=oi[1])


def _scan_array(scans):
    arr = (abi.ScanDesc * max(len(scans), 1))()
    for i, (mi, n, off, s) in enumerate(scans):
        arr[i].map_index = int(mi)
        arr[i].n_points = int(n)
        arr[i].point_offset = int(off)
        arr[i].sensor[:] = [float(x) for x in s]
    return arr


class BatchGraph:
    """The batched pipeline captured in a HIP graph for a fixed set of maps."""

    def __init__(self, engine: Engine, maps, max_points: int, max_scans: int):
        self.engine = engine
        self.maps = list(maps)
        mp = (C.c_void_p * len(self.maps))(*[m.h for m in self.maps])
        h = C.c_void_p()
        check(abi.load().tsdf_batch_graph_create(engine.h, mp, len(self.maps), int(max_points), int(max_scans),
                                                 C.byref(h)))
        self.h = h.value

    def launch(self, scans, points: DeviceBuffer):
        """scans: (map_index, n_points, point_offset, sensor); contiguous and in order."""
        for _, n, off, _ in scans:
            points.check_points(off, n)
        arr = _scan_array(scans)
        check(abi.load().tsdf_batch_graph_launch(self.h, arr, len(scans), C.c_void_p(points.ptr)))

    def close(self):
        if self.h:
            abi.load().tsdf_batch_graph_destroy(self.h)
            self.h = None


def serve(engine: Engine, maps, requests, points: DeviceBuffer, max_batch_scans=32, max_inflight=2,
          max_batch_points=1 << 22, deadline_ms=0.0):
    """Runs the continuous-batching loop.  requests: (arrival_s, map_index,
    n_points, point_offset, sensor), sorted by arrival.  Returns
    (latency_ms array, batch index array, stats dict)."""
    n = len(requests)
    if isinstance(points, DeviceBuffer):
        for _, _, npts, off, _ in requests:
            points.check_points(off, npts)
    arr = (abi.Request * max(n, 1))()
    for i, (t, mi, npts, off, s) in enumerate(requests):
        arr[i].arrival_s = float(t)
        arr[i].map_index = int(mi)
        arr[i].n_points = int(npts)
        arr[i].point_offset = int(off)
        arr[i].sensor[:] = [float(x) for x in s]
    cfg = abi.ServeConfig(int(max_batch_scans), int(max_inflight), int(max_batch_points), float(deadline_ms))
    lat = np.zeros(max(n, 1), np.float64)
    bof = np.zeros(max(n, 1), np.int32)
    st = abi.ServeStats()
    mp = (C.c_void_p * len(maps))(*[m.h for m in maps])
    check(abi.load().tsdf_serve_run(engine.h, mp, len(maps), arr, n, C.c_void_p(points.ptr), C.byref(cfg),
                                     lat.ctypes.data_as(C.POINTER(C.c_double)), bof.ctypes.data_as(abi.i32p),
                                     C.byref(st)))
    stats = {f: getattr(st, f) for f, _ in abi.ServeStats._fields_}
    return lat[:n], bof[:n], stats
