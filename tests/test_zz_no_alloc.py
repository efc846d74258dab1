"""tsdf_integrate_noalloc (pass scratch in a fixed per-GPU device array, no allocation) against
tsdf_integrate, the allocating entry point whose kernels it shares: two calls in a row, a full-size
dense scan and then a smaller one, give bit-identical maps and statistics; a call beyond the array's
size is refused with nothing enqueued."""
import ctypes as C

import numpy as np
import pytest

from ros_tsdf_amd import abi, configs, synth

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return all(np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y) for x, y in zip(a, b))


def test_noalloc_equals_integrate_then_refuses_beyond_its_scratch():
    from ros_tsdf_amd import Engine
    P = configs.DENSE_PARAMS
    # a small initial scratch: the allocating path grows it for the full-size scan, the other never allocates
    with Engine(heap_blocks=1 << 18, total_buckets=1 << 19, max_batch_points=1 << 10) as e:
        a = e.map(P, buckets_log2=17)
        b = e.map(P, buckets_log2=17)
        scans = [synth.make_scan(configs.DENSE, 0)[:2], synth.make_scan(configs.DENSE, 1, n_points=20000)[:2]]
        assert scans[0][0].shape[0] == 131072 and scans[1][0].shape[0] == 20000
        buf = e.buffer(scans[0][0].nbytes)
        for pts, s in scans:
            buf.upload(pts)
            a.integrate(buf, pts.shape[0], s)
            st_a = e.stats()
            b.integrate_noalloc(buf, pts.shape[0], s)
            st_b = e.stats()
            assert st_a == st_b and st_b["voxel_updates"] > 0 and st_b["updates_dropped"] == 0
            assert same_bits(a.export(), b.export())
        # 200,000 points at 32 slots need more than the 96 MiB array: refused before any launch (the
        # buffer holds all of them, so nothing could read past it either way)
        n_big = 200000
        big = e.buffer(16 * n_big)
        big.upload(np.zeros((n_big, 4), np.float32))
        L = abi.load()
        sensor = (C.c_float * 3)(0.0, 0.0, 0.0)
        before = b.export()
        r = L.tsdf_integrate_noalloc(C.c_void_p(b.h), C.c_void_p(big.ptr), n_big, sensor)
        assert r == abi.TSDF_ERR_CAPACITY
        assert b"pass scratch" in L.tsdf_last_error()
        e.synchronize()
        assert same_bits(b.export(), before)
        big.free()
        buf.free()
        a.close()
        b.close()
