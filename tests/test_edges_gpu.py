"""GPU edge cases, each checked against the FP64 oracle or against the engine's own per-scan path.

Inputs: empty and ragged scans, invalid points (NaN, infinity, zero range, beyond the coordinate limit),
points on exact cell and block boundaries, and coordinates near the validity limit.

Capacity cases: heap exhaustion, a chain limit that makes inserts fail, publish capacity truncation, the
Voxel.msg layout, map admission against a heap budget, per-map block quotas, a scan larger than the
initial scratch, and more distinct blocks per binning workgroup than its LDS table holds.  Fold tiers: one
block with exactly n records, n on both sides of every size-class bound.

When blocks cannot be stored (heap, chain or quota exhaustion), the whole block's updates are dropped. Every
block that is stored received all of its updates, so it matches the oracle's block bit-exactly on keys and
within tolerance on values.
"""
import numpy as np
import pytest

from ros_tsdf_amd import abi, configs, synth, workloads
from tests.helpers import compare_maps, compare_publish

pytestmark = pytest.mark.gpu

P = configs.SMALL_PARAMS      # voxel 0.5 m, truncation 0.1 m
D = configs.DENSE_PARAMS      # voxel 0.1 m, truncation 0.4 m


@pytest.fixture(scope="module")
def eng():
    from ros_tsdf_amd import Engine
    e = Engine(heap_blocks=1 << 16, total_buckets=1 << 18, max_batch_points=1 << 16)
    yield e
    e.close()


def tol(params):
    return configs.SDF_TOL_T, configs.WEIGHT_TOL_REL * params[2]


def small_scan(k=0, mi=0, n=None):
    w = configs.map_workload(configs.SMALL, mi)
    p, s, _ = synth.make_scan(w, k, map_id=mi, n_points=n)
    return p, s


def dense_scan(k=0, mi=0, n=None):
    w = configs.map_workload(configs.DENSE, mi)
    p, s, _ = synth.make_scan(w, k, map_id=mi, n_points=n)
    return p, s


def as_points(xyz):
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    out = np.zeros((xyz.shape[0], 4), np.float32)
    out[:, :3] = xyz
    return out


def integrate_both(eng, oracle_lib, params, scans, **map_kw):
    m = eng.map(params, **map_kw)
    om = oracle_lib.OracleMap(params)
    stats = []
    for pts, s in scans:
        buf = eng.buffer(max(pts.nbytes, 16))
        if pts.shape[0]:
            buf.upload(pts)
        m.integrate(buf, pts.shape[0], s)
        stats.append((eng.stats(), om.integrate(pts, s)))
        buf.free()
    return m, om, stats


def compare_stored_blocks(gmap, omap, T, sdf_tol_T, w_tol):
    """Every stored block exists in the oracle and matches it; returns (stored, oracle) block counts."""
    gk, gs, gw = gmap.export()
    ok, os_, ow = omap.export()
    idx = {tuple(k): i for i, k in enumerate(ok.tolist())}
    for j, k in enumerate(gk.tolist()):
        i = idx.get(tuple(k))
        assert i is not None, "block %s is not in the oracle map" % (k,)
        t = ow[i] != 0
        assert np.array_equal(gw[j] != 0, t), "touched voxels of block %s differ" % (k,)
        if t.any():
            assert np.abs(gs[j][t].astype(np.float64) - os_[i][t]).max() / T <= sdf_tol_T
            assert np.abs(gw[j][t].astype(np.float64) - ow[i][t]).max() <= w_tol
    return gk.shape[0], ok.shape[0]


def test_empty_scan_and_empty_batch(eng):
    m = eng.map(P)
    buf = eng.buffer(16)
    m.integrate(buf, 0, (1.0, 2.0, 0.0))
    st = eng.stats()
    assert st["points"] == 0 and st["valid"] == 0 and st["voxel_updates"] == 0
    assert m.num_blocks() == 0
    pos, data, n = m.publish((1.0, 2.0, 0.0))
    assert n == 0 and pos.shape[0] == 0
    assert m.gc((1.0, 2.0, 0.0)) == 0
    eng.integrate_batch([m], [], buf)
    eng.integrate_batch([m], [(0, 0, 0, (0.0, 0.0, 0.0)), (0, 0, 0, (1.0, 0.0, 0.0))], buf)
    eng.synchronize()
    assert m.num_blocks() == 0
    m.close()
    buf.free()


@pytest.mark.parametrize("fmt", [abi.TSDF_STATE_F32, abi.TSDF_STATE_F16, abi.TSDF_STATE_F16_QW])
def test_publish_and_gc_radius_boundaries(oracle_lib, fmt):
    """Blocks centred exactly 4, 8 and 12 m from the sensor with publish and GC radius 8 m: the block on
    the radius is published (withinDistanceSquared is <=, H:750-757, H:812) and kept by GC (H:847); the
    engine's sets equal the oracle's."""
    from ros_tsdf_amd import Engine
    from tests.helpers import RADIUS_PARAMS, RADIUS_POINTS, RADIUS_SENSOR
    with Engine(heap_blocks=1 << 10, total_buckets=1 << 10, max_batch_points=1 << 10, state_format=fmt) as e:
        m, om, _ = integrate_both(e, oracle_lib, RADIUS_PARAMS, [(RADIUS_POINTS, RADIUS_SENSOR)], buckets_log2=6)
        k, _, _ = m.export()
        assert sorted(map(tuple, k.tolist())) == [(1, 0, 0), (2, 0, 0), (3, 0, 0)]
        pos, _, n = m.publish(RADIUS_SENSOR)
        opos, _, _ = om.publish(RADIUS_SENSOR)
        blocks = sorted({tuple(int(c) for c in np.floor(p / 4.0)) for p in pos.astype(np.float64)})
        assert blocks == [(1, 0, 0), (2, 0, 0)]
        rg, ro = pos.view(np.uint32), np.asarray(opos, np.float32).view(np.uint32)
        assert np.array_equal(rg[np.lexsort(rg.T[::-1])], ro[np.lexsort(ro.T[::-1])])
        assert m.gc(RADIUS_SENSOR) == om.gc(RADIUS_SENSOR) == 1
        k, _, _ = m.export()
        assert sorted(map(tuple, k.tolist())) == [(1, 0, 0), (2, 0, 0)]
        m.close()
        om.close()


def test_misaligned_point_buffer_is_refused(eng):
    """Points are float4 rows: a device pointer off a 16-byte boundary is refused before anything runs."""
    import ctypes as C
    pts, s = small_scan(0, n=64)
    m = eng.map(P)
    buf = eng.buffer(pts.nbytes + 16)
    buf.upload(pts)
    sensor = (C.c_float * 3)(*[float(x) for x in s])
    for off in (4, 8, 12):
        assert abi.load().tsdf_integrate(m.h, C.c_void_p(buf.ptr + off), pts.shape[0], sensor) == \
            abi.TSDF_ERR_INVALID_ARG
    assert m.num_blocks() == 0
    m.integrate(buf, pts.shape[0], s)  # the aligned buffer integrates
    eng.synchronize()
    assert m.num_blocks() > 0
    m.close()
    buf.free()


def test_point_ranges_outside_the_buffer_are_refused(eng):
    """The Python API knows each device buffer's size: a scan, batch, graph or serving request whose points
    reach past it (or start before it) is refused before any device work, as are out-of-range copies."""
    from ros_tsdf_amd import BatchGraph, serve
    pts, s = small_scan(0, n=64)
    m = eng.map(P)
    buf = eng.buffer(pts.nbytes)
    buf.upload(pts)
    bad = [(0, 65), (1, 64), (-1, 10)]  # (offset, n): one past the end, shifted, negative
    for off, n in bad:
        with pytest.raises(ValueError):
            m.integrate(buf, n, s, offset_points=off)
        with pytest.raises(ValueError):
            eng.integrate_batch([m], [(0, n, off, s)], buf)
        with pytest.raises(ValueError):
            serve(eng, [m], [(0.0, 0, n, off, s)], buf)
    g = BatchGraph(eng, [m], max_points=128, max_scans=1)
    with pytest.raises(ValueError):
        g.launch([(0, 65, 0, s)], buf)
    g.close()
    with pytest.raises(ValueError):
        buf.upload(np.zeros(pts.nbytes + 1, np.uint8))
    with pytest.raises(ValueError):
        buf.download(np.zeros(16, np.uint8), offset=pts.nbytes - 8)
    assert m.num_blocks() == 0
    m.integrate(buf, 64, s)  # the whole buffer, exactly
    eng.synchronize()
    assert m.num_blocks() > 0
    m.close()
    buf.free()


def test_invalid_points_skipped_like_the_oracle(eng, oracle_lib):
    pts, s = small_scan(0)
    lim = P[0] * 65536.0
    bad = as_points([[np.nan, 1.0, 0.0], [1.0, np.inf, 0.0], [0.0, 0.0, -np.inf],
                     [s[0], s[1], s[2]],                       # zero range
                     [lim * 1.5, 0.0, 0.0], [0.0, -lim * 2.0, 1.0]])  # beyond the coordinate limit
    mixed = np.concatenate([pts[:300], bad, pts[300:]])
    m, om, stats = integrate_both(eng, oracle_lib, P, [(mixed, s)])
    g, o = stats[0]
    assert g["skipped"] == o["n_skipped"] == 6
    assert g["valid"] == o["n_valid"] == pts.shape[0]
    assert g["voxel_updates"] == o["n_voxel_updates"]
    compare_maps(m, om, P[1], *tol(P))
    m.close()
    om.close()


def test_points_on_cell_and_block_boundaries(eng, oracle_lib):
    # exact multiples of the voxel and block sizes on every axis, sensor on a
    # block corner: every DDA starts, ends or steps on a boundary and ties
    vs = D[0]
    g = np.arange(-24, 25, dtype=np.float32)
    xs, ys = np.meshgrid(g * np.float32(vs), g * np.float32(8 * vs))
    zs = np.full_like(xs, np.float32(-2.0))
    pts = as_points(np.stack([xs.ravel(), ys.ravel(), zs.ravel()], 1))
    s = np.array([0.0, 0.0, 0.0], np.float32)
    s2 = np.array([np.float32(8 * vs) * 3, np.float32(vs) * 5, np.float32(0.8)], np.float32)
    m, om, stats = integrate_both(eng, oracle_lib, D, [(pts, s), (pts, s2)])
    for g_, o_ in stats:
        assert g_["voxel_updates"] == o_["n_voxel_updates"] and g_["unique_voxels"] == o_["n_unique_voxels"]
    compare_maps(m, om, D[1], *tol(D))
    m.close()
    om.close()


def test_coordinates_near_the_validity_limit(eng, oracle_lib):
    # the scene shifted to 0.9 x 2^16 voxels from the origin (29.5 km at 0.5 m)
    pts, s = small_scan(1)
    off = np.float32(0.9 * 65536 * P[0])
    p2 = pts.copy()
    p2[:, 0] += off
    p2[:, 1] -= off
    s2 = s.copy()
    s2[0] += off
    s2[1] -= off
    m, om, stats = integrate_both(eng, oracle_lib, P, [(p2, s2)])
    assert stats[0][0]["valid"] == stats[0][1]["n_valid"] == pts.shape[0]
    compare_maps(m, om, P[1], *tol(P))
    m.close()
    om.close()


def test_heap_exhaustion_keeps_complete_blocks(oracle_lib):
    from ros_tsdf_amd import Engine
    with Engine(heap_blocks=64, total_buckets=1 << 12, max_batch_points=1 << 14) as e:
        pts, s = dense_scan(0, n=4096)
        m, om, stats = integrate_both(e, oracle_lib, D, [(pts, s)], buckets_log2=8)
        g, o = stats[0]
        assert g["heap_failures"] > 0 and g["updates_dropped"] > 0
        assert g["new_blocks"] == 64 and m.num_blocks() == 64 and e.free_blocks() == 0
        stored, total = compare_stored_blocks(m, om, D[1], *tol(D))
        assert stored == 64 < total
        # gc from far away returns every block; the heap serves the next scan again
        far = (1e4, 1e4, 0.0)
        assert m.gc(far) == 64 and e.free_blocks() == 64
        pts2, s2 = dense_scan(1, n=256)
        buf = e.buffer(pts2.nbytes)
        buf.upload(pts2)
        m.integrate(buf, pts2.shape[0], s2)
        st = e.stats()
        assert st["heap_failures"] == 0 and st["new_blocks"] == m.num_blocks() > 0
        buf.free()
        m.close()
        om.close()


def test_chain_limit_makes_inserts_fail_cleanly(eng, oracle_lib):
    # 2 home buckets, 2 overflow buckets, chains of at most one overflow bucket: 32 entries in all
    pts, s = dense_scan(2, n=2048)
    m, om, stats = integrate_both(eng, oracle_lib, D, [(pts, s)], buckets_log2=1, overflow_buckets=2, max_chain=1)
    g, o = stats[0]
    assert g["insert_failures"] > 0 and g["updates_dropped"] > 0
    stored, total = compare_stored_blocks(m, om, D[1], *tol(D))
    assert 0 < stored <= 32 < total
    m.close()
    om.close()


def test_publish_capacity_and_msg_layout(eng, oracle_lib):
    pts, s = dense_scan(3, n=8192)
    m, om, _ = integrate_both(eng, oracle_lib, D, [(pts, s)])
    pos, data, n = m.publish(s)
    opos, osdf, ow = om.publish(s)
    compare_publish(pos, data, opos, osdf, ow, D[1], *tol(D))
    cap = n // 3
    pos3, data3, n3 = m.publish(s, capacity=cap)
    assert n3 == n and pos3.shape[0] == cap
    full = {tuple(r) for r in np.concatenate([pos, data], 1).view(np.uint32).tolist()}
    part = {tuple(r) for r in np.concatenate([pos3, data3], 1).view(np.uint32).tolist()}
    assert part <= full and len(part) == cap
    msg, nm = m.publish_msg(s)
    assert nm == n
    # Voxel.msg field order: sdf, weight, x, y, z
    as_msg = np.concatenate([data, pos], 1)
    a = np.lexsort(as_msg.T[::-1])
    b = np.lexsort(msg.T[::-1])
    assert np.array_equal(as_msg[a].view(np.uint32), msg[b].view(np.uint32))
    # caller-owned output arrays (kept between calls): the same voxels; a smaller capacity fills a prefix
    out = (np.full((n + 7, 3), -1.0, np.float32), np.full((n + 7, 2), -1.0, np.float32))
    def rows(*cols):  # a publish result as a sorted array of bit patterns (the voxel order is free)
        r = np.concatenate(cols, 1).view(np.uint32)
        return r[np.lexsort(r.T[::-1])]
    pos_o, data_o, n_o = m.publish(s, out=out)
    assert n_o == n and pos_o.shape[0] == n and np.shares_memory(pos_o, out[0])
    assert np.array_equal(rows(pos_o, data_o), rows(pos, data))
    assert (out[0][n:] == -1.0).all() and (out[1][n:] == -1.0).all()
    pos_c, _, n_c = m.publish(s, capacity=cap, out=out)
    assert n_c == n and pos_c.shape[0] == cap
    with pytest.raises(ValueError):
        m.publish(s, capacity=n + 8, out=out)
    msg_o, nm_o = m.publish_msg(s, out=np.zeros((n, 5), np.float32))
    assert nm_o == n and np.array_equal(rows(msg_o), rows(msg))
    m.close()
    om.close()


def test_map_admission_budget_and_quota(oracle_lib):
    from ros_tsdf_amd import Engine
    with Engine(heap_blocks=1000, total_buckets=1 << 14, max_batch_points=1 << 14) as e:
        a = e.map(D, buckets_log2=8, block_quota=600, admit=True)
        with pytest.raises(abi.TsdfError) as ei:
            e.map(D, buckets_log2=8, block_quota=600, admit=True)
        assert ei.value.code == abi.TSDF_ERR_BUDGET
        b = e.map(D, buckets_log2=8, block_quota=300, admit=True)
        q = e.map(D, buckets_log2=8, block_quota=10)
        pts, s = dense_scan(4, n=4096)
        buf = e.buffer(pts.nbytes)
        buf.upload(pts)
        q.integrate(buf, pts.shape[0], s)
        st = e.stats()
        assert st["quota_failures"] > 0 and q.num_blocks() == 10
        om = oracle_lib.OracleMap(D)
        om.integrate(pts, s)
        stored, total = compare_stored_blocks(q, om, D[1], *tol(D))
        assert stored == 10 < total
        om.close()
        buf.free()
        for mm in (a, b, q):
            mm.close()


def test_ragged_batch_equals_per_scan(eng):
    sizes = [0, 1, 7, 1000, 0, 333, 64, 65]
    scans = []
    for j, n in enumerate(sizes):
        p, s = small_scan(j, mi=j % 3, n=max(n, 1))
        scans.append((p[:n], s, j % 3))
    pool = workloads.Pool(scans)
    buf = eng.buffer(max(pool.points.nbytes, 16))
    if pool.points.shape[0]:
        buf.upload(pool.points)
    mb = [eng.map(P, buckets_log2=10) for _ in range(3)]
    ms = [eng.map(P, buckets_log2=10) for _ in range(3)]
    descs = [(pool.maps[i], pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]
    eng.integrate_batch(mb, descs, buf)
    for (mi, n, off, s) in descs:
        ms[mi].integrate(buf, n, s, offset_points=off)
    eng.synchronize()
    for a, b in zip(mb, ms):
        ka, sa, wa = a.export()
        kb, sb, wb = b.export()
        assert np.array_equal(ka, kb)
        assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))
    for mm in mb + ms:
        mm.close()
    buf.free()


def test_scan_larger_than_initial_scratch(oracle_lib):
    from ros_tsdf_amd import Engine
    with Engine(heap_blocks=1 << 14, total_buckets=1 << 14, max_batch_points=1024) as e:
        pts, s = dense_scan(5, n=20000)
        m, om, _ = integrate_both(e, oracle_lib, D, [(pts, s)], buckets_log2=10)
        compare_maps(m, om, D[1], *tol(D))
        m.close()
        om.close()


@pytest.mark.parametrize("fmt", [abi.TSDF_STATE_F16, abi.TSDF_STATE_F16_QW], ids=["fp16", "fp16_fixed_weight"])
def test_packed_state_publish_and_gc_sequence(oracle_lib, fmt):
    """The reference call sequence (integrate -> publish -> gc, H:1003-1007) with packed voxel state: the
    published and removed sets are exact, the values within the packed formats' tolerance (1e-2 T)."""
    from ros_tsdf_amd import Engine
    with Engine(heap_blocks=1 << 15, total_buckets=1 << 16, max_batch_points=1 << 14, state_format=fmt) as e:
        m = e.map(D, buckets_log2=12)
        om = oracle_lib.OracleMap(D)
        for k in range(3):
            pts, s = dense_scan(k, n=4096)
            pos, data, n = m.process_scan(pts, s)
            om.integrate(pts, s)
            opos, osdf, ow = om.publish(s)
            assert n == opos.shape[0]
            compare_publish(pos, data, opos, osdf, ow, D[1], 1e-2, 2e-2 * D[2])
            om.gc(s)
        k_g, _, w_g = m.export()
        k_o, _, w_o = om.export()
        assert np.array_equal(k_g, k_o) and np.array_equal(w_g != 0, w_o != 0)
        m.close()
        om.close()


def test_pinned_host_arrays():
    """Page-locked host arrays (Engine.host_array) as process_scan input and publish outputs give exactly
    the voxels of pageable arrays; the arrays stay usable, and are released, after the engine is closed."""
    import gc
    from ros_tsdf_amd import Engine
    pts, s = dense_scan(4, n=8192)
    with Engine(heap_blocks=1 << 15, total_buckets=1 << 16, max_batch_points=1 << 14) as e:
        a = e.map(D, buckets_log2=12)
        b = e.map(D, buckets_log2=12)
        pin_pts = e.host_array(pts.shape)
        pin_pts[...] = pts
        out = (e.host_array((1 << 18, 3)), e.host_array((1 << 18, 2)))
        assert out[0].dtype == np.float32 and not out[0].any()
        pos_a, data_a, n_a = a.process_scan(pin_pts, s, out=out)
        pos_b, data_b, n_b = b.process_scan(pts, s)
        assert n_a == n_b > 0 and np.shares_memory(pos_a, out[0])
        ra = np.concatenate([pos_a, data_a], 1).view(np.uint32)
        rb = np.concatenate([pos_b, data_b], 1).view(np.uint32)
        assert np.array_equal(ra[np.lexsort(ra.T[::-1])], rb[np.lexsort(rb.T[::-1])])
        a.close()
        b.close()
    keep = out[0][:n_a].copy()
    assert np.array_equal(keep.view(np.uint32), pos_a.view(np.uint32))
    del pin_pts, out, pos_a, data_a
    gc.collect()


def test_workgroup_block_table_overflow(oracle_lib):
    """More distinct blocks per 256-point workgroup than the binning kernel's LDS table holds (1024 slots):
    2048 points scattered over a 1 km cube, a 1.2 m truncation band (up to 12 blocks and 96 voxels per point).
    Every 256 consecutive points touch about 1440 distinct blocks (the oracle's walks check this), so the table
    fills and the remaining blocks take the direct global path, more than 1024 of them are new (the
    workgroup's new-entry list overflows to the pass list), and points have more runs than the emit kernel
    keeps in registers.  The map equals the oracle's."""
    import oracle
    from ros_tsdf_amd import Engine
    params = (0.1, 1.2, 1000.0, 2500.0, 22500.0)
    rng = np.random.default_rng(11)
    xyz = rng.uniform(-500, 500, size=(2048, 3)).astype(np.float32)
    s = np.zeros(3, np.float32)
    most_runs = 0
    for c in range(0, xyz.shape[0], 256):
        blocks = set()
        for i in range(c, c + 256):
            a, b = oracle.segment(xyz[i], s, params[1])
            keys, _, _ = oracle.traverse(a, b, params[0], 96)
            bk = [tuple(k) for k in (keys >> 3)]
            blocks.update(bk)
            most_runs = max(most_runs, 1 + sum(bk[j] != bk[j - 1] for j in range(1, len(bk))))
        assert len(blocks) > 1024 + 256
    assert most_runs > 6
    with Engine(heap_blocks=1 << 15, total_buckets=1 << 16, max_batch_points=1 << 12) as e:
        m, om, stats = integrate_both(e, oracle_lib, params, [(as_points(xyz), s)], buckets_log2=14)
        g, o = stats[0]
        assert g["insert_failures"] == 0 and g["heap_failures"] == 0 and g["updates_dropped"] == 0
        assert g["voxel_updates"] == o["n_voxel_updates"] and g["unique_voxels"] == o["n_unique_voxels"]
        assert g["new_blocks"] == m.num_blocks() > 8 * 1024
        compare_maps(m, om, params[1], *tol(params))
        m.close()
        om.close()


def test_process_scan_input_buffer_grows():
    """processPointCloudAndUpdateVoxels with host points of growing and shrinking size (1000, then 300000,
    past the engine's first input buffer, then 500), publishing into output arrays kept between calls:
    every call gives exactly the published set of the explicit sequence upload -> integrate -> publish ->
    gc on a twin map (whose publish allocates its arrays)."""
    from ros_tsdf_amd import Engine
    with Engine(heap_blocks=1 << 17, total_buckets=1 << 17, max_batch_points=1 << 19) as e:
        a = e.map(D, buckets_log2=14)
        b = e.map(D, buckets_log2=14)
        buf = e.buffer(300000 * 16)
        out = (np.zeros((1 << 22, 3), np.float32), np.zeros((1 << 22, 2), np.float32))  # kept between calls
        for k, n in enumerate((1000, 300000, 500)):
            pts, s = dense_scan(k, n=n)
            assert pts.shape[0] == n
            pos_a, data_a, n_a = a.process_scan(pts, s, out=out)
            buf.upload(np.ascontiguousarray(pts, np.float32))
            b.integrate(buf, n, s)
            pos_b, data_b, n_b = b.publish(s)
            b.gc(s)
            assert n_a == n_b > 0
            oa, ob = np.lexsort(pos_a.T[::-1]), np.lexsort(pos_b.T[::-1])
            assert np.array_equal(pos_a[oa].view(np.uint32), pos_b[ob].view(np.uint32))
            assert np.array_equal(data_a[oa].view(np.uint32), data_b[ob].view(np.uint32))
        assert a.num_blocks() == b.num_blocks()
        buf.free()
        a.close()
        b.close()


# Fold tiers at their bounds: one block gets exactly n records in one pass, for n on both sides of every
# size-class bound (16/17, 32/33, 64/65 wave tiers; 128/129 wave vs workgroup; 1024/1025 workgroup tiles;
# 4096/4097 counting tile vs voxel ranges; 8192/8193 and 20000 several ranges; a voxel beyond a range).  Axis-aligned rays make the
# count exact: with voxel 0.1 m and truncation 0.25 m, a point at x in (8.36, 8.44) updates voxels 81..86
# (six records, all in block x = 10), and a point at x = 7.8 + 0.1 (r - 1) updates r voxels of block 10
# (the rest of its run lands in block 9).  The points sit in up to 64 (y, z) voxel columns, one scan per
# column (each scan's sensor on the column's axis), and the scans go through one batched pass.
TIER_PARAMS = (0.1, 0.25, 1000.0, 2500.0, 22500.0)


def tier_scans(n, cols=64):
    q, r = divmod(n, 6)
    per = [q // cols + (1 if c < q % cols else 0) for c in range(cols)]
    scans = []
    for c in range(cols):
        y = np.float32(0.1 * (c % 8) + 0.037)
        z = np.float32(0.1 * (c // 8) + 0.061)
        xs = [8.36 + 0.08 * (u + 0.5) / max(per[c], 1) for u in range(per[c])]
        if c == 0 and r:
            xs.append(7.8 + 0.1 * (r - 1))
        if not xs:
            continue
        pts = as_points([[x, y, z] for x in xs])
        scans.append((pts, np.array([0.0, y, z], np.float32)))
    return scans


@pytest.mark.parametrize("n,cols", [(n, 64) for n in (16, 17, 32, 33, 64, 65, 128, 129, 1024, 1025, 4096, 4097,
                                                      8192, 8193, 20000)]
                         + [(1025, 1), (3000, 1), (4000, 1), (4096, 2), (12606, 1)])
def test_fold_tiers_at_their_bounds(eng, oracle_lib, n, cols):
    # (1025, 1) .. (4096, 2) put 171..667 records in each of a few voxels of a 4096-record counting tile.
    # Blocks beyond 4096 records are folded in voxel ranges of about 2048 records; (12606, 1) puts 2101
    # records in each of six voxels, past the range bound, which takes the merge path instead
    scans = tier_scans(n, cols)
    pool = workloads.Pool([(p, s, 0) for p, s in scans])
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    m = eng.map(TIER_PARAMS, buckets_log2=10)
    eng.integrate_batch([m], [(0, pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))], buf)
    eng.synchronize()
    st = eng.stats()
    om = oracle_lib.OracleMap(TIER_PARAMS)
    for p, s in scans:
        om.integrate(p, s)
    # every point's run has six voxels; only the straddling point's run leaves block 10
    assert st["voxel_updates"] == 6 * pool.points.shape[0]
    k, _, w = m.export()
    i10 = [j for j, kk in enumerate(k.tolist()) if kk == [10, 0, 0]]
    assert len(i10) == 1
    compare_maps(m, om, TIER_PARAMS[1], *tol(TIER_PARAMS))
    m.close()
    buf.free()


def test_block_beyond_every_range_folds_in_point_rounds(eng, oracle_lib):
    """270,000 records in one block in one pass, 2,812 in each of its 96 updated voxels: no voxel range fits
    (a voxel holds more than 2,048 records), so the block is folded in rounds of consecutive points.  Nothing
    is dropped and every voxel equals the oracle's."""
    scans = tier_scans(270000, cols=16)
    pool = workloads.Pool([(p, s, 0) for p, s in scans])
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    m = eng.map(TIER_PARAMS, buckets_log2=10)
    eng.integrate_batch([m], [(0, pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))], buf)
    eng.synchronize()
    st = eng.stats()
    assert st["voxel_updates"] == 6 * pool.points.shape[0] == 270000
    assert st["updates_dropped"] == 0 and st["unique_voxels"] == 6 * 16
    om = oracle_lib.OracleMap(TIER_PARAMS)
    for p, s in scans:
        om.integrate(p, s)
    k, _, w = m.export()
    i10 = [j for j, kk in enumerate(k.tolist()) if kk == [10, 0, 0]]
    assert len(i10) == 1 and (w[i10[0]] != 0).sum() == 96
    compare_maps(m, om, TIER_PARAMS[1], *tol(TIER_PARAMS))
    m.close()
    om.close()
    buf.free()


def last_stats_status(e):
    import ctypes as C
    st = abi.ScanStats()
    return abi.load().tsdf_last_stats(C.c_void_p(e.h), C.byref(st)), st.as_dict()


def test_dropped_updates_are_a_status(oracle_lib):
    """Blocks refused for lack of heap make tsdf_last_stats and tsdf_process_scan return TSDF_ERR_DROPPED
    (with the counts filled in and the outputs complete); a pass without refusals returns TSDF_OK."""
    from ros_tsdf_amd import Engine
    with Engine(heap_blocks=64, total_buckets=1 << 12, max_batch_points=1 << 14) as e:
        pts, s = dense_scan(0, n=4096)
        m = e.map(D, buckets_log2=8)
        buf = e.buffer(pts.nbytes)
        buf.upload(pts)
        m.integrate(buf, pts.shape[0], s)
        code, st = last_stats_status(e)
        assert code == abi.TSDF_ERR_DROPPED and st["heap_failures"] > 0 and st["updates_dropped"] > 0
        assert "dropped" in abi.load().tsdf_last_error().decode()
        # the full per-scan call on a second map: the status, and a publish that still ran
        m2 = e.map(D, buckets_log2=8)
        with pytest.warns(abi.DroppedUpdatesWarning):
            pos, data, n = m2.process_scan(pts, s)
        assert n == pos.shape[0] == 0  # the first map holds every heap block
        m.gc((1e4, 1e4, 0.0))          # frees them all
        with pytest.warns(abi.DroppedUpdatesWarning):
            m2.process_scan(pts, s)
        assert m2.num_blocks() == 64
        pos, data, n = m2.publish(s)
        assert n > 0
        m2.gc((1e4, 1e4, 0.0))
        small, s2 = dense_scan(1, n=16)
        buf.upload(small)
        m2.integrate(buf, small.shape[0], s2)
        code, st = last_stats_status(e)
        assert code == abi.TSDF_OK and st["updates_dropped"] == 0
        m.close()
        m2.close()
        buf.free()


def test_process_scan_refuses_rows_that_are_not_float4(eng):
    """Host points must be float4 rows (x, y, z, pad): an (n, 3) array would be read as 16-byte rows."""
    pts, s = dense_scan(0, n=64)
    m = eng.map(D, buckets_log2=8)
    for bad in (pts[:, :3], pts.reshape(-1), pts[:, :2]):
        with pytest.raises(ValueError):
            m.process_scan(bad, s)
    assert m.num_blocks() == 0
    m.process_scan(pts, s)
    assert m.num_blocks() > 0
    m.close()
