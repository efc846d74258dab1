"""GPU: batched pass, HIP-graph replays, serving loop and FP16 state.

* a batch of scans from several maps gives bit-identical maps to the per-scan
  path (same kernels, same per-voxel fold order);
* graph replays match the eager batch path on every replay, with eager work
  and synchronisations between replays;
* a served map ends bit-identical whether its scans ran alone or mixed with
  other maps' scans, and matches the oracle within tolerance;
* FP16 state keeps keys and touched sets exact, SDF within the packed tolerance.
"""
import numpy as np
import pytest

from ros_tsdf_amd import abi, configs, synth, workloads
from tests.helpers import compare_maps

pytestmark = pytest.mark.gpu

P = configs.SMALL_PARAMS


def small_scans(n_maps, per_map, n_points=None):
    out = []
    for mi in range(n_maps):
        w = configs.map_workload(configs.SMALL, mi)
        boxes = synth.make_boxes(w.scene, w.seed, mi)
        for k in range(per_map):
            p, s, _ = synth.make_scan(w, k, map_id=mi, boxes=boxes, n_points=n_points)
            out.append((p, s, mi))
    return out


def exports_equal(a, b):
    ka, sa, wa = a.export()
    kb, sb, wb = b.export()
    return np.array_equal(ka, kb) and np.array_equal(sa.view(np.uint32), sb.view(np.uint32)) and \
        np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


@pytest.fixture(scope="module")
def eng():
    from ros_tsdf_amd import Engine
    e = Engine(heap_blocks=1 << 16, total_buckets=1 << 18, max_batch_points=1 << 16)
    yield e
    e.close()


def test_batch_equals_per_scan(eng):
    scans = small_scans(3, 4)
    # interleave maps: scan order within each map preserved
    order = sorted(range(len(scans)), key=lambda i: (i % 4, scans[i][2]))
    scans = [scans[i] for i in order]
    pool = workloads.Pool(scans)
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    mb = [eng.map(P, buckets_log2=12) for _ in range(3)]
    ms = [eng.map(P, buckets_log2=12) for _ in range(3)]
    descs = [(pool.maps[i], pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]
    eng.integrate_batch(mb, descs, buf)
    for (mi, n, off, s) in descs:
        ms[mi].integrate(buf, n, s, offset_points=off)
    eng.synchronize()
    for a, b in zip(mb, ms):
        assert exports_equal(a, b)
    for m in mb + ms:
        m.close()
    buf.free()


def test_batched_dense_pass_equals_per_scan():
    """Four full-size dense scans of each of two maps in one pass: thousands of blocks of more than 1024
    records (the per-scan count times four), more than the large-block fold has workgroups, so they are
    handed out by its pass counter.  One eager batched pass, then two replays of the captured pass (the
    counter is cleared inside the graph): after each, every map equals the per-scan path's, bit for bit."""
    from ros_tsdf_amd import BatchGraph, Engine
    bp, bs = workloads.generate_maps(configs.DENSE, 2, 12, workers=1)  # no fork once the GPU is in use
    with Engine(heap_blocks=1 << 18, total_buckets=1 << 20, max_batch_points=1 << 20) as e:
        mb = [e.map(configs.DENSE_PARAMS, buckets_log2=16) for _ in range(2)]
        ms = [e.map(configs.DENSE_PARAMS, buckets_log2=16) for _ in range(2)]
        g = BatchGraph(e, mb, max_points=8 * 131072, max_scans=8)
        for third in range(3):
            scans = [(bp[mi][4 * third + k], bs[mi][4 * third + k], mi) for mi in range(2) for k in range(4)]
            pool = workloads.Pool(scans)
            buf = e.buffer(pool.points.nbytes)
            buf.upload(pool.points)
            descs = [(pool.maps[i], pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]
            if third == 0:
                e.integrate_batch(mb, descs, buf)
            else:
                g.launch(descs, buf)
            e.synchronize()
            assert e.stats()["voxel_updates"] > 8 * 1000000
            for (mi, n, off, s) in descs:
                ms[mi].integrate(buf, n, s, offset_points=off)
            e.synchronize()
            buf.free()
            for a, b in zip(mb, ms):
                assert exports_equal(a, b), "pass %d" % third
        g.close()
        for m in mb + ms:
            m.close()


def test_batched_parked_sensor_saturates_like_oracle(oracle_lib):
    """A parked sensor in batches: four copies of the first dense scan per pass of one map, two passes, at
    max_weight 8.  A pass puts four scans' records on each block (blocks of more than 4096 records take the
    voxel-range and point-round folds), and more than a quarter of the voxels reach max_weight, so those
    folds, the 1025-4096-record tier and the in-place tier all run their update-by-update paths.  The map
    matches the FP64 oracle fed the same eight scans, and equals the per-scan path's bit for bit."""
    from ros_tsdf_amd import Engine
    params = (0.1, 0.4, 8.0, 2500.0, 22500.0)
    first = next(iter(synth.scans(configs.DENSE, stop=1)))
    pts, s = first[1], first[2]
    pool = workloads.Pool([(pts, s, 0)] * 4)
    with Engine(heap_blocks=1 << 17, total_buckets=1 << 18, max_batch_points=1 << 20) as e:
        buf = e.buffer(pool.points.nbytes)
        buf.upload(pool.points)
        mb = e.map(params, buckets_log2=16)
        ms = e.map(params, buckets_log2=16)
        descs = [(0, pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]
        for _ in range(2):
            e.integrate_batch([mb], descs, buf)
            e.synchronize()
            st = e.stats()
            assert st["updates_dropped"] == 0 and st["insert_failures"] == 0
            for (_, n, off, sv) in descs:
                ms.integrate(buf, n, sv, offset_points=off)
        e.synchronize()
        assert exports_equal(mb, ms)
        om = oracle_lib.OracleMap(params)
        for _ in range(8):
            om.integrate(pts, s)
        out = compare_maps(mb, om, params[1], configs.SDF_TOL_T, configs.WEIGHT_TOL_REL * params[2])
        _, _, w = mb.export()
        w = np.asarray(w)
        assert np.mean(w[w > 0] >= params[2]) > 0.25, "too few voxels reached max_weight"
        print("batched parked:", out)
        om.close()
        mb.close()
        ms.close()
        buf.free()


def test_graph_replays_match_eager(eng):
    from ros_tsdf_amd import BatchGraph
    scans = small_scans(2, 6)
    pool = workloads.Pool(scans)
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    batches = [[0, 6, 1, 7], [2, 8, 3, 9], [4, 10, 5, 11]]  # 2 scans of each map per batch, in map order
    mg = [eng.map(P, buckets_log2=12) for _ in range(2)]
    me = [eng.map(P, buckets_log2=12) for _ in range(2)]
    g = BatchGraph(eng, mg, max_points=4 * 1000, max_scans=4)
    # batches must be contiguous in the point buffer: build one buffer per batch
    bufs = []
    for b in batches:
        sub = workloads.Pool([scans[i] for i in b])
        bb = eng.buffer(sub.points.nbytes)
        bb.upload(sub.points)
        bufs.append((sub, bb))
    for r, (sub, bb) in enumerate(bufs):
        d = [(sub.maps[i], sub.counts[i], sub.offsets[i], sub.sensors[i]) for i in range(len(sub))]
        g.launch(d, bb)
        eng.synchronize()
        eng.integrate_batch(me, d, bb)  # eager work between replays
        eng.synchronize()
        for a, b in zip(mg, me):
            assert exports_equal(a, b), "replay %d differs from the eager pass" % r
    g.close()
    for m in mg + me:
        m.close()
    for _, bb in bufs:
        bb.free()
    buf.free()


def test_graph_launch_argument_errors(eng):
    """A graph launch with a null or misaligned point buffer, a negative offset, scans out of order or a null
    scan array is refused before anything is enqueued; the map stays empty and a valid launch still runs."""
    import ctypes as C
    from ros_tsdf_amd import BatchGraph
    from ros_tsdf_amd.engine import _scan_array
    scans = small_scans(1, 2)
    pool = workloads.Pool(scans)
    buf = eng.buffer(pool.points.nbytes + 16)
    buf.upload(pool.points)
    m = eng.map(P, buckets_log2=12)
    g = BatchGraph(eng, [m], max_points=4000, max_scans=2)
    d = [(0, pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]
    L = abi.load()

    def launch(desc, ptr, n=None):
        arr = _scan_array(desc)
        return L.tsdf_batch_graph_launch(C.c_void_p(g.h), arr, len(desc) if n is None else n, C.c_void_p(ptr))

    assert launch(d, None) == abi.TSDF_ERR_INVALID_ARG                        # no point buffer
    assert launch(d, buf.ptr + 4) == abi.TSDF_ERR_INVALID_ARG                 # off a 16-byte boundary
    neg = [(0, d[0][1], -1, d[0][3])]
    assert launch(neg, buf.ptr) == abi.TSDF_ERR_INVALID_ARG                   # negative offset
    assert launch(d[::-1], buf.ptr) == abi.TSDF_ERR_INVALID_ARG               # not contiguous and in order
    assert L.tsdf_batch_graph_launch(C.c_void_p(g.h), None, 2, C.c_void_p(buf.ptr)) == abi.TSDF_ERR_INVALID_ARG
    eng.synchronize()
    assert m.num_blocks() == 0
    g.launch(d, buf)
    eng.synchronize()
    assert m.num_blocks() > 0
    g.close()
    m.close()
    buf.free()


def test_serve_argument_errors(eng):
    """The serving loop refuses a null point buffer, a negative point offset and a map of another engine
    before it starts; the map stays empty and a valid run then serves every request."""
    from ros_tsdf_amd import Engine, serve
    scans = small_scans(1, 2)
    pool = workloads.Pool(scans)
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    m = eng.map(P, buckets_log2=12)
    reqs = [(0.0, 0, pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]

    class Raw:  # a bare device pointer: the C ABI's own checks, without the Python buffer's range check
        def __init__(self, ptr):
            self.ptr = ptr

    def refused(maps, requests, points):
        with pytest.raises(abi.TsdfError) as ei:
            serve(eng, maps, requests, points)
        return ei.value.code == abi.TSDF_ERR_INVALID_ARG

    assert refused([m], reqs, Raw(None))
    assert refused([m], [(0.0, 0, reqs[0][2], -16, reqs[0][4])], Raw(buf.ptr))
    with Engine(heap_blocks=1 << 12, total_buckets=1 << 12) as other:
        om = other.map(P, buckets_log2=8)
        assert refused([om], reqs, buf)
        om.close()
    assert m.num_blocks() == 0
    lat, _, st = serve(eng, [m], reqs, buf)
    assert st["requests"] == len(reqs) and (lat > 0).all() and m.num_blocks() > 0
    m.close()
    buf.free()


def test_serving_alone_vs_mixed(eng, oracle_lib):
    from ros_tsdf_amd import serve
    n_maps, per_map = 4, 6
    scans = small_scans(n_maps, per_map)
    pool = workloads.Pool(scans)
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    rng = np.random.default_rng(3)
    # requests: each map's scans in order, maps interleaved, bursty arrivals
    reqs = []
    nxt = [0] * n_maps
    while any(x < per_map for x in nxt):
        mi = int(rng.integers(0, n_maps))
        if nxt[mi] >= per_map:
            continue
        j = mi * per_map + nxt[mi]
        nxt[mi] += 1
        t = 0.002 * (len(reqs) // 6)  # bursts of 6 simultaneous arrivals
        reqs.append((t, mi, pool.counts[j], pool.offsets[j], pool.sensors[j]))
    maps = [eng.map(P, buckets_log2=12) for _ in range(n_maps)]
    lat, bof, st = serve(eng, maps, reqs, buf, max_batch_scans=8, max_inflight=2)
    assert st["requests"] == len(reqs) and np.all(lat > 0)
    assert st["batches"] < len(reqs), "requests were never batched together"
    # map 0 alone, one request at a time
    alone = [eng.map(P, buckets_log2=12)]
    reqs0 = [(t, 0, n, off, s) for (t, mi, n, off, s) in reqs if mi == 0]
    serve(eng, alone, reqs0, buf, max_batch_scans=1, max_inflight=1)
    assert exports_equal(maps[0], alone[0])
    # and the oracle
    om = oracle_lib.OracleMap(P)
    for k in range(per_map):
        om.integrate(scans[k][0], scans[k][1])
    compare_maps(maps[0], om, P[1], configs.SDF_TOL_T, configs.WEIGHT_TOL_REL * P[2])
    om.close()
    for m in maps + alone:
        m.close()
    buf.free()


@pytest.mark.parametrize("fmt", [abi.TSDF_STATE_F16, abi.TSDF_STATE_F16_QW], ids=["fp16", "fp16_fixed_weight"])
def test_fp16_state_error(oracle_lib, fmt):
    from ros_tsdf_amd import Engine
    params = configs.DENSE_PARAMS
    with Engine(heap_blocks=1 << 16, total_buckets=1 << 18, max_batch_points=1 << 17,
                state_format=fmt) as e16:
        m = e16.map(params, buckets_log2=15)
        om = oracle_lib.OracleMap(params)
        buf = e16.buffer(131072 * 16)
        for k, pts, s in synth.scans(configs.DENSE, stop=2):
            buf.upload(pts)
            m.integrate(buf, pts.shape[0], s)
            om.integrate(pts, s, threads=8)
        # FP16 sdf/T has a half-ulp of 2^-11 on [0.5, 1): tolerance 1e-3 T.
        # FP16 weights carry 11 significant bits: tolerance 1 % of the weight.
        # Fixed-point weights have a step of max_weight / 65535 (0.0153 at
        # 1000), half of it lost per store: tolerance 2 stores' worth.  That
        # step is coarse for the first updates of a voxel (W ~ 1), and the
        # next fold's sdf moves by w |S - d| / (W + w)^2 per unit of weight
        # error (|S - d| < 2.5 T here): sdf tolerance 1e-2 T for that format.
        gk, gs, gw = m.export()
        ok, os_, ow = om.export()
        assert np.array_equal(gk, ok)
        t = ow != 0
        assert np.array_equal(gw != 0, t)
        e = np.abs(gs[t] - os_[t]) / params[1]
        we = np.abs(gw[t] - ow[t])
        print("max %.3e rms %.3e T; weight abs max %.3e rel max %.3e" % (
            e.max(), np.sqrt(np.mean(e * e)), we.max(), (we / ow[t]).max()))
        if fmt == abi.TSDF_STATE_F16:
            assert e.max() <= 1e-3
            assert (we / ow[t]).max() <= 1e-2
        else:
            assert e.max() <= 1e-2
            assert we.max() <= 2 * 0.5 * params[2] / 65535 + 1e-4
        buf.free()
        m.close()
        om.close()


def dense_scans(n_maps, per_map, n_points):
    out = []
    for mi in range(n_maps):
        w = configs.map_workload(configs.DENSE, mi)
        boxes = synth.make_boxes(w.scene, w.seed, mi)
        for k in range(per_map):
            p, s, _ = synth.make_scan(w, k, map_id=mi, boxes=boxes, n_points=n_points)
            out.append((p, s, mi))
    return out


def test_autotune_knobs_do_not_change_results():
    """Every launch shape and table shape the autotuner may pick integrates the
    same scans into the same maps, bit for bit (dense scans, so every fold
    size class is exercised)."""
    from ros_tsdf_amd import BatchGraph, Engine, autotune
    n_maps, per_map = 2, 3
    scans = dense_scans(n_maps, per_map, 65536)
    # batches: scan k of every map, scans contiguous per batch
    order = [mi * per_map + k for k in range(per_map) for mi in range(n_maps)]
    pool = workloads.Pool([scans[i] for i in order])
    batches = [[(pool.maps[j], pool.counts[j], pool.offsets[j], pool.sensors[j])
                for j in range(k * n_maps, (k + 1) * n_maps)] for k in range(per_map)]
    params = configs.DENSE_PARAMS
    cfgs = [dict(traverse_block=64, fold_grid=512, buckets_log2=13, max_chain=64),
            dict(traverse_block=256, fold_grid=4096, buckets_log2=16, max_chain=64),
            dict(traverse_block=128, fold_grid=1024, buckets_log2=15, max_chain=16),
            dict(traverse_block=256, rays_per_lane=2, fold_grid=4096, buckets_log2=16, max_chain=64),
            dict(traverse_block=192, rays_per_lane=4, fold_grid=2048, buckets_log2=15, max_chain=64),
            dict(traverse_block=256, rays_per_lane=1, traverse_queue=1, fold_grid=4096, buckets_log2=16, max_chain=64),
            dict(traverse_block=128, rays_per_lane=4, traverse_queue=1, fold_grid=4096, buckets_log2=15, max_chain=16)]
    with Engine(heap_blocks=1 << 18, total_buckets=1 << 19, max_batch_points=1 << 18) as e:
        buf = e.buffer(pool.points.nbytes)
        buf.upload(pool.points)
        ref = None
        for cfg in cfgs:
            e.set_launch(**autotune.launch_kwargs(cfg))
            maps = [e.map(params, buckets_log2=cfg["buckets_log2"], max_chain=cfg["max_chain"])
                    for _ in range(n_maps)]
            g = BatchGraph(e, maps, max_points=n_maps * 65536, max_scans=n_maps)
            for b in batches:
                g.launch(b, buf)
                st = e.stats()
                assert st["insert_failures"] == st["heap_failures"] == st["updates_dropped"] == 0, (cfg, st)
            g.close()
            out = [m.export() for m in maps]
            if ref is None:
                ref = out
            else:
                for (ka, sa, wa), (kb, sb, wb) in zip(ref, out):
                    assert np.array_equal(ka, kb), cfg
                    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32)), cfg
                    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), cfg
            for m in maps:
                m.close()
        # the tuner itself, over a two-value space: it measures, filters and picks
        best, hist = autotune.tune(e, params, n_maps, batches, buf,
                                   space={"traverse_block": (128, 256), "buckets_log2": (14, 15)},
                                   start=dict(traverse_block=256, fold_grid=1024, buckets_log2=15, max_chain=64))
        assert len(hist) >= 3 and all(r["failures"] == 0 and r["points_per_s"] > 0 for _, r in hist)
        assert best["traverse_block"] in (128, 256) and best["buckets_log2"] in (14, 15)
        assert e.get_launch()["traverse_block"] == best["traverse_block"]
        buf.free()


def test_graph_launch_after_a_map_is_destroyed(eng):
    """A graph does not keep its maps alive: after one of them is closed (and its slot reused by a new map),
    a launch with a scan of it is refused with nothing enqueued, while launches with scans of the other map
    still run.  A map of another engine is refused when the graph is created."""
    import ctypes as C
    from ros_tsdf_amd import BatchGraph, Engine
    from ros_tsdf_amd.engine import _scan_array
    scans = small_scans(2, 1)
    pool = workloads.Pool(scans)
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    m0, m1 = eng.map(P, buckets_log2=12), eng.map(P, buckets_log2=12)
    g = BatchGraph(eng, [m0, m1], max_points=pool.total, max_scans=2)
    both = [(pool.maps[i], pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]
    m1.close()
    newcomer = eng.map(P, buckets_log2=12)  # may take the closed map's slot
    L = abi.load()
    arr = _scan_array(both)
    assert L.tsdf_batch_graph_launch(C.c_void_p(g.h), arr, 2, C.c_void_p(buf.ptr)) == abi.TSDF_ERR_INVALID_ARG
    assert "destroyed" in L.tsdf_last_error().decode()
    eng.synchronize()
    assert m0.num_blocks() == 0 and newcomer.num_blocks() == 0
    g.launch(both[:1], buf)  # map 0 alone
    eng.synchronize()
    assert m0.num_blocks() > 0 and newcomer.num_blocks() == 0
    g.close()
    with Engine(heap_blocks=1 << 12, total_buckets=1 << 12) as other:
        om = other.map(P, buckets_log2=8)
        with pytest.raises(abi.TsdfError) as ei:
            BatchGraph(eng, [m0, om], max_points=pool.total, max_scans=2)
        assert ei.value.code == abi.TSDF_ERR_INVALID_ARG
        om.close()
    m0.close()
    newcomer.close()
    buf.free()
