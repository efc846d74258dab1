"""GPU parity at the batched and served workloads' own shapes, against the FP64 oracle.

* batched: launches of 32 scans from 8 maps with their own scenes and trajectories, point counts
  log-uniform in [1,024, 131,072] (workloads.batch_plan), voxel 0.1 m, truncation 0.4 m; each launch is a
  replay of a captured batch graph with the launch and table shapes the autotuner picked on these launches;
* serving: 16 maps, Poisson arrivals on a compressed clock (so that requests overlap on the device and
  are batched together), log-uniform counts, at most 32 scans per batch, maps admitted against the 64 GB
  heap budget.

Every map ends with block keys and touched voxels equal to the oracle's, fed the same scans in the same
order (tsdf_handler.cu:974-1016 is the per-scan sequence each map follows), and sdf within 1e-4 T; a served
map is also bit-identical to the same requests served alone, one per batch.
"""
import numpy as np
import pytest

from ros_tsdf_amd import configs, synth, workloads

pytestmark = pytest.mark.gpu

P = configs.DENSE_PARAMS
T = P[1]
W_TOL = configs.WEIGHT_TOL_REL * P[2]


def map_scans(n_maps, per_map, count_of):
    """scans[m][k] = (points, sensor): scan k of map m, with count_of(m, k) points."""
    out = []
    for m in range(n_maps):
        w = configs.map_workload(configs.DENSE, m)
        boxes = synth.make_boxes(w.scene, w.seed, m)
        row = []
        for k in range(per_map):
            p, s, _ = synth.make_scan(w, k, map_id=m, boxes=boxes, n_points=count_of(m, k))
            row.append((p, s))
        out.append(row)
    return out


def check_export(exp, omap):
    """An engine export against the oracle map: keys and touched voxels exact, values within tolerance."""
    gk, gs, gw = exp
    ok, os_, ow = omap.export()
    assert gk.shape == ok.shape and np.array_equal(gk, ok), "block key sets differ (%d vs %d)" % (gk.shape[0], ok.shape[0])
    t = ow != 0
    assert np.array_equal(gw != 0, t), "touched voxel sets differ"
    if not t.any():
        return int(gk.shape[0]), 0.0
    e = np.abs(gs[t].astype(np.float64) - os_[t]) / T
    we = np.abs(gw[t].astype(np.float64) - ow[t])
    assert e.max() <= configs.SDF_TOL_T, e.max()
    assert we.max() <= W_TOL, we.max()
    return int(gk.shape[0]), float(e.max())


def same_bits(a, b):
    return all(np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y) for x, y in zip(a, b))


def no_failures(st):
    return st["insert_failures"] == 0 and st["heap_failures"] == 0 and st["updates_dropped"] == 0


def test_batched_workload_matches_oracle(oracle_lib):
    from ros_tsdf_amd import BatchGraph, Engine, autotune
    n_maps, n_launch = configs.BATCH_MAPS, 2
    plan = workloads.batch_plan(n_batches=n_launch)  # 32 scans per launch, 4 consecutive scans of each map
    assert all(len(b) == configs.BATCH_SCANS for b in plan)
    counts = {(m, k): n for b in plan for (m, k, n) in b}
    assert min(counts.values()) >= configs.BATCH_MIN_POINTS and max(counts.values()) <= configs.BATCH_MAX_POINTS
    per_map = 1 + max(k for b in plan for (_, k, _) in b)
    scans = map_scans(n_maps, per_map, lambda m, k: counts[(m, k)])
    # one point buffer, launches contiguous in it
    flat = [(scans[m][k][0], scans[m][k][1], m) for b in plan for (m, k, _) in b]
    pool = workloads.Pool(flat)
    descs, j = [], 0
    for b in plan:
        descs.append([(pool.maps[i], pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(j, j + len(b))])
        j += len(b)
    max_pts = max(sum(d[1] for d in ds) for ds in descs)
    with Engine(heap_blocks=1 << 20, total_buckets=1 << 22, max_batch_points=max_pts) as e:
        buf = e.buffer(pool.points.nbytes)
        buf.upload(pool.points)
        # the tuner measures these launches (a two-value space per knob keeps the test short) and keeps
        # only configurations without failures
        best, hist = autotune.tune(e, P, n_maps, descs, buf,
                                   space={"traverse_block": (128, 256), "buckets_log2": (17, 18)},
                                   start=dict(autotune.DEFAULT))
        assert all(r["failures"] == 0 for _, r in hist)
        maps = [e.map(P, buckets_log2=best["buckets_log2"], max_chain=best["max_chain"]) for _ in range(n_maps)]
        g = BatchGraph(e, maps, max_points=max_pts, max_scans=configs.BATCH_SCANS)
        for ds in descs:
            g.launch(ds, buf)
            st = e.stats()
            assert no_failures(st), st
            assert st["points"] == sum(d[1] for d in ds)
        g.close()
        for m in range(n_maps):
            om = oracle_lib.OracleMap(P)
            for k in range(per_map):
                om.integrate(*scans[m][k], threads=8)
            check_export(maps[m].export(), om)
            if m == 0:
                s = scans[0][per_map - 1][1]
                gpos, gdata, n = maps[0].publish(s)
                opos, osdf, ow = om.publish(s)
                assert n == opos.shape[0] > 0
                go, oo = np.lexsort(gpos.T[::-1]), np.lexsort(opos.T[::-1])
                assert np.array_equal(gpos[go].view(np.uint32), opos[oo].view(np.uint32))
                assert np.abs(gdata[go, 0].astype(np.float64) - osdf[oo]).max() <= configs.SDF_TOL_T * T
            om.close()
        for mm in maps:
            mm.close()
        buf.free()


def test_serving_workload_matches_oracle_and_alone(oracle_lib):
    from ros_tsdf_amd import Engine, serve
    n_maps, n_req = configs.SERVE_MAPS, 96
    rate = 8000.0  # scans/s: the arrival clock compressed 20x, so that requests meet on the device
    raw = workloads.poisson_requests(n_req, rate, n_maps, counts="loguniform")  # (t, map, k, n)
    per_map = 1 + max(k for (_, _, k, _) in raw)
    counts = {(m, k): n for (_, m, k, n) in raw}
    scans = map_scans(n_maps, per_map, lambda m, k: counts.get((m, k), 1))
    flat, where = [], {}
    for (_, m, k, _) in raw:
        where[(m, k)] = len(flat)
        flat.append((scans[m][k][0], scans[m][k][1], m))
    pool = workloads.Pool(flat)
    reqs = []
    for (t, m, k, n) in raw:
        i = where[(m, k)]
        reqs.append((t, m, pool.counts[i], pool.offsets[i], pool.sensors[i]))
    budget_blocks = configs.SERVE_HEAP_BUDGET // (512 * 8)
    with Engine(heap_blocks=budget_blocks, total_buckets=1 << 21, max_batch_points=1 << 22) as e:
        buf = e.buffer(pool.points.nbytes)
        buf.upload(pool.points)
        quota = budget_blocks // n_maps
        maps = [e.map(P, buckets_log2=16, block_quota=quota, admit=True) for _ in range(n_maps)]
        lat, bof, st = serve(e, maps, reqs, buf, max_batch_scans=32, max_inflight=2)
        assert st["requests"] == n_req and (lat > 0).all()
        assert st["batches"] < n_req, "requests were never batched together"
        sizes = np.bincount(bof, minlength=int(st["batches"]))
        assert sizes.max() > 1
        for m in range(n_maps):
            mixed = maps[m].export()
            # the oracle, fed map m's scans in arrival order
            om = oracle_lib.OracleMap(P)
            for (_, mm, k, _) in raw:
                if mm == m:
                    om.integrate(*scans[m][k], threads=8)
            check_export(mixed, om)
            om.close()
            # the same requests served alone, one per batch, on a fresh map
            alone = e.map(P, buckets_log2=16)
            serve(e, [alone], [(t, 0, n, off, s) for (t, mm, n, off, s) in reqs if mm == m], buf,
                  max_batch_scans=1, max_inflight=1)
            assert same_bits(alone.export(), mixed), "map %d differs between the mix and alone" % m
            alone.close()
        for mm in maps:
            mm.close()
        buf.free()
