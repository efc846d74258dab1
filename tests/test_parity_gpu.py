"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

Bar: block keys, touched-voxel sets and published positions bit-exact; SDF
within SDF_TOL_T * truncation and weight within WEIGHT_TOL_REL * max_weight
(ros_tsdf_amd.configs).
"""
import pytest

from ros_tsdf_amd import configs, synth
from tests.helpers import compare_maps, compare_publish

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from ros_tsdf_amd import Engine
    e = Engine(heap_blocks=1 << 17, total_buckets=1 << 19, max_batch_points=1 << 17)
    yield e
    e.close()


def _tol(params):
    return configs.SDF_TOL_T, configs.WEIGHT_TOL_REL * params[2]


def run_sequence(eng, oracle, params, w, scans, map_kwargs=None, publish=True, gc=True, check_every=1):
    """Per scan: integrate -> publish -> gc on both sides, comparing as we go."""
    m = eng.map(params, **(map_kwargs or {}))
    om = oracle.OracleMap(params)
    T = params[1]
    st, wt = _tol(params)
    buf = None
    results = []
    try:
        for j, (k, pts, s) in enumerate(scans):
            if buf is None or buf.nbytes < pts.nbytes:
                if buf is not None:
                    buf.free()
                buf = eng.buffer(max(pts.nbytes, 16))
            buf.upload(pts)
            m.integrate(buf, pts.shape[0], s)
            ost = om.integrate(pts, s)
            gst = eng.stats()
            assert gst["voxel_updates"] == ost["n_voxel_updates"], (gst, ost)
            assert gst["unique_voxels"] == ost["n_unique_voxels"], (gst, ost)
            assert gst["valid"] == ost["n_valid"] and gst["skipped"] == ost["n_skipped"]
            assert gst["block_cells"] == ost["n_block_cells"]
            assert gst["insert_failures"] == 0 and gst["heap_failures"] == 0 and gst["updates_dropped"] == 0
            if publish:
                gpos, gdata, n = m.publish(s)
                opos, osdf, ow = om.publish(s)
                compare_publish(gpos, gdata, opos, osdf, ow, T, st, wt)
            if gc:
                assert m.gc(s) == om.gc(s)
            if (j + 1) % check_every == 0:
                results.append(compare_maps(m, om, T, st, wt))
        results.append(compare_maps(m, om, T, st, wt))
    finally:
        if buf is not None:
            buf.free()
        m.close()
        om.close()
    return results


def test_small_workload_parity(eng, oracle_lib):
    w = configs.SMALL
    r = run_sequence(eng, oracle_lib, configs.SMALL_PARAMS, w, synth.scans(w, stop=40), check_every=10)
    print("small:", r[-1])


def test_dense_workload_parity(eng, oracle_lib):
    w = configs.DENSE
    r = run_sequence(eng, oracle_lib, configs.DENSE_PARAMS, w, synth.scans(w, stop=3))
    print("dense:", r[-1])


def test_parked_sensor_saturates_like_oracle(eng, oracle_lib):
    """A parked sensor: the first dense scan integrated six times at max_weight 8, so that the voxels of blocks
    of every size class reach max_weight within a few passes and their updates take the update-by-update fold
    from there: in place in the first fold launch for blocks of up to 768 records, handed to the second
    launch for larger ones.  Every pass is compared with the FP64 oracle."""
    import numpy as np
    params = (0.1, 0.4, 8.0, 2500.0, 22500.0)
    w = configs.DENSE
    first = next(iter(synth.scans(w, stop=1)))
    r = run_sequence(eng, oracle_lib, params, w, [first] * 6, publish=False, gc=False)
    print("parked:", r[-1])
    om = oracle_lib.OracleMap(params)
    for _ in range(6):
        om.integrate(first[1], first[2])
    _, _, wts = om.export()
    om.close()
    wts = np.asarray(wts)
    assert np.mean(wts[wts > 0] >= params[2]) > 0.2, "too few voxels reached max_weight"


def test_tiny_hash_overflow_chains(eng, oracle_lib):
    """4 home buckets force long overflow chains (and chain-extension races)."""
    w = configs.SMALL
    r = run_sequence(eng, oracle_lib, configs.SMALL_PARAMS, w, synth.scans(w, stop=8),
                     map_kwargs=dict(buckets_log2=2, overflow_buckets=4096, max_chain=4096), check_every=4)
    print("tiny-hash:", r[-1])


def test_gc_recycles_blocks(eng, oracle_lib):
    """A 30 m GC radius removes blocks every scan; freed heap blocks are reused zeroed."""
    params = (0.5, 0.1, 1000.0, 2500.0, 900.0)
    w = configs.SMALL
    free0 = eng.free_blocks()
    run_sequence(eng, oracle_lib, params, w, synth.scans(w, stop=30), check_every=5)
    assert eng.free_blocks() == free0


@pytest.mark.parametrize("launch", [dict(rays_per_lane=4), dict(rays_per_lane=2, traverse_queue=1),
                                    dict(traverse_block=64, rays_per_lane=3, traverse_queue=1)])
def test_dense_parity_across_traversal_shapes(eng, oracle_lib, launch):
    """The traversal's chunking (rays per lane, the persistent work-queue grid) changes which rays share an LDS
    block table, and with 4 x 256 rays per table the table overflows into the global fallback paths; none of it
    may change a key, a touched voxel or a value beyond the tolerance."""
    w = configs.DENSE
    eng.set_launch(**launch)
    try:
        got = eng.get_launch()
        for k, v in launch.items():
            assert got[k] == v
        r = run_sequence(eng, oracle_lib, configs.DENSE_PARAMS, w, synth.scans(w, stop=2), publish=False, gc=False)
    finally:
        eng.set_launch()
    print("dense", launch, r[-1])
