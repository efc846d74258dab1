"""The fold of a pass by (voxel, scan) groups.

A group's updates are summed exactly (fixed point) and applied as one weighted mean while the voxel's
weight stays within max_weight; from the update that would cross it, the group is folded update by update
in point order.  The cases below put many updates on few voxels of one block in one pass, in every fold tier
(16/17 .. 12606 records per block, the bounds of test_edges_gpu's tier cases, and 768/769), with a max_weight
that groups cross inside the pass (4) or never reach (1000), and with 1, 3 or 10 scans of the map in the pass.
Each map is compared with the FP64 oracle, which folds update by update in point order, and the batched passes
with the same scans one pass each, bit for bit, in all three state formats.
"""
import numpy as np
import pytest

from ros_tsdf_amd import abi, workloads
from tests.helpers import compare_maps
from tests.test_edges_gpu import tier_scans, tol

pytestmark = pytest.mark.gpu

FORMATS = (abi.TSDF_STATE_F32, abi.TSDF_STATE_F16, abi.TSDF_STATE_F16_QW)


@pytest.fixture(scope="module")
def engines():
    from ros_tsdf_amd import Engine
    es = {f: Engine(heap_blocks=1 << 12, total_buckets=1 << 14, max_batch_points=1 << 16, state_format=f)
          for f in FORMATS}
    yield es
    for e in es.values():
        e.close()


def group_scans(n, cols, nscans):
    """tier_scans(n, cols) with every column's points dealt round-robin to nscans scans: scan k of column c
    holds its points k, k + nscans, ...  Scans are ordered share-major, so each scan of the pass adds
    records to the same voxels."""
    cols_scans = tier_scans(n, cols)
    out = []
    for k in range(nscans):
        for pts, s in cols_scans:
            part = pts[k::nscans]
            if part.shape[0]:
                out.append((part, s))
    return out


def run(eng, params, scans, passes, batched):
    pool = workloads.Pool([(p, s, 0) for p, s in scans])
    buf = eng.buffer(pool.points.nbytes)
    buf.upload(pool.points)
    m = eng.map(params, buckets_log2=10)
    descs = [(0, pool.counts[i], pool.offsets[i], pool.sensors[i]) for i in range(len(pool))]
    for _ in range(passes):
        if batched:
            eng.integrate_batch([m], descs, buf)
        else:
            for d in descs:
                m.integrate(buf, d[1], d[3], offset_points=d[2])
    eng.synchronize()
    st = eng.stats()
    buf.free()
    return m, st


# (768, 1) and (769, 1): the largest block whose clamped groups the first fold launch folds in order itself,
# and the smallest it hands to the second launch
CASES = [(n, cols, ns, maxw)
         for (n, cols) in [(16, 1), (17, 1), (64, 1), (65, 1), (128, 1), (129, 1), (768, 1), (769, 1), (1024, 1),
                           (1025, 1), (4096, 1), (4097, 2), (8193, 2), (12606, 1)]
         for (ns, maxw) in [(1, 4.0), (3, 4.0), (3, 1000.0)]]
CASES += [(768, 1, 10, 4.0), (1024, 1, 10, 4.0), (1024, 1, 10, 1000.0), (4096, 1, 10, 4.0), (4096, 1, 10, 1000.0),
          (8193, 2, 10, 1000.0)]


@pytest.mark.parametrize("n,cols,ns,maxw", CASES)
def test_groups_match_oracle_and_per_scan(engines, oracle_lib, n, cols, ns, maxw):
    params = (0.1, 0.25, maxw, 2500.0, 22500.0)
    scans = group_scans(n, cols, ns)
    passes = 3
    eng = engines[abi.TSDF_STATE_F32]
    mb, st = run(eng, params, scans, passes, batched=True)
    assert st["updates_dropped"] == 0
    om = oracle_lib.OracleMap(params)
    for _ in range(passes):
        for p, s in scans:
            om.integrate(p, s)
    out = compare_maps(mb, om, params[1], *tol(params))
    _, _, w = mb.export()
    if maxw < 10:
        assert (w == np.float32(maxw)).any(), "no voxel reached max_weight"
    # the same scans one pass each: the same map, bit for bit
    ms, _ = run(eng, params, scans, passes, batched=False)
    for a, b in zip(mb.export(), ms.export()):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    mb.close()
    ms.close()
    om.close()
    assert out["voxels"] > 0


@pytest.mark.parametrize("fmt", FORMATS[1:])
@pytest.mark.parametrize("n,cols,ns", [(64, 1, 3), (1024, 1, 3), (4096, 1, 10), (8193, 2, 3), (12606, 1, 3)])
def test_packed_formats_batched_equals_per_scan(engines, n, cols, ns, fmt):
    """The state is rounded to its stored format between a voxel's groups, as a store and a load between
    passes would: a batch gives the per-scan map bit for bit in the packed formats too."""
    params = (0.1, 0.25, 4.0, 2500.0, 22500.0)
    scans = group_scans(n, cols, ns)
    eng = engines[fmt]
    mb, _ = run(eng, params, scans, 3, batched=True)
    ms, _ = run(eng, params, scans, 3, batched=False)
    for a, b in zip(mb.export(), ms.export()):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    mb.close()
    ms.close()
