"""Map snapshot / evict / restore: tsdf_map_import is the inverse of tsdf_map_export.

Properties checked (none depends on the map's size):
- export -> import into an empty map -> export gives back the same blocks, bit for bit;
- a restored map and the original, fed the same further scans, stay bit-identical (integration,
  publish) although their hash and heap layouts differ;
- evicting (export, destroy) returns the blocks to the heap, restoring takes them again;
- packed state formats re-encode exported values to the same values;
- a block quota refuses the excess blocks, which leave no key behind;
- importing a map's own snapshot over it changes nothing;
- a map evicted, restored into another table shape and integrated further equals the FP64 oracle fed
  every scan, in all three state formats.
"""
import numpy as np
import pytest

from ros_tsdf_amd import Engine, abi, configs, synth

pytestmark = pytest.mark.gpu

SCANS = 5
STRIDE = 3  # every third point of a dense scan keeps the test light


@pytest.fixture(scope="module")
def scans():
    out = []
    for k in range(SCANS):
        pts, sensor, _ = synth.make_scan(configs.DENSE, k)
        out.append((np.ascontiguousarray(pts[::STRIDE]), np.asarray(sensor, np.float32)))
    return out


def integrate(eng, m, scans):
    buf = eng.buffer(max(p.shape[0] for p, _ in scans) * 16)
    for pts, sensor in scans:
        buf.upload(pts)
        m.integrate(buf, pts.shape[0], sensor)
    eng.synchronize()
    buf.free()


def same_export(a, b):
    (ka, sa, wa), (kb, sb, wb) = a, b
    assert np.array_equal(ka, kb)
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


def sorted_publish(m, sensor):
    pos, data, _ = m.publish(sensor)
    order = np.lexsort(pos.T[::-1])
    return pos[order], data[order]


@pytest.mark.parametrize("fmt", [abi.TSDF_STATE_F32, abi.TSDF_STATE_F16, abi.TSDF_STATE_F16_QW])
def test_roundtrip_and_continue(scans, fmt):
    with Engine(heap_blocks=1 << 18, total_buckets=1 << 18, state_format=fmt) as eng:
        a = eng.map(configs.DENSE_PARAMS, buckets_log2=15)
        integrate(eng, a, scans[:3])
        snap = a.export()
        n = snap[0].shape[0]
        assert n > 1000
        b = eng.map(configs.DENSE_PARAMS, buckets_log2=15)
        assert b.import_blocks(*snap) == (n, 0)
        assert b.num_blocks() == n
        same_export(b.export(), snap)
        # the same further scans keep both maps identical
        integrate(eng, a, scans[3:])
        integrate(eng, b, scans[3:])
        ea, eb = a.export(), b.export()
        assert ea[0].shape[0] > n
        same_export(ea, eb)
        pa, pb = sorted_publish(a, scans[-1][1]), sorted_publish(b, scans[-1][1])
        assert pa[0].shape[0] > 0
        assert np.array_equal(pa[0].view(np.uint32), pb[0].view(np.uint32))
        assert np.array_equal(pa[1].view(np.uint32), pb[1].view(np.uint32))
        a.close()
        b.close()


def test_evict_and_restore(scans):
    with Engine(heap_blocks=1 << 18, total_buckets=1 << 18) as eng:
        free0 = eng.free_blocks()
        a = eng.map(configs.DENSE_PARAMS, buckets_log2=15)
        integrate(eng, a, scans[:2])
        snap = a.export()
        n = snap[0].shape[0]
        assert eng.free_blocks() == free0 - n
        a.close()  # evicted: the blocks return to the heap
        assert eng.free_blocks() == free0
        c = eng.map(configs.DENSE_PARAMS, buckets_log2=15)
        assert c.import_blocks(*snap) == (n, 0)
        assert eng.free_blocks() == free0 - n
        same_export(c.export(), snap)
        # reimporting its own snapshot overwrites every block with the same values
        assert c.import_blocks(*snap) == (n, 0)
        assert c.num_blocks() == n
        same_export(c.export(), snap)
        c.close()


def test_quota_refuses_excess(scans):
    with Engine(heap_blocks=1 << 18, total_buckets=1 << 18) as eng:
        a = eng.map(configs.DENSE_PARAMS, buckets_log2=15)
        integrate(eng, a, scans[:1])
        keys, sdf, w = a.export()
        n = keys.shape[0]
        q = n // 2
        b = eng.map(configs.DENSE_PARAMS, buckets_log2=15, block_quota=q)
        assert b.import_blocks(keys, sdf, w) == (q, n - q)
        kb, sb, wb = b.export()
        assert kb.shape[0] == q == b.num_blocks()
        # the stored blocks hold their snapshot values; the refused ones left no key behind
        row = {tuple(k): i for i, k in enumerate(keys.tolist())}
        idx = np.array([row[tuple(k)] for k in kb.tolist()])
        assert np.array_equal(sb.view(np.uint32), sdf[idx].view(np.uint32))
        assert np.array_equal(wb.view(np.uint32), w[idx].view(np.uint32))
        a.close()
        b.close()


def test_import_argument_errors(scans):
    with Engine(heap_blocks=1 << 14, total_buckets=1 << 14) as eng:
        m = eng.map(configs.DENSE_PARAMS, buckets_log2=12)
        keys = np.array([[1, 2, 3], [4, 5, 6]], np.int32)
        vals = np.zeros((2, 512), np.float32)
        with pytest.raises(abi.TsdfError) as ei:
            m.import_blocks(np.array([[1, 2, 3], [1, 2, 3]], np.int32), vals, vals)
        assert ei.value.code == abi.TSDF_ERR_INVALID_ARG
        with pytest.raises(abi.TsdfError) as ei:
            m.import_blocks(np.array([[1, 2, 3], [1 << 14, 0, 0]], np.int32), vals, vals)
        assert ei.value.code == abi.TSDF_ERR_INVALID_ARG
        assert m.num_blocks() == 0
        assert m.import_blocks(keys, vals + 0.25, vals + 2.0) == (2, 0)
        k, s, w = m.export()
        assert np.array_equal(k, keys)
        assert (s == 0.25).all() and (w == 2.0).all()
        m.close()


# Tolerances of each state format against the FP64 oracle after SCANS passes: max sdf error in units of T,
# and the weight bound as (relative, absolute).  FP16 sdf/T is rounded once per pass, by at most half an
# ulp of a value below 2 in magnitude (2^-11), and the fold never amplifies an earlier rounding (the old
# sdf is weighted by W / (W + w) < 1): SCANS * 2^-11.  A weight's rounding is not damped (W + w carries it
# on), so the fixed-point weight may be off by half a step (max_weight / 65535) per pass: SCANS half steps.
# FP16 weights keep 11 significant bits (2^-12 relative per pass, bounded here by 1e-2), and the
# fixed-point format's sdf bound of 1e-2 T is that of the packed-state error test in
# tests/test_batch_serve_gpu.py.
FORMAT_TOL = {
    abi.TSDF_STATE_F32: (configs.SDF_TOL_T, (0.0, configs.WEIGHT_TOL_REL * configs.DENSE_PARAMS[2])),
    abi.TSDF_STATE_F16: (SCANS * 2.0 ** -11, (1e-2, 0.0)),
    abi.TSDF_STATE_F16_QW: (1e-2, (0.0, SCANS * 0.5 * configs.DENSE_PARAMS[2] / 65535 + 1e-4)),
}


@pytest.mark.parametrize("fmt", [abi.TSDF_STATE_F32, abi.TSDF_STATE_F16, abi.TSDF_STATE_F16_QW])
def test_evict_restore_continue_matches_oracle(scans, fmt, oracle_lib):
    """Three scans, export, destroy (evict), import into a map of another table shape, two more scans: the
    map equals an oracle map fed all five scans (keys and touched voxels exact, values within the format's
    tolerance), and so does its publish."""
    params = configs.DENSE_PARAMS
    T = params[1]
    sdf_tol, (w_rel, w_abs) = FORMAT_TOL[fmt]
    with Engine(heap_blocks=1 << 18, total_buckets=1 << 18, state_format=fmt) as eng:
        om = oracle_lib.OracleMap(params)
        for p, s in scans:
            om.integrate(p, s, threads=8)
        a = eng.map(params, buckets_log2=15)
        integrate(eng, a, scans[:3])
        snap = a.export()
        a.close()
        b = eng.map(params, buckets_log2=13, max_chain=16)
        assert b.import_blocks(*snap) == (snap[0].shape[0], 0)
        integrate(eng, b, scans[3:])
        gk, gs, gw = b.export()
        ok, os_, ow = om.export()
        assert np.array_equal(gk, ok)
        t = ow != 0
        assert np.array_equal(gw != 0, t)
        e = np.abs(gs[t].astype(np.float64) - os_[t]) / T
        we = np.abs(gw[t].astype(np.float64) - ow[t])
        assert e.max() <= sdf_tol, e.max()
        assert (we <= w_rel * ow[t] + w_abs).all(), we.max()
        s = scans[-1][1]
        gpos, gdata, n = b.publish(s)
        opos, osdf, ow_p = om.publish(s)
        assert n == opos.shape[0] > 0
        go, oo = np.lexsort(gpos.T[::-1]), np.lexsort(opos.T[::-1])
        assert np.array_equal(gpos[go].view(np.uint32), opos[oo].view(np.uint32))
        assert np.abs(gdata[go, 0].astype(np.float64) - osdf[oo]).max() <= sdf_tol * T
        b.close()
        om.close()
