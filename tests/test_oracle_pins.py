"""Pins of the CPU oracle (no GPU).

The reference ships no tests or golden vectors, so each oracle function is
checked against something independent of its own code: exact rational
arithmetic, closed forms, exhaustive FP32 sweeps and invariants chosen so that
a dropped term, a wrong sign or a transposed index fails.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from tests.helpers import RADIUS_PARAMS, RADIUS_POINTS, RADIUS_SENSOR

f32 = np.float32


# ---- initGlobalVars derived constants (H:1289-1317) ------------------------------
@pytest.mark.parametrize("vs,T", [(0.5, 0.1), (0.1, 0.4), (0.5, 2.0), (1.0, 4.0), (0.25, 0.3), (0.2, 1.0)])
def test_derived_constants(vs, T):
    d = oracle.derived([vs, T, 1000, 2500, 250000])
    vs32, T32 = Fr(float(f32(vs))), Fr(float(f32(T)))
    # ceil(2T/vs)*4 and ceil(2T/(8 vs))*4 in exact arithmetic; the FP32 quotient
    # can only disagree when it lies within one ulp of an integer
    q = 2 * T32 / vs32
    qb = 2 * T32 / (8 * vs32)
    if abs(q - round(q)) > Fr(1, 1 << 20):
        assert d["max_vox"] == 4 * -(-q.numerator // q.denominator)
    if abs(qb - round(qb)) > Fr(1, 1 << 20):
        assert d["max_blk"] == 4 * -(-qb.numerator // qb.denominator)
    assert d["BL"] == float(f32(vs) * f32(8))
    assert d["half_BL"] * 2 == d["BL"] and d["blk_eps"] * 4 == d["BL"] and d["vox_eps"] * 4 == float(f32(vs))


def test_derived_constants_at_workloads():
    assert oracle.derived([0.5, 0.1, 1000, 2500, 250000])["max_vox"] == 4
    assert oracle.derived([0.5, 0.1, 1000, 2500, 250000])["max_blk"] == 4
    assert oracle.derived([0.1, 0.4, 1000, 2500, 22500])["max_vox"] == 32
    assert oracle.derived([0.1, 0.4, 1000, 2500, 22500])["max_blk"] == 4


# ---- cell centre / floor index (H:65-85) ----------------------------------------------
@pytest.mark.parametrize("L", [0.1, 0.5, 0.8, 4.0])
def test_cell_center_sweep(L):
    """x lies in cell k of the grid of spacing G = 1/fl(1/L) (exact rational
    test), except where fl(x * fl(1/L)) rounds across an integer, i.e. within
    one ulp of x*s of a cell boundary; the centre is the cell midpoint."""
    rng = np.random.default_rng(1)
    L32 = f32(L)
    sc = Fr(float(f32(1) / L32))
    G = 1 / sc
    xs = np.concatenate([rng.uniform(-600, 600, 2000), rng.uniform(-3, 3, 2000),
                         np.arange(-3000, 3000, dtype=np.float32) * f32(0.1)]).astype(np.float32)
    n_edge = 0
    for i, x in enumerate(xs):
        c, k = oracle.cell_center(float(x), float(L32))
        xr = Fr(float(x))
        inside = k * G <= xr < (k + 1) * G
        if not inside:
            ulp = Fr(float(abs(np.spacing(f32(x) * f32(float(sc))))))
            edge = min(abs(xr - k * G), abs(xr - (k + 1) * G))
            assert edge * sc <= ulp, (x, k)
            n_edge += i < 4000  # random samples; the last 6000 sit on boundaries by design
        assert abs(Fr(float(c)) - (k * G + G / 2)) <= (abs(k) + 2) * G * Fr(1, 1 << 22)
    assert n_edge < 40


def test_cell_center_exact_scales():
    # 1/L is exact in FP32 at the workload sizes, so centres depend on k only
    for L in (0.1, 0.5, 0.8, 4.0):
        s = f32(1) / f32(L)
        assert s in (f32(10), f32(2), f32(1.25), f32(0.25))


# ---- truncation segment (H:112-133) -------------------------------------------------------
def test_segment_properties():
    rng = np.random.default_rng(2)
    for _ in range(500):
        s = rng.uniform(-30, 30, 3).astype(np.float32)
        p = (s + rng.normal(0, 20, 3)).astype(np.float32)
        T = float(rng.choice([0.1, 0.4, 2.0]))
        a, b = oracle.segment(p, s, T)
        v = (p.astype(np.float64) - s) / np.linalg.norm(p.astype(np.float64) - s)
        assert np.linalg.norm(b.astype(np.float64) - a) == pytest.approx(2 * T, rel=1e-4, abs=1e-5)
        assert np.linalg.norm(a.astype(np.float64) - s) <= np.linalg.norm(b.astype(np.float64) - s)
        # a lies before p along the ray, b beyond it
        assert np.allclose(a, p - T * v, atol=1e-4) and np.allclose(b, p + T * v, atol=1e-4)


def test_segment_axis_aligned_exact():
    a, b = oracle.segment([5.0, 1.0, 2.0], [1.0, 1.0, 2.0], 0.25)
    assert list(a) == [4.75, 1.0, 2.0] and list(b) == [5.25, 1.0, 2.0]
    a, b = oracle.segment([1.0, 1.0, -3.0], [1.0, 1.0, 2.0], 0.5)  # ray towards -z
    assert list(a) == [1.0, 1.0, -2.5] and list(b) == [1.0, 1.0, -3.5]


# ---- DDA (H:146-245) ---------------------------------------------------------------------------
def exact_cells(a, b, L):
    """Cells crossed by the segment a->b, by exact rational arithmetic.
    Returns (cells, min gap between distinct crossing parameters)."""
    A = [Fr(float(x)) for x in a]
    B = [Fr(float(x)) for x in b]
    Lr = Fr(float(f32(L)))
    k = [int((A[i] / Lr).__floor__()) for i in range(3)]
    events = []
    for i in range(3):
        d = B[i] - A[i]
        if d == 0:
            continue
        kb = int((B[i] / Lr).__floor__())
        step = 1 if d > 0 else -1
        j = k[i]
        while j != kb:
            plane = (j + 1) * Lr if step > 0 else j * Lr
            events.append(((plane - A[i]) / d, i, step))
            j += step
    events.sort()
    cells = [tuple(k)]
    cur = list(k)
    for t, i, st in events:
        cur[i] += st
        cells.append(tuple(cur))
    ts = [e[0] for e in events]
    gap = min((ts[i + 1] - ts[i] for i in range(len(ts) - 1)), default=Fr(1))
    return cells, gap


def test_dda_axis_aligned_kat():
    # along +x through 9 cells of 0.1: 2T/L + 1 with T = 0.4
    keys, cen, tr = oracle.traverse([0.13, 0.27, 0.33], [0.93, 0.27, 0.33], 0.1, 32)
    assert not tr
    assert [tuple(k) for k in keys] == [(i, 2, 3) for i in range(1, 10)]


def test_dda_diagonal_tie_kat():
    # exact ties of tMax step two axes at once (the reference's tie branches)
    keys, _, tr = oracle.traverse([0.25, 0.25, 0.25], [3.75, 3.75, 0.25], 0.5, 32)
    assert [tuple(k) for k in keys] == [(i, i, 0) for i in range(8)]
    keys, _, _ = oracle.traverse([0.25, 0.25, 0.25], [3.75, 3.75, 3.75], 0.5, 32)
    assert [tuple(k) for k in keys] == [(i, i, i) for i in range(8)]


def test_dda_negative_direction_kat():
    keys, _, _ = oracle.traverse([-0.05, 0.05, 0.05], [-0.75, 0.05, 0.05], 0.1, 32)
    assert [tuple(k) for k in keys] == [(-i, 0, 0) for i in range(1, 9)]


def test_dda_truncation_bound():
    keys, _, tr = oracle.traverse([0.05, 0.05, 0.05], [5.05, 0.05, 0.05], 0.1, 32)
    assert tr and len(keys) == 32 and tuple(keys[-1]) == (31, 0, 0)


@pytest.mark.parametrize("L,T", [(0.1, 0.4), (0.5, 0.1), (0.8, 0.4)])
def test_dda_matches_exact_brute_force(L, T):
    rng = np.random.default_rng(int(L * 100) + 7)
    n_cmp = 0
    n_excused = 0
    for _ in range(600):
        s = rng.uniform(-20, 20, 3).astype(np.float32)
        p = (s + rng.normal(0, 15, 3)).astype(np.float32)
        a, b = oracle.segment(p, s, T)
        keys, _, tr = oracle.traverse(a, b, L, 128)
        got = [tuple(int(x) for x in k) for k in keys]
        ref, gap = exact_cells(a, b, L)
        # consecutive cells move by one cell on every stepped axis
        for u, v in zip(got, got[1:]):
            d = [abs(v[i] - u[i]) for i in range(3)]
            assert max(d) == 1
        if got == ref:
            n_cmp += 1
            continue
        # disagreements only where FP32 stepping cannot resolve the exact order:
        # crossings closer than the accumulated FP32 error, or an end point
        # within FP32 rounding of a cell boundary
        near = min(abs(Fr(float(x)) / Fr(float(f32(L))) - round(Fr(float(x)) / Fr(float(f32(L)))))
                   for x in list(a) + list(b))
        assert gap < Fr(1, 1 << 12) or near < Fr(1, 1 << 12), (a, b, got, ref)
        n_excused += 1
    assert n_excused <= 6, n_excused
    assert n_cmp >= 590


# ---- voxel -> block / local index (H:612, H:622-626, H:507-512) ------------------------------
@pytest.mark.parametrize("vs", [0.1, 0.5, 0.25])
def test_voxel_to_block_closed_form(vs):
    params = [vs, 0.4, 1000, 2500, 22500]
    for k in list(range(-1100, 1100, 7)) + list(range(-20, 20)):
        # a centre of voxel k on each axis (centres from the oracle's own cell centre)
        c, kk = oracle.cell_center(float((k + 0.5) * f32(vs)), vs)
        assert kk == k
        ok, kb, loc = oracle.voxel_to_block(params, [c, c, c])
        assert ok
        assert list(kb) == [k // 8] * 3
        assert loc == (k % 8) * (1 + 8 + 64)
    # distinct axes: local index is x + 8 y + 64 z
    cx, _ = oracle.cell_center(float(3.5 * f32(vs)), vs)
    cy, _ = oracle.cell_center(float(9.5 * f32(vs)), vs)
    cz, _ = oracle.cell_center(float(-2.5 * f32(vs)), vs)
    ok, kb, loc = oracle.voxel_to_block(params, [cx, cy, cz])
    assert ok and list(kb) == [0, 1, -1] and loc == 3 + 8 * 1 + 64 * 5  # k = (3, 9, -3)


# ---- SDF update (H:562-593, H:643-654) --------------------------------------------------------
def test_signed_distance_closed_form():
    s = np.array([0, 0, 0], np.float32)
    p = np.array([10, 0, 0], np.float32)
    assert oracle.signed_distance(p, s, [9.75, 0, 0]) == pytest.approx(0.25)    # in front of the surface
    assert oracle.signed_distance(p, s, [10.25, 0, 0]) == pytest.approx(-0.25)  # behind it
    d = oracle.signed_distance(p, s, [9.5, 0.3, 0.4])
    assert d == pytest.approx(np.sqrt(0.25 + 0.09 + 0.16), rel=1e-7)


def test_fold_closed_forms():
    S, W = oracle.fold(0.0, 0.0, 0.3, 1000.0)
    assert S == pytest.approx(0.3) and W == pytest.approx(1 / 1.3)
    S2, W2 = oracle.fold(S, W, -0.1, 1000.0)
    w1, w2 = 1 / 1.3, 1 / 1.1
    assert S2 == pytest.approx((w1 * 0.3 + w2 * -0.1) / (w1 + w2)) and W2 == pytest.approx(w1 + w2)
    # the clamp: the mean uses the unclamped W + w, the stored weight is capped
    S3, W3 = oracle.fold(0.2, 999.9, 0.0, 1000.0)
    assert W3 == 1000.0 and S3 == pytest.approx(999.9 * 0.2 / (999.9 + 1.0))
    # repeated identical updates keep the value and saturate the weight
    S, W = 0.0, 0.0
    for _ in range(3000):
        S, W = oracle.fold(S, W, 0.05, 1000.0)
    assert S == pytest.approx(0.05) and W == 1000.0


# ---- map-level invariants ----------------------------------------------------------------------
def _single_point_map(p, s, params):
    om = oracle.OracleMap(params)
    pts = np.array([[*p, 1.0]], np.float32)
    st = om.integrate(pts, s)
    return om, st


def test_single_point_update():
    params = [0.1, 0.4, 1000, 2500, 22500]
    p, s = [12.34, -3.21, -1.98], [20.0, 0.0, 0.0]
    om, st = _single_point_map(p, s, params)
    a, b = oracle.segment(p, s, 0.4)
    vk, vc, _ = oracle.traverse(a, b, 0.1, 32)
    bk, _, _ = oracle.traverse(a, b, 0.8, 4)
    keys, sdf, w = om.export()
    blocks = {tuple(k) for k in bk} | {(k[0] // 8, k[1] // 8, k[2] // 8) for k in vk}
    assert {tuple(k) for k in keys} == blocks
    assert st["n_voxel_updates"] == len(vk) == int((w != 0).sum())
    key_index = {tuple(k): i for i, k in enumerate(keys)}
    for kv, c in zip(vk, vc):
        i = key_index[(kv[0] // 8, kv[1] // 8, kv[2] // 8)]
        loc = (kv[0] % 8) + 8 * (kv[1] % 8) + 64 * (kv[2] % 8)
        d = oracle.signed_distance(p, s, c)
        assert sdf[i, loc] == pytest.approx(d, abs=1e-12)   # one update: sdf = d
        assert w[i, loc] == pytest.approx(1 / (abs(d) + 1))  # weight = 1/(|d|+1)
        assert abs(d) <= 0.4 * 1.8 + 0.1                    # inside the truncation band
    om.close()


def _scan(n=400, seed=0):
    from ros_tsdf_amd import configs, synth
    w = configs.SMALL
    p, s, _ = synth.make_scan(w, seed, n_points=n)
    return p, s


def test_publish_and_gc_brute_force():
    params = [0.5, 0.1, 1000.0, 400.0, 900.0]
    om = oracle.OracleMap(params)
    p, s = _scan()
    om.integrate(p, s)
    keys, sdf, w = om.export()
    pos, psdf, pw = om.publish(s)
    # brute force: block centres (k + 1/2) * 4 m within 20 m, voxels with weight
    cen = (keys.astype(np.float64) + 0.5) * 4.0
    d2 = ((cen - s.astype(np.float64)) ** 2).sum(1)
    assert not np.any(np.abs(d2 - 400.0) < 1e-3)
    inr = d2 <= 400.0
    expect = int((w[inr] != 0).sum())
    assert pos.shape[0] == expect
    # published positions are voxel centres of published blocks
    vk = np.floor(pos.astype(np.float64) / 0.5).astype(np.int64)
    assert np.allclose(pos, (vk + 0.5) * 0.5, atol=1e-4)
    assert set(map(tuple, vk // 8)) <= set(map(tuple, keys[inr]))
    assert np.all(pw > 0)
    removed = om.gc(s)
    assert removed == int((d2 > 900.0).sum())
    keys2, _, _ = om.export()
    assert {tuple(k) for k in keys2} == {tuple(k) for k in keys[d2 <= 900.0]}
    om.close()


def test_publish_and_gc_radius_boundaries():
    """withinDistanceSquared is `<=` (H:750-757): a block centred exactly on the publish radius is published
    (H:812), and one exactly on the GC radius is kept, since GC removes only blocks outside it (H:847)."""
    om = oracle.OracleMap(RADIUS_PARAMS)
    om.integrate(RADIUS_POINTS, RADIUS_SENSOR)
    keys, _, _ = om.export()
    assert sorted(map(tuple, keys.tolist())) == [(1, 0, 0), (2, 0, 0), (3, 0, 0)]
    pos, _, _ = om.publish(RADIUS_SENSOR)
    assert sorted({tuple(int(c) for c in np.floor(p / 4.0)) for p in pos.astype(np.float64)}) == [(1, 0, 0), (2, 0, 0)]
    assert om.gc(RADIUS_SENSOR) == 1
    keys, _, _ = om.export()
    assert sorted(map(tuple, keys.tolist())) == [(1, 0, 0), (2, 0, 0)]
    om.close()


def test_invalid_points_skipped():
    params = [0.5, 0.1, 1000, 2500, 250000]
    om = oracle.OracleMap(params)
    pts = np.array([[np.nan, 1, 1, 1], [0, 0, 0, 1], [np.inf, 0, 0, 1], [1e9, 0, 0, 1], [3.3, 4.4, -1.9, 1]],
                   np.float32)
    st = om.integrate(pts, np.zeros(3, np.float32))
    assert st["n_skipped"] == 4 and st["n_valid"] == 1
    assert oracle.point_valid([3.3, 4.4, -1.9], [0, 0, 0], 0.5)
    assert not oracle.point_valid([0, 0, 0], [0, 0, 0], 0.5)
    om.close()


def test_multithreaded_path_bit_identical():
    params = [0.1, 0.4, 1000, 2500, 22500]
    from ros_tsdf_amd import configs, synth
    a = oracle.OracleMap(params)
    b = oracle.OracleMap(params)
    for k, p, s in synth.scans(configs.DENSE, stop=2):
        p = p[:20000]
        sa = a.integrate(p, s)
        sb = b.integrate(p, s, threads=5)
        assert sa == sb
    ka, xa, wa = a.export()
    kb, xb, wb = b.export()
    assert np.array_equal(ka, kb) and np.array_equal(xa, xb) and np.array_equal(wa, wb)
    a.close()
    b.close()
