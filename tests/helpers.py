"""Shared comparison helpers: engine state vs the FP64 oracle."""
import numpy as np


def compare_maps(gmap, omap, T, sdf_tol_T, w_tol):
    """Block keys and touched voxels bit-exact; sdf within sdf_tol_T * T; weight within w_tol.

    Returns a dict of error figures."""
    gk, gs, gw = gmap.export()
    ok, os_, ow = omap.export()
    assert gk.shape == ok.shape, "block count %d vs oracle %d" % (gk.shape[0], ok.shape[0])
    assert np.array_equal(gk, ok), "block key sets differ"
    tg, to = gw != 0, ow != 0
    assert np.array_equal(tg, to), "touched voxel sets differ (%d vs %d)" % (tg.sum(), to.sum())
    out = dict(blocks=int(gk.shape[0]), voxels=int(tg.sum()), max_sdf_err_T=0.0, rms_sdf_err_T=0.0, max_w_err=0.0)
    if tg.any():
        e = np.abs(gs[tg].astype(np.float64) - os_[tg]) / T
        we = np.abs(gw[tg].astype(np.float64) - ow[tg])
        out.update(max_sdf_err_T=float(e.max()), rms_sdf_err_T=float(np.sqrt(np.mean(e * e))), max_w_err=float(we.max()))
        assert out["max_sdf_err_T"] <= sdf_tol_T, out
        assert out["max_w_err"] <= w_tol, out
    return out


def compare_publish(gpos, gdata, opos, osdf, ow, T, sdf_tol_T, w_tol):
    """Published voxel sets: positions bit-exact as sets, values within tolerance."""
    assert gpos.shape[0] == opos.shape[0], "published %d vs oracle %d" % (gpos.shape[0], opos.shape[0])
    if gpos.shape[0] == 0:
        return
    g = np.lexsort((gpos[:, 2], gpos[:, 1], gpos[:, 0]))
    o = np.lexsort((opos[:, 2], opos[:, 1], opos[:, 0]))
    assert np.array_equal(gpos[g].view(np.uint32), opos[o].view(np.uint32)), "published positions differ"
    e = np.abs(gdata[g, 0].astype(np.float64) - osdf[o]) / T
    assert e.max() <= sdf_tol_T, e.max()
    we = np.abs(gdata[g, 1].astype(np.float64) - ow[o])
    assert we.max() <= w_tol, we.max()


def load_golden():
    """The closed-form cases of tests/golden/closed_form.json (written by tests/golden/make_golden.py)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "closed_form.json")) as f:
        return json.load(f)


def golden_points(scan):
    pts = np.zeros((len(scan["points"]), 4), np.float32)
    pts[:, :3] = np.asarray(scan["points"], np.float32)
    return pts, np.asarray(scan["sensor"], np.float32)


def compare_to_golden(keys, sdf, w, expected, T, sdf_tol_T, w_tol_rel):
    """An exported map (keys [n,3], sdf/weight [n,512]) against a case's closed-form expectation."""
    exp_blocks = {tuple(b) for b in expected["blocks"]}
    got_blocks = {tuple(k) for k in np.asarray(keys).tolist()}
    assert got_blocks == exp_blocks, (sorted(got_blocks), sorted(exp_blocks))
    row = {tuple(k): i for i, k in enumerate(np.asarray(keys).tolist())}
    touched = set()
    for (k, S, W) in expected["voxels"]:
        kb = tuple(v >> 3 for v in k)
        loc = (k[0] & 7) + 8 * (k[1] & 7) + 64 * (k[2] & 7)
        touched.add((kb, loc))
        i = row[kb]
        assert abs(float(sdf[i][loc]) - S) <= sdf_tol_T * T, (k, float(sdf[i][loc]), S)
        assert abs(float(w[i][loc]) - W) <= w_tol_rel * W, (k, float(w[i][loc]), W)
    got_touched = {(tuple(np.asarray(keys)[i].tolist()), int(l)) for i in range(len(keys))
                   for l in np.nonzero(np.asarray(w)[i])[0]}
    assert got_touched == touched


def compare_publish_to_golden(pos, sdf, w, expected, T, sdf_tol_T, w_tol_rel, pos_tol=1e-5):
    exp = expected["published"]
    assert len(pos) == len(exp), (len(pos), len(exp))
    if not exp:
        return
    e_pos = np.asarray([p for p, _, _ in exp], np.float64)
    g = np.lexsort(np.asarray(pos, np.float64).T[::-1])
    o = np.lexsort(e_pos.T[::-1])
    assert np.abs(np.asarray(pos, np.float64)[g] - e_pos[o]).max() <= pos_tol
    e_s = np.asarray([s for _, s, _ in exp])[o]
    e_w = np.asarray([x for _, _, x in exp])[o]
    assert np.abs(np.asarray(sdf, np.float64)[g] - e_s).max() <= sdf_tol_T * T
    assert (np.abs(np.asarray(w, np.float64)[g] - e_w) <= w_tol_rel * e_w).all()


# Publish / GC radius boundary case: 4 m blocks, publish and GC radius 8 m.  One point in each of blocks
# (1,0,0), (2,0,0), (3,0,0), whose centres (6,2,2), (10,2,2), (14,2,2) lie exactly 4, 8 and 12 m from the
# sensor in FP32 (squared distances 16, 64, 144).
RADIUS_PARAMS = [0.5, 0.1, 1000.0, 64.0, 64.0]
RADIUS_SENSOR = np.array([2.0, 2.0, 2.0], np.float32)
RADIUS_POINTS = np.array([[6.25, 2.25, 2.25, 1], [10.25, 2.25, 2.25, 1], [14.25, 2.25, 2.25, 1]], np.float32)
