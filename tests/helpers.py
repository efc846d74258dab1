"""Shared comparison helpers: engine state vs the FP64 oracle."""
import numpy as np

from ros_tsdf_amd import configs


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
