"""The oracle against the closed-form golden cases (tests/golden/closed_form.json).

The fixtures come from tests/golden/make_golden.py, which computes them from the geometry of axis-aligned rays
without calling the oracle. The oracle must reproduce:
- the block and touched-voxel sets exactly;
- SDF and weight to within the FP32 rounding of the voxel centres (1e-5 T);
- the published voxels;
- the GC count.
"""
import json
import os
import sys

import numpy as np
import pytest

from tests.helpers import compare_publish_to_golden, compare_to_golden, golden_points, load_golden

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fixture_matches_its_generator():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    try:
        import make_golden
    finally:
        sys.path.pop(0)
    regenerated = json.loads(json.dumps(make_golden.build()))
    assert regenerated == load_golden()


@pytest.mark.parametrize("case", load_golden(), ids=lambda c: c["name"])
def test_oracle_reproduces_closed_forms(case, oracle_lib):
    params = case["params"]
    T = params[1]
    om = oracle_lib.OracleMap(params)
    for scan in case["scans"]:
        pts, s = golden_points(scan)
        om.integrate(pts, s)
    k, sdf, w = om.export()
    compare_to_golden(k, sdf, w, case["expected"], T, 1e-5, 1e-6)
    if "publish_sensor" in case:
        pos, psdf, pw = om.publish(np.asarray(case["publish_sensor"], np.float32))
        compare_publish_to_golden(pos, psdf, pw, case["expected"], T, 1e-5, 1e-6)
    if "gc_sensor" in case:
        assert om.gc(np.asarray(case["gc_sensor"], np.float32)) == case["expected"]["gc_removed"]
    om.close()
