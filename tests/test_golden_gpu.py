"""GPU: the engine against the closed-form golden cases (tests/golden/closed_form.json), to the parity bar:
block keys and touched voxels exact, SDF within 1e-4 T, weight within 1e-5 relative."""
import numpy as np
import pytest

from ros_tsdf_amd import configs
from tests.helpers import compare_publish_to_golden, compare_to_golden, golden_points, load_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", load_golden(), ids=lambda c: c["name"])
def test_engine_reproduces_closed_forms(case):
    from ros_tsdf_amd import Engine
    params = case["params"]
    T = params[1]
    with Engine(heap_blocks=1 << 10, total_buckets=1 << 10, max_batch_points=1 << 10) as e:
        m = e.map(params, buckets_log2=6)
        buf = e.buffer(1 << 12)
        for scan in case["scans"]:
            pts, s = golden_points(scan)
            buf.upload(pts)
            m.integrate(buf, pts.shape[0], s)
        k, sdf, w = m.export()
        compare_to_golden(k, sdf, w, case["expected"], T, configs.SDF_TOL_T, configs.WEIGHT_TOL_REL)
        if "publish_sensor" in case:
            pos, data, n = m.publish(np.asarray(case["publish_sensor"], np.float32))
            compare_publish_to_golden(pos, data[:, 0], data[:, 1], case["expected"], T, configs.SDF_TOL_T,
                                      configs.WEIGHT_TOL_REL)
        if "gc_sensor" in case:
            assert m.gc(np.asarray(case["gc_sensor"], np.float32)) == case["expected"]["gc_removed"]
        buf.free()
        m.close()
