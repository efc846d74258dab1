"""Writes closed_form.json: integration cases whose expected results follow from closed forms.

The expected values are computed here from the geometry of each case, in float64. Neither the oracle nor the
engine is called, so the fixtures check both of them.

Every case has axis-aligned rays, so that:
- the unit direction is exact (sqrt(x*x) == |x| in binary IEEE arithmetic);
- the truncation segment is [p - T, p + T] along the ray axis, in FP32;
- the voxels walked are the cells floor(x / vs) between the segment ends, with the other two indices fixed;
- the blocks are the voxels' cells >> 3, together with the block walk's cells floor(x / 8vs);
- the signed distance is |p - c| for voxel centre c = (k + 1/2) vs, negated when the centre lies beyond the
  point along the ray (H:562-573);
- the weight is 1 / (|d| + 1) (H:579-582), folded per voxel in point order as S <- (W S + w d) / (W + w),
  W <- min(W + w, maxW) (H:648-654).

Segment ends are placed well inside cells, so FP32 rounding cannot move a floor.

Publish (H:767-790) emits every touched voxel of blocks within the radius, at the voxel centres. GC
(H:836-868) removes the blocks whose centre lies beyond the GC radius.

Run:  python tests/golden/make_golden.py   (rewrites tests/golden/closed_form.json)
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "closed_form.json")

DENSE = [0.1, 0.4, 1000.0, 2500.0, 22500.0]


def f32(x):
    return float(np.float32(x))


def axis_ray_cells(p, s, params):
    """Voxel keys and block keys of one point whose ray from s runs along one axis."""
    vs, T = params[0], params[1]
    p32 = [f32(v) for v in p]
    diff = [p32[i] - f32(s[i]) for i in range(3)]
    ax = [i for i in range(3) if diff[i] != 0.0]
    assert len(ax) == 1, "the ray must run along one axis"
    ax = ax[0]
    sign = 1.0 if diff[ax] > 0 else -1.0
    # segment ends in FP32: p -+ T along the ray (the nearer end first)
    a = f32(np.float32(p32[ax]) - np.float32(sign) * np.float32(T))
    b = f32(np.float32(p32[ax]) + np.float32(sign) * np.float32(T))
    sc = f32(np.float32(1.0) / np.float32(vs))
    sb = f32(np.float32(1.0) / (np.float32(vs) * np.float32(8)))
    for x in (a, b):
        for scale in (sc, sb):
            frac = (x * scale) - math.floor(x * scale)
            assert 0.05 < frac < 0.95, "segment end too close to a cell boundary"
    k_other = [math.floor(p32[i] * sc) for i in range(3)]
    ka, kb_ = math.floor(a * sc), math.floor(b * sc)
    step = 1 if kb_ >= ka else -1
    voxels = []
    for k in range(ka, kb_ + step, step):
        kk = list(k_other)
        kk[ax] = k
        voxels.append(kk)
    blocks = {tuple(v >> 3 for v in kk) for kk in voxels}
    ba, bb = math.floor(a * sb), math.floor(b * sb)
    for kbx in range(min(ba, bb), max(ba, bb) + 1):
        kk = [v >> 3 for v in k_other]
        kk[ax] = kbx
        blocks.add(tuple(kk))
    return voxels, blocks, ax, sign


def signed_distance(p, s, kk, vs, ax, sign):
    c = [(k + 0.5) * vs for k in kk]
    d = math.sqrt(sum((p[i] - c[i]) ** 2 for i in range(3)))
    beyond = (c[ax] - p[ax]) * sign > 0
    return -d if beyond else d


def integrate(case):
    """Closed-form voxel states after all scans of a case: {voxel key: [S, W]} and block keys."""
    params = case["params"]
    vs, maxW = params[0], params[2]
    state = {}
    blocks = set()
    for scan in case["scans"]:
        s = scan["sensor"]
        for p in scan["points"]:
            voxels, bl, ax, sign = axis_ray_cells(p, s, params)
            blocks |= bl
            p32 = [f32(v) for v in p]
            for kk in voxels:
                d = signed_distance(p32, s, kk, vs, ax, sign)
                w = 1.0 / (abs(d) + 1.0)
                S, W = state.get(tuple(kk), (0.0, 0.0))
                S = (W * S + w * d) / (W + w)
                W = min(W + w, maxW)
                state[tuple(kk)] = (S, W)
    return state, blocks


def case_records(case):
    state, blocks = integrate(case)
    vox = [[list(k), S, W] for k, (S, W) in sorted(state.items())]
    out = {"blocks": sorted(list(b) for b in blocks), "voxels": vox}
    if "publish_sensor" in case:
        s = case["publish_sensor"]
        vs = case["params"][0]
        pub = []
        for k, (S, W) in sorted(state.items()):
            kb = [v >> 3 for v in k]
            c = [(v + 0.5) * 8 * vs for v in kb]
            if sum((s[i] - c[i]) ** 2 for i in range(3)) <= case["params"][3]:
                pub.append([[(v + 0.5) * vs for v in k], S, W])
        out["published"] = pub
    if "gc_sensor" in case:
        s = case["gc_sensor"]
        vs = case["params"][0]
        far = 0
        for kb in blocks:
            c = [(v + 0.5) * 8 * vs for v in kb]
            if sum((s[i] - c[i]) ** 2 for i in range(3)) > case["params"][4]:
                far += 1
        out["gc_removed"] = far
    return out


CASES = [
    {"name": "one point, ray along +x",
     "params": DENSE,
     "scans": [{"sensor": [0.0, 0.262, 0.271], "points": [[10.137, 0.262, 0.271]]}],
     "publish_sensor": [0.0, 0.262, 0.271], "gc_sensor": [200.0, 0.0, 0.0]},
    {"name": "one point, ray along -y, negative cells",
     "params": DENSE,
     "scans": [{"sensor": [0.337, 0.0, 0.155], "points": [[0.337, -5.128, 0.155]]}],
     "publish_sensor": [0.0, 0.0, 0.0], "gc_sensor": [0.0, -200.0, 0.0]},
    {"name": "two points on one ray: weighted mean in point order",
     "params": DENSE,
     "scans": [{"sensor": [0.0, 0.262, 0.271], "points": [[10.137, 0.262, 0.271], [10.183, 0.262, 0.271]]}]},
    {"name": "two scans from opposite sides along z",
     "params": DENSE,
     "scans": [{"sensor": [1.234, -2.345, 20.0], "points": [[1.234, -2.345, 3.117]]},
               {"sensor": [1.234, -2.345, -20.0], "points": [[1.234, -2.345, 3.163]]}]},
    {"name": "weight clamp at max_weight",
     "params": [0.1, 0.4, 2.5, 2500.0, 22500.0],
     "scans": [{"sensor": [0.0, 0.262, 0.271], "points": [[10.137, 0.262, 0.271]] * 5}]},
    {"name": "sub-voxel truncation (0.5 m voxels, 0.1 m truncation)",
     "params": [0.5, 0.1, 1000.0, 2500.0, 250000.0],
     "scans": [{"sensor": [3.1, 0.0, 0.6], "points": [[3.1, 7.45, 0.6], [3.1, -6.14, 0.6]]}],
     "publish_sensor": [3.1, 0.0, 0.6]},
]


def build():
    return [dict(case, expected=case_records(case)) for case in CASES]


def main():
    with open(OUT, "w") as f:
        json.dump(build(), f, indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
