"""Seeded synthetic lidar workloads: ground plane + axis-aligned boxes, a sensor
on a circle, and ray-cast returns with Gaussian range noise.

The reference ships no datasets (its benchmarks used AirSim recordings that are
not in the repository, README.md:107-113), so every workload here is generated.
Open choices of the workload descriptions are fixed as follows:

* ground plane z = -2 m; boxes rest on the ground, sides U[1, 5] m per axis,
  centres uniform in the square; a box whose footprint comes within 1 m of the
  sensor circle is redrawn;
* the sensor sits at z = 0 on a circle about the origin and makes one
  revolution over the run (scan k at angle 2*pi*k/n_scans), 4 Hz time stamps;
* rays are uniform in azimuth and in elevation over +-15 deg ("random"
  pattern), or 64 beams x 2048 azimuths ("grid" pattern);
* noise is applied along the ray; returns beyond max range are misses;
* misses are not placed at max range: uniformly random rays are drawn until
  the scan holds exactly the requested number of returns;
* the per-scan random stream is numpy's PCG64 seeded with
  (seed, map_id, scan_index), so any scan can be generated on its own.

Points are float32 (x, y, z, 1) rows, the 16-byte pcl::PointXYZ layout the
reference integrates (tsdf_handler.cu:977-986).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GROUND_Z = -2.0


@dataclass(frozen=True)
class SceneSpec:
    square: float = 100.0          # side of the square the box centres are drawn in (m)
    n_boxes: int = 20
    side_min: float = 1.0
    side_max: float = 5.0
    radius: float = 20.0           # sensor circle radius (m)
    sensor_z: float = 0.0


@dataclass(frozen=True)
class LidarSpec:
    pattern: str = "random"        # "random" or "grid"
    n_points: int = 1000
    beams: int = 64
    azimuths: int = 2048
    elev_deg: float = 15.0
    max_range: float = 50.0
    sigma: float = 0.02


@dataclass(frozen=True)
class Workload:
    scene: SceneSpec = field(default_factory=SceneSpec)
    lidar: LidarSpec = field(default_factory=LidarSpec)
    n_scans: int = 371
    seed: int = 0


def make_boxes(spec: SceneSpec, seed: int, map_id: int = 0) -> np.ndarray:
    """Return (n_boxes, 6) float64 [xmin, ymin, zmin, xmax, ymax, zmax]."""
    rng = np.random.default_rng([seed, map_id, 0xB0CE5])
    out = []
    while len(out) < spec.n_boxes:
        cx, cy = rng.uniform(-spec.square / 2, spec.square / 2, size=2)
        sx, sy, sz = rng.uniform(spec.side_min, spec.side_max, size=3)
        x0, x1 = cx - sx / 2, cx + sx / 2
        y0, y1 = cy - sy / 2, cy + sy / 2
        # distance range from the origin to the footprint grown by 1 m
        gx0, gx1, gy0, gy1 = x0 - 1, x1 + 1, y0 - 1, y1 + 1
        nx = min(max(0.0, gx0), gx1) if not (gx0 <= 0 <= gx1) else 0.0
        ny = min(max(0.0, gy0), gy1) if not (gy0 <= 0 <= gy1) else 0.0
        dmin = np.hypot(nx, ny)
        dmax = max(np.hypot(fx, fy) for fx in (gx0, gx1) for fy in (gy0, gy1))
        if dmin <= spec.radius <= dmax:
            continue
        out.append([x0, y0, GROUND_Z, x1, y1, GROUND_Z + sz])
    return np.asarray(out, dtype=np.float64)


def sensor_pose(spec: SceneSpec, k: int, n_scans: int, phase: float = 0.0) -> np.ndarray:
    th = 2.0 * np.pi * k / n_scans + phase
    return np.array([spec.radius * np.cos(th), spec.radius * np.sin(th), spec.sensor_z], dtype=np.float64)


def _wrap(a):
    return (a + np.pi) % (2.0 * np.pi) - np.pi


def ray_cast(origin: np.ndarray, dirs: np.ndarray, boxes: np.ndarray) -> np.ndarray:
    """Distance along unit rays to the first surface (inf for a miss).

    Rays are culled per box by the azimuth interval its footprint subtends
    (the sensor is never inside a footprint), then slab-tested."""
    t = np.full(dirs.shape[0], np.inf)
    az = np.arctan2(dirs[:, 1], dirs[:, 0])
    with np.errstate(divide="ignore", invalid="ignore"):
        dz = dirs[:, 2]
        tp = (GROUND_Z - origin[2]) / dz
        ok = (dz < 0) & (tp > 0)
        t = np.where(ok, tp, t)
        for b in boxes:
            cx, cy = 0.5 * (b[0] + b[3]) - origin[0], 0.5 * (b[1] + b[4]) - origin[1]
            c = np.arctan2(cy, cx)
            corners = np.array([[b[0], b[1]], [b[0], b[4]], [b[3], b[1]], [b[3], b[4]]]) - origin[:2]
            dl = _wrap(np.arctan2(corners[:, 1], corners[:, 0]) - c)
            rel = _wrap(az - c)
            idx = np.nonzero((rel >= dl.min() - 1e-9) & (rel <= dl.max() + 1e-9))[0]
            if idx.size == 0:
                continue
            d = dirs[idx]
            inv = 1.0 / d
            t1 = (b[:3] - origin) * inv
            t2 = (b[3:] - origin) * inv
            tn = np.nanmax(np.minimum(t1, t2), axis=1)
            tf = np.nanmin(np.maximum(t1, t2), axis=1)
            hit = (tn <= tf) & (tn > 0) & (tn < t[idx])
            t[idx[hit]] = tn[hit]
    return t


def _dirs(az: np.ndarray, el: np.ndarray) -> np.ndarray:
    ce = np.cos(el)
    return np.stack([ce * np.cos(az), ce * np.sin(az), np.sin(el)], axis=1)


def make_scan(w: Workload, k: int, map_id: int = 0, n_points: int | None = None,
              boxes: np.ndarray | None = None, phase: float = 0.0):
    """Scan k of map `map_id`: returns (points float32 (n,4), sensor float32 (3,), fill_fraction)."""
    lid = w.lidar
    n = lid.n_points if n_points is None else int(n_points)
    if boxes is None:
        boxes = make_boxes(w.scene, w.seed, map_id)
    s = sensor_pose(w.scene, k, w.n_scans, phase)
    rng = np.random.default_rng([w.seed, map_id, 1 + k])
    el_max = np.deg2rad(lid.elev_deg)
    pts = []
    have = 0
    n_fill = 0
    if lid.pattern == "grid":
        el = np.repeat(np.linspace(-el_max, el_max, lid.beams), lid.azimuths)
        az = np.tile(np.arange(lid.azimuths) * (2.0 * np.pi / lid.azimuths), lid.beams)
        d = _dirs(az, el)
        t = ray_cast(s, d, boxes)
        hit = t <= lid.max_range
        d, t = d[hit], t[hit]
        if d.shape[0] > n:
            d, t = d[:n], t[:n]
        tn = np.maximum(t + rng.normal(0.0, lid.sigma, size=t.shape), 1e-2)
        pts.append(s + tn[:, None] * d)
        have = d.shape[0]
    while have < n:
        m = max(256, 2 * (n - have))
        az = rng.uniform(0.0, 2.0 * np.pi, size=m)
        el = rng.uniform(-el_max, el_max, size=m)
        d = _dirs(az, el)
        t = ray_cast(s, d, boxes)
        hit = t <= lid.max_range
        d, t = d[hit][: n - have], t[hit][: n - have]
        tn = np.maximum(t + rng.normal(0.0, lid.sigma, size=t.shape), 1e-2)
        pts.append(s + tn[:, None] * d)
        have += d.shape[0]
        n_fill += d.shape[0]
    p = np.concatenate(pts, axis=0) if pts else np.zeros((0, 3))
    out = np.ones((n, 4), dtype=np.float32)
    out[:, :3] = p[:n].astype(np.float32)
    return out, s.astype(np.float32), (n_fill / n if n else 0.0)


def scans(w: Workload, map_id: int = 0, start: int = 0, stop: int | None = None, phase: float = 0.0):
    """Yield (k, points, sensor) for scans start..stop-1 of one map."""
    boxes = make_boxes(w.scene, w.seed, map_id)
    stop = w.n_scans if stop is None else stop
    for k in range(start, stop):
        p, s, _ = make_scan(w, k, map_id=map_id, boxes=boxes, phase=phase)
        yield k, p, s
