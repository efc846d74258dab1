"""Autotuner of the batched pipeline.

It measures the batched pipeline on a sample workload and picks the launch shapes and hash-table shape with the
highest throughput. The sample is a list of batches, each a list of scans of independent maps.

Knobs:
  traverse_block  threads per workgroup of the binning kernel (64, 128, 256)
  rays_per_lane   rays each lane of the binning kernel walks per chunk (1, 2, 4), i.e. 64, 128 or 256 rays per
                  wavefront: a chunk of traverse_block x rays_per_lane rays gathers its blocks in one LDS table
                  before touching the hash
  traverse_queue  0: one binning workgroup per chunk; 1: a persistent grid of resident workgroups taking chunks
                  from a device work queue
  fold_grid       workgroups of the first fold launch (work items are spread over them)
  buckets_log2    home buckets per map; a bucket holds 8 entries, so 17..21 spans 2^20..2^24 entries
  max_chain       longest overflow chain a map may grow (overflow-list length)

Each candidate is measured on fresh maps: one untimed warm-up replay of the first batch, then every batch
replayed through a HIP graph captured for that candidate. A candidate is discarded if any insert, heap
allocation or update failed on the sample. That keeps the choice from changing results: every surviving
configuration integrates the same scans into the same maps, bit for bit (tests/test_batch_serve_gpu.py).

The search is coordinate descent from DEFAULT: one knob at a time, in SPACE order, keeping the best value
found so far for the knobs already swept.
"""
from __future__ import annotations

import time

from .engine import BatchGraph, Engine

SPACE = {
    "traverse_block": (64, 128, 256),
    "rays_per_lane": (1, 2, 4),
    "traverse_queue": (0, 1),
    "fold_grid": (2048, 4096, 8192, 16384),
    "buckets_log2": (17, 18, 19, 20, 21),
    "max_chain": (4, 16, 64),
}
DEFAULT = {"traverse_block": 256, "rays_per_lane": 1, "traverse_queue": 0, "fold_grid": 4096, "buckets_log2": 17,
           "max_chain": 64}
LAUNCH_KNOBS = ("traverse_block", "rays_per_lane", "traverse_queue", "fold_grid")


def launch_kwargs(cfg: dict) -> dict:
    """The launch-shape knobs of a configuration, as Engine.set_launch arguments (absent ones at DEFAULT)."""
    return {k: cfg.get(k, DEFAULT[k]) for k in LAUNCH_KNOBS}


def measure(engine: Engine, params, nmaps: int, batches, points, cfg: dict, reps: int = 1) -> dict:
    """Batched throughput of one configuration.

    batches: lists of (map_index, n_points, point_offset, sensor), each batch holding at most one scan per map.
    points: a DeviceBuffer holding every batch's points.
    Returns a dict with points_per_s, ms_per_batch and failures (the summed failure counters).
    """
    engine.set_launch(**launch_kwargs(cfg))
    best = None
    failures = 0
    max_pts = max(sum(s[1] for s in b) for b in batches)
    max_scans = max(len(b) for b in batches)
    # failures: every batch once, on fresh maps, with the counters read after each
    maps = [engine.map(params, buckets_log2=cfg["buckets_log2"], max_chain=cfg["max_chain"]) for _ in range(nmaps)]
    g = BatchGraph(engine, maps, max_points=max_pts, max_scans=max_scans)
    try:
        for b in batches:
            g.launch(b, points)
            failures += _failures(engine.stats())
    finally:
        g.close()
        for m in maps:
            m.close()
    # throughput: fresh maps, warm-up on the first batch, the others timed
    dt = 0.0
    for _ in range(reps):
        maps = [engine.map(params, buckets_log2=cfg["buckets_log2"], max_chain=cfg["max_chain"])
                for _ in range(nmaps)]
        g = BatchGraph(engine, maps, max_points=max_pts, max_scans=max_scans)
        try:
            g.launch(batches[0], points)
            engine.synchronize()
            npts = 0
            t0 = time.perf_counter()
            for b in batches[1:]:
                g.launch(b, points)
                npts += sum(s[1] for s in b)
            engine.synchronize()
            dt = time.perf_counter() - t0
        finally:
            g.close()
            for m in maps:
                m.close()
        pps = npts / dt if dt > 0 else 0.0
        if best is None or pps > best:
            best = pps
    return {"points_per_s": best, "ms_per_batch": 1e3 * dt / max(1, len(batches) - 1), "failures": failures}


def _failures(st: dict) -> int:
    return int(st["insert_failures"] + st["heap_failures"] + st["updates_dropped"] + st["quota_failures"])


def tune(engine: Engine, params, nmaps: int, batches, points, space: dict | None = None, start: dict | None = None,
         reps: int = 1, log=None) -> tuple[dict, list]:
    """Coordinate descent over space (default SPACE) from start (default DEFAULT).

    Returns (best configuration, list of (configuration, measurement) in the order measured).
    """
    space = dict(SPACE if space is None else space)
    cur = dict(DEFAULT if start is None else start)
    seen = {}
    history = []

    def run(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            r = measure(engine, params, nmaps, batches, points, cfg, reps=reps)
            seen[key] = r
            history.append((dict(cfg), r))
            if log:
                log("autotune %s -> %.3e pts/s%s" % (cfg, r["points_per_s"],
                                                     "" if r["failures"] == 0 else " (%d failures)" % r["failures"]))
        return seen[key]

    def score(r):
        return r["points_per_s"] if r["failures"] == 0 else -1.0

    best_r = run(cur)
    for knob, values in space.items():
        for v in values:
            cand = dict(cur, **{knob: v})
            r = run(cand)
            if score(r) > score(best_r):
                best_r, cur = r, cand
    engine.set_launch(**launch_kwargs(cur))
    return cur, history
