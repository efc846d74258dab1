"""Materialised workloads: scans generated in parallel, batch plans and
serving request streams built from them."""
from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import configs, synth


def _gen_one(args):
    w, k, map_id, n = args
    boxes = synth.make_boxes(w.scene, w.seed, map_id)
    p, s, fill = synth.make_scan(w, k, map_id=map_id, n_points=n, boxes=boxes)
    return p, s, fill


def generate(w: synth.Workload, ks, map_id: int = 0, n_points=None, workers: int | None = None):
    """Scans ks of one map as (points list, sensors array, fill fractions)."""
    ks = list(ks)
    workers = workers or min(16, os.cpu_count() or 1, max(1, len(ks)))
    jobs = [(w, k, map_id, n_points) for k in ks]
    if workers <= 1 or len(ks) <= 1:
        out = [_gen_one(j) for j in jobs]
    else:
        import multiprocessing as mp
        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork")) as ex:
            out = list(ex.map(_gen_one, jobs, chunksize=max(1, len(jobs) // (4 * workers))))
    pts = [o[0] for o in out]
    sens = np.stack([o[1] for o in out]) if out else np.zeros((0, 3), np.float32)
    return pts, sens, [o[2] for o in out]


def generate_maps(base: synth.Workload, n_maps: int, scans_per_map: int, workers: int | None = None):
    """Full-size scans 0..scans_per_map-1 of n_maps maps with their own scenes."""
    jobs = []
    for m in range(n_maps):
        w = configs.map_workload(base, m)
        for k in range(scans_per_map):
            jobs.append((w, k, m, None))
    workers = workers or min(16, os.cpu_count() or 1)
    import multiprocessing as mp
    if workers <= 1:
        out = [_gen_one(j) for j in jobs]
    else:
        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork")) as ex:
            out = list(ex.map(_gen_one, jobs, chunksize=max(1, len(jobs) // (4 * workers))))
    pts = [[None] * scans_per_map for _ in range(n_maps)]
    sens = np.zeros((n_maps, scans_per_map, 3), np.float32)
    for (w, k, m, _), o in zip(jobs, out):
        pts[m][k] = o[0]
        sens[m, k] = o[1]
    return pts, sens


class Pool:
    """All points of a set of scans packed in one array (one device upload)."""

    def __init__(self, scans):
        """scans: list of (points float32 (n,4), sensor (3,), map_index)."""
        self.offsets = []
        self.counts = []
        self.sensors = []
        self.maps = []
        off = 0
        for p, s, m in scans:
            self.offsets.append(off)
            self.counts.append(int(p.shape[0]))
            self.sensors.append(np.asarray(s, np.float32))
            self.maps.append(int(m))
            off += int(p.shape[0])
        self.points = np.concatenate([p for p, _, _ in scans], axis=0) if scans else np.zeros((0, 4), np.float32)
        self.total = off

    def __len__(self):
        return len(self.counts)


def batch_plan(n_maps=configs.BATCH_MAPS, scans_per_batch=configs.BATCH_SCANS, n_batches=4, seed=7,
               scans_per_map=None):
    """Batches of scans drawn from n_maps maps: each batch holds
    scans_per_batch / n_maps consecutive scans of every map, with log-uniform
    point counts.  Returns a list of batches, each a list of (map, scan_k, n)."""
    rng = np.random.default_rng(seed)
    per_map = scans_per_batch // n_maps
    nxt = [0] * n_maps
    plan = []
    for _ in range(n_batches):
        b = []
        for m in range(n_maps):
            for _ in range(per_map):
                k = nxt[m]
                nxt[m] += 1
                if scans_per_map:
                    k %= scans_per_map
                b.append((m, k, int(configs.log_uniform_counts(rng, 1)[0])))
        plan.append(b)
    return plan


def poisson_requests(n_requests, rate, n_maps, seed=5, counts="loguniform", scans_per_map=None):
    """(arrival_s, map, scan_k, n) for a Poisson stream over n_maps maps; each
    map's requests carry its scans in order."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    arr = np.cumsum(gaps)
    arr -= arr[0]
    maps = rng.integers(0, n_maps, size=n_requests)
    if counts == "loguniform":
        ns = configs.log_uniform_counts(rng, n_requests)
    else:
        ns = np.full(n_requests, int(counts), np.int64)
    nxt = [0] * n_maps
    out = []
    for i in range(n_requests):
        m = int(maps[i])
        k = nxt[m]
        nxt[m] += 1
        if scans_per_map:
            k %= scans_per_map
        out.append((float(arr[i]), m, k, int(ns[i])))
    return out
