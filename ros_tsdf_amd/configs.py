"""The five workloads the engine is measured and checked on.

Parameters are the five arguments of initGlobalVars (voxel size, truncation
distance, max weight, publish distance^2, GC distance^2).

* small      371 scans x 1,000 random-pattern points, 100 m scene, 50 m range,
             voxel 0.5 m, truncation 0.1 m (the reference README's first setting)
* dense      200 scans x 131,072 points (64 x 2048 grid pattern, misses
             replaced by random rays), 400 m scene, 120 m range, voxel 0.1 m,
             truncation 0.4 m, GC at 150 m, one scan per launch
* batched    32 scans per launch from 8 maps with their own scenes, point
             counts log-uniform in [1,024, 131,072]
* fp16       the dense workload with 4-byte packed voxel state
* serving    16 maps, Poisson arrivals at 400 scans/s, counts as batched
             (and a fixed-131,072 variant)

max_weight for the dense-family workloads and the publish radius (50 m) are
not stated by the workload descriptions; they inherit the small workload's
values.
"""
from __future__ import annotations

import numpy as np

from .synth import LidarSpec, SceneSpec, Workload

SMALL_PARAMS = (0.5, 0.1, 1000.0, 2500.0, 250000.0)
DENSE_PARAMS = (0.1, 0.4, 1000.0, 2500.0, 22500.0)

SMALL = Workload(scene=SceneSpec(square=100.0), lidar=LidarSpec(pattern="random", n_points=1000, max_range=50.0),
                 n_scans=371, seed=11)
DENSE = Workload(scene=SceneSpec(square=400.0), lidar=LidarSpec(pattern="grid", n_points=131072, max_range=120.0),
                 n_scans=200, seed=12)

BATCH_MAPS = 8
BATCH_SCANS = 32
BATCH_MIN_POINTS = 1024
BATCH_MAX_POINTS = 131072

SERVE_MAPS = 16
SERVE_RATE = 400.0            # scans per second, all maps together
SERVE_HEAP_BUDGET = 64 << 30  # bytes of voxel heap the maps are admitted against

# Parity tolerances of the FP32-state engine against the FP64 oracle.
SDF_TOL_T = 1e-4              # max |sdf - sdf_oracle| in units of truncation distance
WEIGHT_TOL_REL = 1e-5         # max |w - w_oracle| relative to max_weight


def map_workload(base: Workload, map_id: int) -> Workload:
    """Each batched / served map gets its own scene seed."""
    return Workload(scene=base.scene, lidar=base.lidar, n_scans=base.n_scans, seed=base.seed * 1000 + map_id)


def log_uniform_counts(rng: np.random.Generator, n: int, lo=BATCH_MIN_POINTS, hi=BATCH_MAX_POINTS) -> np.ndarray:
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(np.int64)
