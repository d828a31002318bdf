"""Per-size-bucket tuning of a simulation's launch parameters by measurement.

For each bucket of simulations (grouped by particle count), one representative is stepped with
every candidate setting (after a few warm-up steps, timed over a fixed number of steps on the GPU)
and the fastest setting is used for the whole bucket.  Settings are plain lsdem_config entries:
results do not depend on them (see include/lsdem.h), only the time per step does.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

DEFAULT_SPACE = {
    "block_size": (64, 128, 256),      # threads per block of the body kernels
    "narrow_batch": (0, 1, 4),         # candidate pairs per wavefront batch (0: chosen per step)
    "skin_factor": (0.04, 0.08, 0.16), # candidate-pair skin: fraction of the largest node radius x scale
}


def _settings(space: dict):
    keys = sorted(space)
    for values in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, values))


def _config(scene, base: dict, setting: dict) -> dict:
    cfg = {**scene.cfg, **base}
    for k, v in setting.items():
        if k == "skin_factor":
            r = max(float(np.max(np.linalg.norm(s["nodes"], axis=1))) for s in scene.shapes if len(s["nodes"]))
            cfg["skin"] = v * r * float(np.max(scene.particles["scale"]))
        else:
            cfg[k] = v
    return cfg


def tune(L, scene, base: dict | None = None, space: dict | None = None, warmup: int = 20, steps: int = 200):
    """Time every setting of `space` on `scene`; returns (best setting, {setting tuple: ms/step})."""
    base = base or {}
    space = space or DEFAULT_SPACE
    times = {}
    for setting in _settings(space):
        sim = scene.load(L.Sim(_config(scene, base, setting)))
        sim.step(warmup)
        t = time.perf_counter()
        sim.step(steps)
        times[tuple(sorted(setting.items()))] = 1e3 * (time.perf_counter() - t) / steps
        del sim
    best = min(times, key=times.get)
    return dict(best), times


def tune_buckets(L, scenes: list, edges=(300, 600), base: dict | None = None, space: dict | None = None,
                 warmup: int = 20, steps: int = 200):
    """Buckets by particle count (edges split them); the largest simulation of each bucket is tuned.
    Returns the setting per simulation and the per-bucket timings."""
    def bucket(n):
        return sum(n >= e for e in edges)
    groups: dict = {}
    for i, sc in enumerate(scenes):
        groups.setdefault(bucket(sc.n_particles), []).append(i)
    chosen, report = [None] * len(scenes), {}
    for b, idx in sorted(groups.items()):
        rep = max(idx, key=lambda i: scenes[i].n_particles)
        best, times = tune(L, scenes[rep], base, space, warmup, steps)
        report[b] = {"representative": rep, "particles": scenes[rep].n_particles, "best": best,
                     "ms_per_step": {str(dict(k)): v for k, v in times.items()}}
        for i in idx:
            chosen[i] = best
    return chosen, report


def apply(scene, base: dict, setting: dict) -> dict:
    """The lsdem_config entries of `scene` with `setting` applied."""
    return _config(scene, base, setting)
