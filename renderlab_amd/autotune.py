"""Autotuner for the traversal launches.

Sweeps the launch knobs the kernels expose -- workgroup size (the tile of
rays a workgroup holds), LDS stack entries per lane, the refill threshold
(how many idle lanes a wave waits for before fetching new rays) and the
workgroups per CU (waves per CU), and for CW8 whether lanes with triangles
pending join the wave's node steps -- over a sample of the caller's own rays,
timing each candidate with HIP events, and installs the fastest with
rq_set_tuning / rq_set_batch_tuning. Results are cached in a JSON file keyed
by (layout, mode, scene-size class, ray kind), so a service pays for the
sweep once per workload class.
"""
import json
import os
import time

import numpy as np

from . import rq

CANDIDATES = [(128, 8), (128, 16), (128, 32), (256, 8), (256, 16), (256, 32), (64, 16), (64, 32)]
REFILLS = [1, 8, 16, 32, 64]
GRIDS = [0, 2, 4]


def size_class(triangles):
    return int(np.clip(np.floor(np.log10(max(triangles, 1)) * 2) / 2, 0, 9) * 10)


def _time(fn, reps=2, stream=None):
    """Mean time of fn() in ms; fn must launch on `stream` (None: the default stream), where the events go."""
    start, end = rq.Event(), rq.Event()
    fn()  # warm
    rq.synchronize()
    start.record(stream)
    for _ in range(reps):
        fn()
    end.record(stream)
    end.synchronize()
    return start.elapsed_ms(end) / reps


def _speculation(layout, best, set_fn, time_fn):
    """CW8 only (swept with speculation on): time the best launch with it off and on; installs the faster."""
    ms, block, stack, refill, grid = best
    if layout != "cw8":
        set_fn(block, stack, refill, grid, 0)
        return ms, 0
    timed = []
    for spec in (0, 1):
        set_fn(block, stack, refill, grid, spec)
        timed.append((time_fn(), spec))
    ms, spec = min(timed)
    set_fn(block, stack, refill, grid, spec)
    return ms, spec


def tune_trace(scene, layout, rays_dev, count, mode, hits=None, vis=None, quick=False):
    """Sweep the single-scene launch for (layout, mode) on these rays; installs and returns the best."""
    results = []
    grids = [0] if quick else GRIDS
    refills = [1, 16, 64] if quick else REFILLS
    for block, stack in CANDIDATES:
        for refill in refills:
            for grid in grids:
                rq.set_tuning(layout, mode, block, stack, refill, grid, speculate=int(layout == "cw8"))
                ms = _time(lambda: scene.trace(layout, rays_dev, count, mode, hits=hits, vis=vis))
                results.append((ms, block, stack, refill, grid))
    results.sort()
    _, block, stack, refill, grid = results[0]
    ms, spec = _speculation(layout, results[0],
                            lambda b, s, r, g, x: rq.set_tuning(layout, mode, b, s, r, g, speculate=x),
                            lambda: _time(lambda: scene.trace(layout, rays_dev, count, mode, hits=hits, vis=vis)))
    return dict(ms=ms, block=block, stack=stack, refill=refill, grid_per_cu=grid, speculate=spec,
                candidates=len(results))


def tune_ranges(scene, layout, ranges, quick=False, stream=None):
    """Sweep the one-scene, several-range launch (it runs with the closest-hit tuning of `layout`) on
    these ranges; installs and returns the best."""
    results = []
    grids = [0] if quick else GRIDS
    refills = [1, 16, 64] if quick else REFILLS
    for block, stack in CANDIDATES:
        for refill in refills:
            for grid in grids:
                rq.set_tuning(layout, 0, block, stack, refill, grid, speculate=int(layout == "cw8"))
                ms = _time(lambda: scene.trace_ranges(layout, ranges, stream=stream), stream=stream)
                results.append((ms, block, stack, refill, grid))
    results.sort()
    _, block, stack, refill, grid = results[0]
    ms, spec = _speculation(layout, results[0],
                            lambda b, s, r, g, x: rq.set_tuning(layout, 0, b, s, r, g, speculate=x),
                            lambda: _time(lambda: scene.trace_ranges(layout, ranges, stream=stream), stream=stream))
    return dict(ms=ms, block=block, stack=stack, refill=refill, grid_per_cu=grid, speculate=spec,
                candidates=len(results))


def tune_batches(scenes, layout, batches, quick=False):
    """Sweep the multi-scene launch on these batches; installs and returns the best."""
    results = []
    grids = [0] if quick else GRIDS
    refills = [1, 16, 64] if quick else REFILLS
    for block, stack in CANDIDATES:
        for refill in refills:
            for grid in grids:
                rq.set_batch_tuning(layout, block, stack, refill, grid, speculate=int(layout == "cw8"))
                ms = _time(lambda: rq.trace_batches(scenes, layout, batches))
                results.append((ms, block, stack, refill, grid))
    results.sort()
    _, block, stack, refill, grid = results[0]
    ms, spec = _speculation(layout, results[0],
                            lambda b, s, r, g, x: rq.set_batch_tuning(layout, b, s, r, g, speculate=x),
                            lambda: _time(lambda: rq.trace_batches(scenes, layout, batches)))
    return dict(ms=ms, block=block, stack=stack, refill=refill, grid_per_cu=grid, speculate=spec,
                candidates=len(results))


class Cache:
    """JSON cache of tuned launches keyed by workload class."""

    def __init__(self, path):
        self.path = path
        self.data = {}
        if os.path.exists(path):
            with open(path) as f:
                self.data = json.load(f)

    @staticmethod
    def key(layout, mode, triangles, kind):
        return f"{layout}/{mode}/{size_class(triangles)}/{kind}"

    def apply(self, layout, mode, triangles, kind):
        e = self.data.get(self.key(layout, mode, triangles, kind))
        if e:
            rq.set_tuning(layout, mode, e["block"], e["stack"], e["refill"], e["grid_per_cu"],
                          speculate=e.get("speculate", 0))
        return e

    def store(self, layout, mode, triangles, kind, entry):
        e = {k: entry[k] for k in ("block", "stack", "refill", "grid_per_cu", "ms")}
        e["speculate"] = entry.get("speculate", 0)
        e["when"] = time.time()
        self.data[self.key(layout, mode, triangles, kind)] = e
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)
