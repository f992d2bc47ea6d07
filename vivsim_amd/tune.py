"""Tile-shape autotuner of the batched VIV step.

The bulk tiles of a batched step hold 1024 cells as (1024 / tile_y) rows of
tile_y cells along the contiguous y axis.  Which width streams best depends on
the grid (a narrow grid wastes lanes on wide tiles; short tiles re-read more
halo rows through L2), so each distinct case configuration is timed once on the
device: every candidate width runs a one-case batch for a few steps, timed with
HIP events on the current stream, and the fastest median wins.  Results are
kept per process (and optionally in the JSON file named by VSX_TUNE_CACHE).
"""

import json
import os

import torch

CANDIDATES = (64, 128, 256, 512, 1024)
_CACHE = {}


def candidates(ny):
    """Widths that do not exceed the grid's y extent rounded up to a wavefront."""
    top = max(64, ((ny + 63) // 64) * 64)
    return [t for t in CANDIDATES if t <= top]


def _key(case, storage):
    return f"{case.nx}x{case.ny}/{case.collision}/{case.forcing}/{storage}/m{case.n_markers}/w{case.ib_size}"


def _load():
    path = os.environ.get("VSX_TUNE_CACHE")
    if path and os.path.exists(path) and not _CACHE:
        with open(path) as fh:
            _CACHE.update(json.load(fh))


def _save():
    path = os.environ.get("VSX_TUNE_CACHE")
    if path:
        with open(path, "w") as fh:
            json.dump(_CACHE, fh, indent=1, sort_keys=True)


def time_tile_y(case, storage, tile_y, steps=20, reps=3):
    """Median milliseconds per step of a one-case batch with the given tile width."""
    from .sim import VIVBatch
    b = VIVBatch([case], storage=storage, tile_y=tile_y)
    f = b.initial_f()
    out = torch.empty_like(f)
    b.run(f, 2, out=out)                    # warm-up (first launches, clocks)
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.run(f, steps, out=out)
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1) / steps)
    times.sort()
    return times[len(times) // 2]


def best_tile_y(case, storage="f32", steps=20, reps=3):
    """The fastest tile width for this case configuration (timed once per process)."""
    _load()
    k = _key(case, storage)
    if k not in _CACHE:
        res = {t: time_tile_y(case, storage, t, steps, reps) for t in candidates(case.ny)}
        _CACHE[k] = {"tile_y": min(res, key=res.get), "ms_per_step": {str(t): v for t, v in res.items()}}
        _save()
    return int(_CACHE[k]["tile_y"])


def results():
    return dict(_CACHE)
