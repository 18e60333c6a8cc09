"""Autotuner of libfsx's launch shapes (no counterpart in the reference).

Tunables (fsx_tune_set keys): orientation blocks per image, refinement blocks per image, the
float matcher's targets per LDS tile and blocks of 32 queries per wave, the pyramid blur's
row strip (output rows per tile) and the guided matcher's contraction workgroups (one per
chunk of 32 queries, or a fixed number of persistent ones striding over all pairs' chunks).  Every choice gives identical results (they only change how
work is split over workgroups), so the tuner times the candidates on the caller's own
workload with HIP events and keeps the fastest.  Choices are cached per (device, key,
workload signature) in memory, and in the JSON file named by FSX_TUNE_CACHE if set.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib

ORIENT_BLOCKS, REFINE_BLOCKS, FSCREEN_TILE, FSCREEN_QB, BLUR_ROWS, MATCH_BLOCKS = 0, 1, 2, 3, 4, 5
CANDIDATES = {ORIENT_BLOCKS: (1024, 2048, 4096, 8192), REFINE_BLOCKS: (64, 128, 256, 512),
              FSCREEN_TILE: (64, 128), FSCREEN_QB: (1, 2), BLUR_ROWS: (8, 16, 24, 32),
              MATCH_BLOCKS: (0, 2048, 4096, 8192)}
_cache = {}


def get(key: int) -> int:
    v = C.c_int()
    _lib.call("fsx_tune_get", key, C.byref(v))
    return v.value


def set(key: int, value: int) -> None:   # noqa: A001  (mirrors fsx_tune_set)
    _lib.call("fsx_tune_set", key, int(value))


def _device():
    import torch
    return torch.cuda.get_device_name(0)


def _load_file():
    path = os.environ.get("FSX_TUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as f:
            _cache.update({tuple(json.loads(k)): v for k, v in json.load(f).items()})


def _save_file():
    path = os.environ.get("FSX_TUNE_CACHE")
    if path:
        with open(path, "w") as f:
            json.dump({json.dumps(list(k)): v for k, v in _cache.items()}, f, indent=1)


def _time(run, stage, reps):
    """Median milliseconds of `stage` (an index pair into the run's events) over reps runs."""
    import torch
    ts = []
    for _ in range(reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(stage[2])]
        run(ev)
        ev[-1].synchronize()
        ts.append(ev[stage[0]].elapsed_time(ev[stage[1]]))
    return float(np.median(ts))


def tune(key, run, stage, signature, reps=3):
    """Try every candidate of `key` on `run(events)` and keep the fastest for `stage`
    ((first event, last event, number of events)).  Returns {value: ms}."""
    _load_file()
    ck = (_device(), int(key), str(signature))
    if ck in _cache:
        set(key, _cache[ck])
        return {}
    times = {}
    run(None)                      # warm-up
    for v in CANDIDATES[key]:
        set(key, v)
        times[v] = _time(run, stage, reps)
    best = min(times, key=times.get)
    set(key, best)
    _cache[ck] = best
    _save_file()
    return times


def tune_features(fb, imgs, reps=3):
    """Pyramid row strip, orientation and refinement grids for a FeatureBatch on its images
    (capture its graph after tuning: a captured graph keeps its launch shapes)."""
    sig = ("features", tuple(fb.shapes), fb.max_kp)
    run = lambda ev: fb.run(imgs, events=ev)     # noqa: E731
    out = {"blur_rows": tune(BLUR_ROWS, run, (0, 1, 8), sig, reps),
           "orientation_blocks": tune(ORIENT_BLOCKS, run, (3, 4, 8), sig, reps),
           "refine_blocks": tune(REFINE_BLOCKS, run, (2, 3, 8), sig, reps)}
    out["chosen"] = {"blur_rows": get(BLUR_ROWS), "orientation_blocks": get(ORIENT_BLOCKS),
                     "refine_blocks": get(REFINE_BLOCKS)}
    return out


def tune_float_matcher(fm, reps=2):
    """Blocks of queries per wave, then targets per LDS tile, of the float matcher's screening
    kernel."""
    sig = ("fmatch", fm.precision, fm.K, tuple(fm.nk[:8]), len(fm.nk))
    run = lambda ev: fm.run(events=ev)           # noqa: E731
    qb = tune(FSCREEN_QB, run, (1, 2, 5), sig, reps)
    times = tune(FSCREEN_TILE, run, (1, 2, 5), sig, reps)
    return {"query_blocks": qb, "tile": times, "chosen": get(FSCREEN_TILE),
            "chosen_query_blocks": get(FSCREEN_QB)}


def tune_matcher(pm, form="mfma", reps=3):
    """Workgroups of a PairMatcher's MFMA contraction: one per (chunk, pair), or 2048 / 4096 /
    8192 persistent ones (capture its graph after tuning: a captured graph keeps its launch
    shapes)."""
    sig = ("match", form, len(pm.pairs), tuple(pm.nk[:8]), len(pm.nk))
    run = lambda ev: pm.run(form, events=ev)     # noqa: E731
    times = tune(MATCH_BLOCKS, run, (1, 2, 4), sig, reps)
    return {"workgroups": times, "chosen": get(MATCH_BLOCKS)}
