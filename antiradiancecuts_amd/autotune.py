"""Autotuner of the gather's schedule knobs, per view and per batch.

A view's frame time depends on knobs that leave its result (nearly) alone:
the error-bound cut's traversal (one cut per 8x8 tile, or per pixel), the
depth of the estimate cut that sets each pixel's threshold, and the
traversal stack of the solid-angle cut.  The tuner times every candidate on
the view itself (HIP events on the library's stream, median of `reps`
frames after one warm-up) and keeps the fastest.  The accuracy knobs
(errorBound, clusterRefinementThreshold) are the caller's and are not
touched.
"""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np

from . import ffi
from .runtime import default_params


def candidates(base: ffi.arc_gather_params):
    """The schedule variants of `base` the tuner tries."""
    out = []
    if base.useLightCuts and base.cutMode == ffi.ARC_CUT_ERROR_BOUND:
        # the moment estimator runs per 8x8 tile only
        tiles = (1,) if base.cutEstimator == ffi.ARC_ESTIMATOR_MOMENTS else (1, 0)
        for per_tile, d0 in itertools.product(tiles, (4, 6, 8)):
            p = _copy(base)
            p.cutPerTile, p.cutEstimateDepth = per_tile, d0
            out.append(p)
    else:
        out.append(_copy(base))
    return out


def _copy(p: ffi.arc_gather_params) -> ffi.arc_gather_params:
    q = ffi.arc_gather_params()
    C.memmove(C.addressof(q), C.addressof(p), C.sizeof(ffi.arc_gather_params))
    return q


def tune(gather, avpls, cam, extent, base: ffi.arc_gather_params | None = None, reps: int = 3):
    """Returns (best params, [(cutPerTile, cutEstimateDepth, median ms), ...])."""
    base = base or default_params()
    flags = ffi.ARC_RUN_PROFILE | ffi.ARC_TRUST_INPUT
    table = []
    best, best_ms = None, np.inf
    for p in candidates(base):
        gather.run(avpls, cam, extent, p, ffi.ARC_TRUST_INPUT)
        ms = float(np.median([gather.run(avpls, cam, extent, p, flags).msTotal for _ in range(reps)]))
        table.append((p.cutPerTile, p.cutEstimateDepth, ms))
        if ms < best_ms:
            best, best_ms = p, ms
    return best, table


def tune_batch(batch, views, reps: int = 3):
    """The batch's schedule: the views' moment cuts as one launch over all
    their tiles, or one launch per view.  Captures the views both ways, times
    `reps` graph launches of each (HIP events on the batch's stream, median
    after one warm-up launch) and leaves the batch captured with the faster.
    Returns (one_launch, {mode: median ms})."""
    import torch

    table = {}
    for one in (True, False):
        batch.set_one_cut_launch(one)
        batch.capture(views)
        batch.launch()
        times = []
        for _ in range(reps):
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            batch.launch()
            stop.record()
            torch.cuda.synchronize()
            times.append(start.elapsed_time(stop))
        table["one cut launch" if one else "a cut per view"] = float(np.median(times))
    best = table["one cut launch"] <= table["a cut per view"]
    if not best:
        return False, table
    batch.set_one_cut_launch(True)
    batch.capture(views)
    return True, table
