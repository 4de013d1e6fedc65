"""Autotuner of the gather's schedule, per view and per batch.

A view's frame time depends on knobs that leave its image alone (to
rounding):
  * launch shapes (Gather.set_launch_shape): the moment cut's traversal
    stack, 64, 96 or 128 nodes per 8x8 tile (the deepest cut a tile can
    refine to), and the clustered antiradiance's list entries filed into LDS
    per pass, 768, 1536 or 3072 (the AVPL block of a pass);
  * where the moment cut's kept nodes are estimated: in the traversal, or
    handed to the matrix-core pass (Gather.set_estimate_pool, 0 or 32
    chunks per tile);
  * for the centroid estimator, the cut per 8x8 tile or per pixel and the
    depth of the estimate cut that sets each pixel's threshold (these move
    the cut a little).
The tuner times the candidates on the view itself (HIP events on the
library's stream, median of `reps` frames after one warm-up), one knob at a
time with the others at their best so far (coordinate descent), and rejects
a candidate whose frame reports errorPixels (a stack too shallow for the
cut).  The accuracy knobs (errorBound, cutImageFloor, cutExactCount,
clusterRefinementThreshold) are the caller's and are not touched.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import itertools

import numpy as np

from . import ffi
from .runtime import default_params

STACKS = (96, 64, 128)
CHUNKS = (1536, 768, 3072)
POOLS = (0, 32)


@dataclasses.dataclass
class Schedule:
    """A gather's launch shapes and estimate pool."""
    cut_stack: int = 96
    anti_chunk: int = 1536
    estimate_pool: int = 0

    def apply(self, gather) -> None:
        gather.set_launch_shape(self.cut_stack, self.anti_chunk)
        gather.set_estimate_pool(self.estimate_pool)


def candidates(base: ffi.arc_gather_params):
    """The parameter variants of `base` the tuner tries (the centroid
    estimator's traversal and estimate depth; one for other settings)."""
    out = []
    if (base.useLightCuts and base.cutMode == ffi.ARC_CUT_ERROR_BOUND
            and base.cutEstimator != ffi.ARC_ESTIMATOR_MOMENTS):
        for per_tile, d0 in itertools.product((1, 0), (4, 6, 8)):
            p = _copy(base)
            p.cutPerTile, p.cutEstimateDepth = per_tile, d0
            out.append(p)
    else:
        out.append(_copy(base))
    return out


def schedule_axes(base: ffi.arc_gather_params):
    """The launch-shape knobs that apply to `base`: (name, values), the
    current default first."""
    moments = (base.useLightCuts and base.cutMode == ffi.ARC_CUT_ERROR_BOUND
               and base.cutEstimator == ffi.ARC_ESTIMATOR_MOMENTS)
    axes = []
    if moments:
        axes += [("cut_stack", STACKS), ("estimate_pool", POOLS)]
    if base.antiradianceMode == ffi.ARC_ANTI_CLUSTERED:
        axes.append(("anti_chunk", CHUNKS))
    return axes


def _copy(p: ffi.arc_gather_params) -> ffi.arc_gather_params:
    q = ffi.arc_gather_params()
    C.memmove(C.addressof(q), C.addressof(p), C.sizeof(ffi.arc_gather_params))
    return q


def _time(gather, avpls, cam, extent, p, reps):
    """Median frame ms of p with the gather's current schedule, or None when
    the frame reports error pixels."""
    st = gather.run(avpls, cam, extent, p, ffi.ARC_RUN_COUNTERS | ffi.ARC_TRUST_INPUT)
    if st.errorPixels:
        return None
    flags = ffi.ARC_RUN_PROFILE | ffi.ARC_TRUST_INPUT
    return float(np.median([gather.run(avpls, cam, extent, p, flags).msTotal for _ in range(reps)]))


def tune(gather, avpls, cam, extent, base: ffi.arc_gather_params | None = None, reps: int = 3):
    """Returns (best params, best Schedule, table of (candidate, median ms));
    the gather is left set to the best schedule."""
    base = base or default_params()
    table = []
    sched = Schedule()
    sched.apply(gather)
    best, best_ms = None, np.inf
    for p in candidates(base):
        ms = _time(gather, avpls, cam, extent, p, reps)
        table.append((f"cutPerTile={p.cutPerTile},cutEstimateDepth={p.cutEstimateDepth}", ms))
        if ms is not None and ms < best_ms:
            best, best_ms = p, ms
    if best is None:  # every candidate reported error pixels: the default schedule stands
        return _copy(base), sched, table
    for name, values in schedule_axes(best):
        current = getattr(sched, name)  # timed already: best_ms
        for v in values:
            if v == current:
                continue
            trial = dataclasses.replace(sched, **{name: v})
            trial.apply(gather)
            ms = _time(gather, avpls, cam, extent, best, reps)
            table.append((f"{name}={v}", ms))
            if ms is not None and ms < best_ms:
                sched, best_ms = trial, ms
        sched.apply(gather)
    return best, sched, table


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
