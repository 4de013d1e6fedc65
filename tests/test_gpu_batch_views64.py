"""The batched form at its full shapes: 64 views, seeds 100..163, each a
seeded variant of the 1080p workload's scene (Cornell walls + 200K random
triangles) with W x H drawn uniformly from [256, 1920] x [256, 1080] and an
AVPL count log-uniform over [16K, 1M], all at the workload's setting
(Lightcuts bound 0.01 with moment estimates, nodes of at most 16 AVPLs
exact, 10-degree cone antiradiance over clustered lists), dealt over four
lanes and captured into one graph.

Every replay must give each view's eager result, both when the views' cuts
run as one launch over all their tiles and when each view's cut is a launch
of its own (the default): the cut's tiles sum the same terms in the same
order wherever they run; the clustered antiradiance sums its lists in an
order the appends leave open, so images agree to 1e-5 relative L2 rather
than bit for bit.  The views' frame times alone and each graph launch's time
are printed (p50 / p99 over the views).
"""
import math
import time

import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import Batch, DeviceAvpls, Gather, default_params, device_array

pytestmark = pytest.mark.gpu

NVIEWS = 64
LANES = 4


def view_shapes(nviews=NVIEWS, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nviews):
        w = int(rng.integers(256, 1921)) // 8 * 8
        h = int(rng.integers(256, 1081)) // 8 * 8
        n = int(round(math.exp(rng.uniform(math.log(16384), math.log(1_000_000)))))
        out.append((100 + k, w, h, n))
    return out


def workload_params():
    return default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=1,
                          cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS, cutCriterion=ffi.ARC_CRITERION_BOUND,
                          errorBound=0.01, cutImageFloor=0.0, cutExactCount=16,
                          antiradianceModel=ffi.ARC_ANTIRAD_CONE, antiradianceMode=ffi.ARC_ANTI_CLUSTERED)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_64_views_replay_matches_eager(gpu_ctx):
    import torch

    shapes = view_shapes()
    sizes = np.array([w * h for _, w, h, _ in shapes])
    counts = np.array([n for *_, n in shapes])
    assert sizes.min() >= 256 * 256 and sizes.max() <= 1920 * 1080
    assert counts.min() >= 16384 and counts.max() <= 1_000_000
    b = Batch(gpu_ctx, lanes=LANES)
    p = workload_params()
    views, ref, alone = [], [], []
    try:
        t0 = time.time()
        for k, (seed, w, h, n) in enumerate(shapes):
            sc = S.clutter_1080p(seed=seed, width=w, height=h, num_avpls=n)
            g = Gather(b.lane(k), w, h, sc.materials)
            g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
            av = DeviceAvpls.from_scene(sc)
            g.run(av, sc.cam, sc.extent, p, ffi.ARC_TRUST_INPUT)
            st = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_PROFILE | ffi.ARC_TRUST_INPUT)
            alone.append(st.msTotal)
            torch.cuda.synchronize()
            ref.append(g.result[:, :3].double().cpu().numpy())
            assert np.isfinite(ref[-1]).all() and st.errorPixels == 0
            views.append((g, av, sc.cam, sc.extent, p))
            print(f"[views64] view {k}: {w}x{h}, {n} AVPLs, alone {alone[-1]:.3f} ms", flush=True)
        print(f"[views64] setup {time.time() - t0:.1f} s", flush=True)
        for one in (True, False):
            b.set_one_cut_launch(one)
            b.capture(views)
            replay_and_check(gpu_ctx, b, views, ref, sizes, counts, "one cut launch" if one else "a cut per view")
        a = np.array(alone)
        print(f"[views64] views alone: p50 {np.percentile(a, 50):.2f} ms, p99 {np.percentile(a, 99):.2f} ms, "
              f"sum {a.sum():.1f} ms", flush=True)
    finally:
        for g, *_ in views:
            g.close()
        b.close()


def replay_and_check(gpu_ctx, b, views, ref, sizes, counts, what):
    import torch

    for rep in range(2):
        for g, *_ in views:
            g.result.fill_(-1.0)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        b.launch()
        stop.record()
        gpu_ctx.sync()
        worst = 0.0
        for k, (g, *_rest) in enumerate(views):
            e = rel_l2(g.result[:, :3].double().cpu().numpy(), ref[k])
            worst = max(worst, e)
            assert e <= 1e-5, (k, e)
        pairs = float((sizes * counts).sum())
        ms = start.elapsed_time(stop)
        print(f"[views64] {what}, replay {rep}: {ms:.1f} ms for {NVIEWS} views "
              f"({pairs / (ms * 1e-3):.3e} pairs/s), worst rel-L2 {worst:.2e}", flush=True)
