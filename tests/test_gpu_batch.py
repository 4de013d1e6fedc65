"""Batched views in one HIP graph (arc_batch_*), against the same views run
one by one.

Each view is a complete frame (LBVH rebuild, cut, antiradiance, combine) of
its own G-buffer, AVPL set and parameters.  A replay of the graph must give
every view's eager result; the clustered antiradiance sums its lists in an
order that can differ between runs, so images agree to 1e-5 relative L2
rather than bit for bit.  Between replays the caller's stream does other
work (clears, an eager frame on another gather) and the AVPL arrays change
in place: the graph reads its inputs live, so doubling every AVPL's radiance
and antiradiance doubles the image (a power-of-two scale is exact in FP32
and leaves every cut decision unchanged).
"""
import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import Batch, DeviceAvpls, Gather, default_params, device_array

pytestmark = pytest.mark.gpu

VIEWS = [
    # (width, height, num_avpls, seed, params)
    (64, 48, 2048, 11, dict(useLightCuts=0)),
    (96, 80, 4096, 12, dict(useLightCuts=1, cutMode=ffi.ARC_CUT_SOLID_ANGLE,
                            antiradianceMode=ffi.ARC_ANTI_CLUSTERED)),
    (80, 64, 3000, 13, dict(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=1)),
    (72, 56, 5000, 16, dict(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=1,
                            cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS, cutCriterion=ffi.ARC_CRITERION_BOUND,
                            cutImageFloor=0.0,
                            antiradianceModel=ffi.ARC_ANTIRAD_CONE, antiradianceMode=ffi.ARC_ANTI_CLUSTERED)),
    (128, 96, 8192, 14, dict(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=0,
                             antiradianceMode=ffi.ARC_ANTI_CLUSTERED)),
    (48, 40, 1024, 15, dict(useLightCuts=0, antiradianceMode=ffi.ARC_ANTI_CLUSTERED)),
]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def rig(gpu_ctx):
    import torch

    b = Batch(gpu_ctx, lanes=3)
    views, ref = [], []
    for k, (w, h, n, seed, kw) in enumerate(VIEWS):
        sc = S.cornell_256(seed=seed, width=w, height=h, num_avpls=n)
        lane = b.lane(k)
        g = Gather(lane, w, h, sc.materials)
        g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
        av = DeviceAvpls.from_scene(sc)
        p = default_params(**kw)
        torch.cuda.synchronize()
        g.run(av, sc.cam, sc.extent, p)
        lane.sync()
        ref.append(g.result[:, :3].double().cpu().numpy())
        views.append((g, av, sc.cam, sc.extent, p))
    yield b, views, ref
    gpu_ctx.sync()
    for g, *_ in views:
        g.close()
    b.close()


def test_launch_before_capture_is_an_error(gpu_ctx):
    b = Batch(gpu_ctx, lanes=1)
    with pytest.raises(ffi.ArcError) as e:
        b.launch()
    assert e.value.status == ffi.ARC_E_STATE
    b.close()


def test_gather_off_lane_is_rejected(gpu_ctx):
    sc = S.cornell_256(seed=1, width=32, height=32, num_avpls=256)
    g = Gather(gpu_ctx, 32, 32, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    b = Batch(gpu_ctx, lanes=2)
    with pytest.raises(ffi.ArcError) as e:
        b.capture([(g, av, sc.cam, sc.extent, default_params())])
    assert e.value.status == ffi.ARC_E_INVALID
    b.close()
    g.close()


def test_shared_gather_is_rejected(gpu_ctx):
    """Two views of one gather would overwrite each other's outputs inside
    the graph: arc_batch_capture rejects them."""
    sc = S.cornell_256(seed=2, width=32, height=32, num_avpls=256)
    b = Batch(gpu_ctx, lanes=2)
    g = Gather(b.lane(0), 32, 32, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    view = (g, av, sc.cam, sc.extent, default_params())
    with pytest.raises(ffi.ArcError) as e:
        b.capture([view, view])
    assert e.value.status == ffi.ARC_E_INVALID
    g.close()
    b.close()


def test_one_cut_launch_matches_eager(gpu_ctx):
    """Views whose radiance is the moment cut: the graph runs their cuts as
    one launch over all their tiles (heaviest first from the second replay
    on), and each replay gives every view's eager result."""
    shapes = [(72, 56, 5000, 21), (200, 120, 60_000, 22), (96, 64, 20_000, 23), (40, 40, 3000, 24),
              (160, 96, 30_000, 25)]
    p = default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=1,
                       cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS, cutCriterion=ffi.ARC_CRITERION_BOUND,
                       cutImageFloor=0.0, antiradianceModel=ffi.ARC_ANTIRAD_CONE,
                       antiradianceMode=ffi.ARC_ANTI_CLUSTERED)
    b = Batch(gpu_ctx, lanes=2)
    views, ref = [], []
    try:
        for k, (w, h, n, seed) in enumerate(shapes):
            sc = S.clutter_1080p(seed=seed, width=w, height=h, num_avpls=n, num_tris=20_000)
            lane = b.lane(k)
            g = Gather(lane, w, h, sc.materials)
            g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
            av = DeviceAvpls.from_scene(sc)
            g.run(av, sc.cam, sc.extent, p)
            lane.sync()  # the frame ran on the lane's stream
            ref.append(g.result[:, :3].double().cpu().numpy())
            views.append((g, av, sc.cam, sc.extent, p))
        print(f"[batch] one cut launch over {len(views)} views", flush=True)
        b.set_one_cut_launch(True)
        b.capture(views)
        for rep in range(3):
            for g, *_ in views:
                g.result.fill_(-1.0)
            b.launch()
            gpu_ctx.sync()
            for k, (g, *_rest) in enumerate(views):
                e = rel_l2(g.result[:, :3].double().cpu().numpy(), ref[k])
                print(f"[batch] one cut launch, replay {rep} view {k}: rel-L2 {e:.2e}", flush=True)
                assert e <= 1e-5
    finally:
        for g, *_ in views:
            g.close()
        b.close()


def test_pooled_estimates_replay_matches_eager(gpu_ctx):
    """Moment-cut views whose kept nodes go to the matrix-core pass (pool of
    32 chunks per tile; one view with a pool of one chunk per tile, which
    fills and leaves the rest to the traversal) and other launch shapes,
    captured and replayed: the pool is cleared by a kernel inside the graph,
    so every replay gives each view's eager image."""
    shapes = [(72, 56, 5000, 31, 32, (96, 1536)), (203, 117, 40_000, 32, 1, (128, 768)),
              (96, 64, 20_000, 33, 32, (64, 3072))]
    p = default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=1,
                       cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS, cutCriterion=ffi.ARC_CRITERION_BOUND,
                       cutImageFloor=0.0, antiradianceModel=ffi.ARC_ANTIRAD_CONE,
                       antiradianceMode=ffi.ARC_ANTI_CLUSTERED)
    b = Batch(gpu_ctx, lanes=2)
    views, ref = [], []
    try:
        for k, (w, h, n, seed, pool, shape) in enumerate(shapes):
            sc = S.clutter_1080p(seed=seed, width=w, height=h, num_avpls=n, num_tris=20_000)
            lane = b.lane(k)
            g = Gather(lane, w, h, sc.materials)
            g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
            g.set_estimate_pool(pool)
            g.set_launch_shape(*shape)
            av = DeviceAvpls.from_scene(sc)
            g.run(av, sc.cam, sc.extent, p)
            lane.sync()
            ref.append(g.result[:, :3].double().cpu().numpy())
            views.append((g, av, sc.cam, sc.extent, p))
        b.capture(views)
        for rep in range(3):
            for g, *_ in views:
                g.result.fill_(-1.0)
            b.launch()
            gpu_ctx.sync()
            for k, (g, *_rest) in enumerate(views):
                e = rel_l2(g.result[:, :3].double().cpu().numpy(), ref[k])
                print(f"[batch] pooled estimates, replay {rep} view {k}: rel-L2 {e:.2e}", flush=True)
                assert e <= 1e-5
    finally:
        for g, *_ in views:
            g.close()
        b.close()


def test_replays_match_eager(gpu_ctx, rig):
    import torch

    b, views, ref = rig
    print(f"[batch] capture {len(views)} views on {len(b.lanes)} lanes", flush=True)
    b.capture(views)
    side = S.cornell_256(seed=99, width=40, height=40, num_avpls=512)
    gs = Gather(gpu_ctx, 40, 40, side.materials)
    gs.set_gbuffer(device_array(side.pos4), device_array(side.nrm4), device_array(side.albedo4))
    avs = DeviceAvpls.from_scene(side)
    gs.run(avs, side.cam, side.extent, default_params())
    gpu_ctx.sync()
    side_ref = gs.result.clone()
    for rep in range(3):
        for g, *_ in views:
            g.result.fill_(-1.0)   # on the caller's stream, before the launch
        b.launch()
        gs.run(avs, side.cam, side.extent, default_params())  # eager work between replays
        gpu_ctx.sync()
        for k, (g, *_rest) in enumerate(views):
            e = rel_l2(g.result[:, :3].double().cpu().numpy(), ref[k])
            print(f"[batch] replay {rep} view {k}: rel-L2 {e:.2e}", flush=True)
            assert e <= 1e-5
        assert torch.equal(gs.result, side_ref)


def test_replay_reads_inputs_live(gpu_ctx, rig):
    b, views, ref = rig
    b.capture(views)
    for _, av, *_rest in views:
        av.tensors["rad"].mul_(2.0)
        av.tensors["anti"].mul_(2.0)
    try:
        b.launch()
        gpu_ctx.sync()
        for k, (g, *_rest) in enumerate(views):
            e = rel_l2(g.result[:, :3].double().cpu().numpy(), 2.0 * ref[k])
            print(f"[batch] doubled AVPLs view {k}: rel-L2 {e:.2e}", flush=True)
            assert e <= 1e-5
    finally:
        for _, av, *_rest in views:
            av.tensors["rad"].mul_(0.5)
            av.tensors["anti"].mul_(0.5)
        gpu_ctx.sync()
