"""Error-bound light cut (ARC_CUT_ERROR_BOUND) through the C ABI.

What fixes the result without a reference implementation (the reference has
only the solid-angle cut, cudaGather.cu:405-414):
  * errorBound = 0 with no image floor refines every node that can
    contribute, down to the leaves, so the cut radiance equals the exact
    radiance (FP64 image, relative L2 <= 1e-5), per pixel and per tile;
  * a node whose box lies behind the pixel's tangent plane contributes
    exactly 0 in both, so the equality holds with such nodes kept;
  * tightening errorBound lowers the error and raises the work;
  * a spent cut budget is counted, never silent.
"""
import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import Bvh, DeviceAvpls, Gather, default_params, device_array

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    return S.cornell_256(seed=3, width=96, height=80, num_avpls=4096)


@pytest.fixture(scope="module")
def rig(gpu_ctx, scene):
    g = Gather(gpu_ctx, scene.width, scene.height, scene.materials)
    g.set_gbuffer(device_array(scene.pos4), device_array(scene.nrm4), device_array(scene.albedo4))
    av = DeviceAvpls.from_scene(scene)
    r64, _ = g.exact_f64(av, scene.cam, scene.extent, default_params())
    yield g, av, r64.cpu().numpy()
    g.close()


def eb(**kw):
    return default_params(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, **kw)


def run_rad(gpu_ctx, g, av, sc, p):
    st = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    return g.radiance[:, :3].double().cpu().numpy(), st


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("per_tile", [0, 1])
@pytest.mark.parametrize("criterion", [ffi.ARC_CRITERION_BOUND, ffi.ARC_CRITERION_ESTIMATE])
def test_zero_bound_is_exact(gpu_ctx, rig, scene, per_tile, criterion):
    g, av, r64 = rig
    print(f"[eb] zero bound tile={per_tile} criterion={criterion}", flush=True)
    rad, st = run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.0, cutImageFloor=0.0, cutPerTile=per_tile,
                                                 cutCriterion=criterion))
    e = rel_l2(rad, r64)
    print(f"[eb] rel-L2 {e:.3e}, evaluated/px {st.evaluatedNodes / scene.npix:.1f}", flush=True)
    assert st.errorPixels == 0 and st.budgetHits == 0
    assert st.evaluatedNodes <= scene.npix * (2 * scene.num_avpls - 1)
    assert e <= 1e-5


def test_tighter_bound_lowers_error(gpu_ctx, rig, scene):
    g, av, r64 = rig
    errs, work = [], []
    for eps in (0.3, 0.03, 0.003):
        rad, st = run_rad(gpu_ctx, g, av, scene, eb(errorBound=eps))
        errs.append(rel_l2(rad, r64))
        work.append(st.evaluatedNodes)
        print(f"[eb] eps {eps}: rel-L2 {errs[-1]:.3e}, evaluated/px {work[-1] / scene.npix:.1f}", flush=True)
    assert work[0] < work[1] < work[2]
    assert errs[2] < errs[0]
    assert errs[0] < 0.5


def test_image_floor_limits_work(gpu_ctx, rig, scene):
    g, av, r64 = rig
    _, st0 = run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.01, cutImageFloor=0.0))
    _, st1 = run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.01, cutImageFloor=4.0))
    assert st1.evaluatedNodes < st0.evaluatedNodes


def test_budget_hits_are_counted(gpu_ctx, rig, scene):
    g, av, r64 = rig
    rad, st = run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.0, cutImageFloor=0.0, cutBudget=8))
    assert st.budgetHits > 0
    assert np.isfinite(rad).all()
    assert st.evaluatedNodes <= scene.npix * (8 + 64)


def test_rejects_specular_materials(gpu_ctx):
    sc = S.cornell_256(seed=4, width=32, height=32, num_avpls=256)
    sc.materials[4, 6:9] = [0.4, 0.4, 0.4]
    sc.materials[4, 9] = 30.0
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4))
    with pytest.raises(ffi.ArcError) as ei:
        g.run(DeviceAvpls.from_scene(sc), sc.cam, sc.extent, eb())
    assert ei.value.status == ffi.ARC_E_UNSUPPORTED
    g.close()


@pytest.mark.parametrize("field,value", [("cutCriterion", 7), ("cutImageFloor", -1.0), ("cutBudget", 0),
                                         ("cutEvalDepth", 5), ("cutEvalDepth", -1)])
def test_invalid_parameters(gpu_ctx, rig, scene, field, value):
    g, av, _ = rig
    with pytest.raises(ffi.ArcError) as ei:
        g.run(av, scene.cam, scene.extent, eb(**{field: value}))
    assert ei.value.status == ffi.ARC_E_INVALID


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
@pytest.mark.parametrize("per_tile", [0, 1])
def test_descendant_sums(gpu_ctx, rig, scene, depth, per_tile):
    """cutEvalDepth sums a kept node as its descendants `depth` levels down.
    With no tolerance every kept node is a leaf (or contributes exactly 0),
    so the image stays exact; at a loose bound the same cut is summed finer:
    the node tests do not change, the estimates multiply (at most 2^depth
    per kept node) and the error falls."""
    g, av, r64 = rig
    rad, st = run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.0, cutImageFloor=0.0, cutPerTile=per_tile,
                                                 cutEvalDepth=depth))
    assert rel_l2(rad, r64) <= 1e-5
    rad0, st0 = run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.1, cutPerTile=per_tile))
    radd, std = run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.1, cutPerTile=per_tile, cutEvalDepth=depth))
    e0, ed = rel_l2(rad0, r64), rel_l2(radd, r64)
    print(f"[eb] depth {depth} tile={per_tile}: rel-L2 {e0:.3e} -> {ed:.3e}, "
          f"estimates/px {st0.evaluatedNodes / scene.npix:.1f} -> {std.evaluatedNodes / scene.npix:.1f}", flush=True)
    assert std.visitedNodes == st0.visitedNodes
    assert st0.evaluatedNodes < std.evaluatedNodes <= st0.evaluatedNodes * 2 ** depth
    assert ed < e0


def test_kernel_spans_of_the_two_stream_frame(gpu_ctx, rig, scene):
    """arc_gather_kernel_spans: in a cut frame with clustered antiradiance the
    antiradiance kernel runs on the second stream, forked before the LBVH
    rebuild, so it starts before the cut kernel; both spans lie inside the
    frame, and the history durations are the spans' lengths."""
    g, av, _ = rig
    p = eb(errorBound=0.1, cutPerTile=1, cutEvalDepth=3, antiradianceMode=ffi.ARC_ANTI_CLUSTERED)
    frames = 3
    for _ in range(frames):
        g.run(av, scene.cam, scene.extent, p, ffi.ARC_RUN_TIMELINE)
    gpu_ctx.sync()
    spans = np.zeros((frames, 4), np.float32)
    hist = np.zeros((frames, 6), np.float32)
    kern = np.zeros((frames, 2), np.float32)
    assert g.lib.arc_gather_kernel_spans(g.h, frames, spans.ctypes.data) == frames
    assert g.lib.arc_gather_profile_history(g.h, frames, hist.ctypes.data) == frames
    assert g.lib.arc_gather_kernel_history(g.h, frames, kern.ctypes.data) == frames
    for (r0, r1, a0, a1), h, k in zip(spans, hist, kern):
        assert 0.0 <= r0 <= r1 <= h[5] + 1e-3 and 0.0 <= a0 <= a1 <= h[5] + 1e-3
        assert a0 < r0  # forked before the rebuild
        assert abs((r1 - r0) - k[0]) < 1e-2 and abs((a1 - a0) - k[1]) < 1e-2


def test_frame_tree_is_the_avpl_bvh(gpu_ctx, rig, scene):
    """The cut's rebuild sums only what the cut reads, but its tree is the
    AvplBvh: the same codes, order, topology, leaf ranges, boxes and sorted
    leaves as arc_bvh_build_avpls over the same AVPLs, without node data."""
    g, av, _ = rig
    run_rad(gpu_ctx, g, av, scene, eb(errorBound=0.01, cutEvalDepth=1))
    frame = g.bvh_to_host()
    own = Bvh(gpu_ctx)
    own.build_avpls(av, trust=True)
    gpu_ctx.sync()
    ref = own.to_host()
    own.close()
    assert frame["n"] == ref["n"] == scene.num_avpls
    for k in ("codes", "perm", "parents", "ranges", "positions", "normals"):
        np.testing.assert_array_equal(frame[k], ref[k], err_msg=k)
    for k in ("left", "right", "bbMin", "bbMax", "visited"):
        np.testing.assert_array_equal(frame["nodes"][k], ref["nodes"][k], err_msg=k)
    assert "nd_size" in ref and "nd_size" not in frame
