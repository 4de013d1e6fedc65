"""The benchmark frame at its full size: 1920x1080 pixels, 1,000,000 AVPLs,
10-degree cone antiradiance (Shaders/Gather.frag:149-188) over the clustered
lists on the gather's second stream, and the light-cut radiance chosen by
the analytic (Lightcuts) error bound at 1 % with no image floor, kept nodes
summed with second-order moments and nodes of at most 16 AVPLs exactly: the
workload as written.  Two more settings of the same frame reach the 1e-3
target: bound 0.0045 with an image floor of 4, and the workload's own
bound and floor, each with nodes of at most 128 AVPLs summed exactly.

Size-independent properties stand in for an oracle that cannot run this
frame in seconds:
  * the antiradiance is exact: the FP32 clustered sum equals the FP64 exact
    gather on a pixel sample (relative L2 <= 1e-5);
  * the gather is linear in the AVPL intensities: doubling every AVPL's
    radiance and antiradiance doubles the result (a power of two is exact in
    FP32, and the cut's thresholds double with the estimates);
  * the schedule settings (kept nodes on the matrix cores, traversal stacks
    of 64 and 128, antiradiance blocks of 768 and 3072) give the same image
    to rounding;
  * the result's error against the FP64 exact gather stays within the level
    measured for each setting on this test's sample (every 8th pixel in x
    and y): <= 2e-3 for the workload as written (1.77e-3 measured), <= 1e-3
    for the other two (9.67e-4 and 8.56e-4 measured).
"""
import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import DeviceAvpls, Gather, default_params, device_array

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def frame_params(**kw):
    base = dict(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=1,
                cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS, cutCriterion=ffi.ARC_CRITERION_BOUND, errorBound=0.01,
                cutImageFloor=0.0, cutExactCount=16, antiradianceModel=ffi.ARC_ANTIRAD_CONE,
                antiradianceMode=ffi.ARC_ANTI_CLUSTERED)
    base.update(kw)
    return default_params(**base)


def test_full_frame(gpu_ctx):
    print("[frame] building the 1080p x 1M scene", flush=True)
    sc = S.clutter_1080p()
    assert (sc.width, sc.height, sc.num_avpls) == (1920, 1080, 1_000_000)
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    av = DeviceAvpls.from_scene(sc)
    p = frame_params()
    st = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    res = g.result[:, :3].double().cpu().numpy()
    anti = g.antiradiance[:, :3].double().cpu().numpy()
    assert np.isfinite(res).all() and st.errorPixels == 0

    # FP64 exact gather on every 8th pixel in x and y
    k = 8
    ys, xs = np.arange(0, sc.height, k), np.arange(0, sc.width, k)
    pix = (ys[:, None] * sc.width + xs[None, :]).ravel()
    sub = S.Scene(**{**sc.__dict__, "width": len(xs), "height": len(ys), "pos4": sc.pos4[pix].copy(),
                     "nrm4": sc.nrm4[pix].copy(), "albedo4": sc.albedo4[pix].copy()})
    gs = Gather(gpu_ctx, sub.width, sub.height, sub.materials)
    gs.set_gbuffer(device_array(sub.pos4), device_array(sub.nrm4), device_array(sub.albedo4))
    print(f"[frame] FP64 exact gather on {len(pix)} pixels", flush=True)
    r64, a64 = gs.exact_f64(av, sc.cam, sc.extent, p)
    gpu_ctx.sync()
    r64, a64 = r64.cpu().numpy(), a64.cpu().numpy()
    gs.close()
    e_anti = rel_l2(anti[pix], a64)
    e_res = rel_l2(res[pix], r64 - a64)
    print(f"[frame] cone antiradiance rel-L2 {e_anti:.2e}, result rel-L2 {e_res:.3e}", flush=True)
    assert e_anti <= 1e-5
    assert e_res <= 2e-3

    # the schedule settings at full size: the kept nodes on the matrix cores,
    # the deepest and shallowest traversal stacks and the other antiradiance
    # blocks give the same image (to rounding) with no error pixels
    rad_ref = g.radiance[:, :3].double().cpu().numpy()
    for stack, chunk, pool in ((96, 1536, 32), (128, 768, 0), (64, 3072, 0)):
        g.set_launch_shape(stack, chunk)
        g.set_estimate_pool(pool)
        st_s = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_COUNTERS)
        gpu_ctx.sync()
        e_rad = rel_l2(g.radiance[:, :3].double().cpu().numpy(), rad_ref)
        e_all = rel_l2(g.result[:, :3].double().cpu().numpy(), res)
        print(f"[frame] schedule stack {stack}, block {chunk}, pool {pool}: radiance rel-L2 {e_rad:.2e}, "
              f"result {e_all:.2e}, error pixels {st_s.errorPixels}", flush=True)
        if st_s.errorPixels == 0:
            assert e_rad <= 1e-5 and e_all <= 1e-5
        else:  # a stack too shallow for some tile: the tuner's reason to reject it
            assert stack == 64
    g.set_launch_shape(96, 1536)
    g.set_estimate_pool(0)

    # two settings within the 1e-3 target: the cheapest measured, and the
    # workload's own bound and floor with nodes of at most 128 AVPLs exact
    for bound, floor in ((0.0045, 4.0), (0.01, 0.0)):
        p2 = frame_params(errorBound=bound, cutImageFloor=floor, cutExactCount=128, cutBudget=1 << 20)
        st2 = g.run(av, sc.cam, sc.extent, p2, ffi.ARC_RUN_COUNTERS)
        gpu_ctx.sync()
        res_acc = g.result[:, :3].double().cpu().numpy()
        e_acc = rel_l2(res_acc[pix], r64 - a64)
        print(f"[frame] bound {bound:g}, floor {floor:g}, exact <= 128: result rel-L2 {e_acc:.3e}", flush=True)
        assert st2.errorPixels == 0 and st2.budgetHits == 0
        assert e_acc <= 1e-3

    # linearity: doubled intensities, doubled image
    for key in ("rad", "anti"):
        av.tensors[key].mul_(2.0)
    g.run(av, sc.cam, sc.extent, p)
    gpu_ctx.sync()
    res2 = g.result[:, :3].double().cpu().numpy()
    e_lin = rel_l2(res2, 2.0 * res)
    print(f"[frame] doubled AVPLs: rel-L2 vs twice the image {e_lin:.2e}", flush=True)
    assert e_lin <= 1e-5
    g.close()
