"""GPU gather vs the oracle (through the C ABI).

Exact gather: FP32 (compensated) result image against the FP64 image, and the
GPU FP64 image against the CPU FP64 oracle on a pixel sample.  Tolerance:
relative L2 <= 1e-5 for FP32 vs FP64.  The disk / cone tests are
discontinuous; a pixel whose FP32 and FP64 decisions differ for some pair is
a legitimate difference, so the max-error check excludes pixels the FP64
oracle marks as within rounding of a decision threshold.

Cut gather (solid angle, kernel_bvh): every traversal decision is made with
the same FP32 arithmetic as the REF32 oracle, so visited / evaluated counts
match exactly and the per-pixel radiance agrees to FP32 summation error.
"""
import math

import numpy as np
import pytest
import torch

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import DeviceAvpls, Gather, default_params, device_array, photon_radius

pytestmark = pytest.mark.gpu

REL_L2_F32 = 1e-5


@pytest.fixture(scope="module")
def cornell():
    return S.cornell_256()


def make_gather(ctx, sc):
    g = Gather(ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    return g


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("model", [ffi.ARC_ANTIRAD_DISK, ffi.ARC_ANTIRAD_CONE])
def test_exact_gather_cornell(gpu_ctx, orc, cornell, model):
    sc = cornell
    print(f"[gather] exact cornell model={model}", flush=True)
    g = make_gather(gpu_ctx, sc)
    av = DeviceAvpls.from_scene(sc)
    p = default_params(useLightCuts=0, antiradianceModel=model, coneHalfAngle=math.pi / 18)
    g.run(av, sc.cam, sc.extent, p)
    r64, a64 = g.exact_f64(av, sc.cam, sc.extent, p)
    gpu_ctx.sync()
    res32 = g.result[:, :3].double().cpu().numpy()
    rad32 = g.radiance[:, :3].double().cpu().numpy()
    r64, a64 = r64.cpu().numpy(), a64.cpu().numpy()
    res64 = r64 - a64
    e_res, e_rad = rel_l2(res32, res64), rel_l2(rad32, r64)
    print(f"[gather] rel-L2 result {e_res:.3e} radiance {e_rad:.3e}", flush=True)
    assert e_rad <= REL_L2_F32
    assert e_res <= REL_L2_F32
    # w channels: radiance 1, antiradiance 1, result 0 (kernel_combine)
    assert torch.all(g.radiance[:, 3] == 1) and torch.all(g.result[:, 3] == 0)
    # GPU FP64 against the CPU FP64 oracle on a pixel sample
    pix = np.random.default_rng(0).choice(sc.npix, 1024, replace=False).astype(np.int32)
    ref = orc.SceneRef(sc, photon_radius(p, sc.extent), model=model, cone_half_angle=math.pi / 18)
    orad, oanti = orc.gather_exact(ref, 64, pix)
    assert rel_l2(r64[pix], orad) < 1e-11
    assert np.linalg.norm(a64[pix] - oanti) <= 1e-9 * max(1e-30, np.linalg.norm(oanti))
    g.close()


def test_exact_gather_specular_material(gpu_ctx, orc):
    """Phong lobes on both sides take the general pair path."""
    sc = S.cornell_256(seed=4, width=64, height=48, num_avpls=1500)
    sc.materials[4, 6:9] = [0.4, 0.4, 0.4]   # boxes get a lobe (cb-specular.mtl: Ks 0.5, Ns 30)
    sc.materials[4, 9] = 30.0
    sc.materials[2, 6:9] = [0.2, 0.1, 0.1]
    sc.materials[2, 9] = 8.0
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4))  # material table diffuse
    sc_nalb = S.Scene(**{**sc.__dict__, "albedo4": None})
    av = DeviceAvpls.from_scene(sc)
    p = default_params(useLightCuts=0)
    g.run(av, sc.cam, sc.extent, p)
    gpu_ctx.sync()
    ref = orc.SceneRef(sc_nalb, photon_radius(p, sc.extent))
    orad, oanti = orc.gather_exact(ref, 64)
    res = g.result[:, :3].double().cpu().numpy()
    e = rel_l2(res, orad - oanti)
    print(f"[gather] specular rel-L2 {e:.3e}", flush=True)
    assert e <= REL_L2_F32
    g.close()


def test_host_input_path(gpu_ctx, cornell):
    """AVPLs handed over as host arrays (AvplsGpu upload) give the same image."""
    sc = S.cornell_256(seed=5, width=64, height=64, num_avpls=999)
    g = make_gather(gpu_ctx, sc)
    p = default_params(useLightCuts=0)
    av = DeviceAvpls.from_scene(sc)
    g.run(av, sc.cam, sc.extent, p)
    gpu_ctx.sync()
    dev = g.result.clone()
    host = ffi.arc_avpls()
    keep = [np.ascontiguousarray(a) for a in (sc.avpl_pos, sc.avpl_nrm, sc.avpl_dir, sc.avpl_rad, sc.avpl_anti,
                                                sc.avpl_mat, sc.avpl_bounce)]
    (host.position, host.normal, host.incDirection, host.incRadiance, host.antiradiance, host.materialIndex,
     host.bounce) = [k.ctypes.data for k in keep]
    host.numAvpls = sc.num_avpls
    st = ffi.arc_gather_stats()
    import ctypes as C
    ffi.check(g.lib.arc_gather_run(g.h, C.byref(host), ffi.arc_float3(*map(float, sc.cam)), sc.extent,
                                   C.byref(p), ffi.ARC_INPUT_HOST, C.byref(st)), "run", gpu_ctx.h)
    gpu_ctx.sync()
    assert torch.equal(dev, g.result)
    g.close()


@pytest.mark.parametrize("thresh", [math.pi / 10, 0.05, 1.0])
def test_cut_solid_angle_matches_ref32(gpu_ctx, orc, cornell, thresh):
    sc = cornell
    print(f"[gather] cut solid angle thresh={thresh}", flush=True)
    g = make_gather(gpu_ctx, sc)
    av = DeviceAvpls.from_scene(sc)
    p = default_params(useLightCuts=1, clusterRefinementThreshold=thresh)
    st = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    tree = g.bvh_to_host()
    ref = orc.SceneRef(sc, photon_radius(p, sc.extent))
    orad, oerr, ovis, oev = orc.gather_cut_solid_angle(ref, tree, thresh, 32)
    print(f"[gather] visited {st.visitedNodes} / {ovis} evaluated {st.evaluatedNodes} / {oev} "
          f"errors {st.errorPixels} / {int(oerr.sum())}", flush=True)
    assert st.visitedNodes == ovis
    assert st.evaluatedNodes == oev
    assert st.errorPixels == int(oerr.sum())
    rad = g.radiance[:, :3].double().cpu().numpy()
    ok = oerr == 0
    assert rel_l2(rad[ok], orad[ok]) <= REL_L2_F32
    np.testing.assert_allclose(rad[ok], orad[ok], rtol=1e-4, atol=1e-6 * np.abs(orad).max())
    if (~ok).any():
        np.testing.assert_array_equal(rad[~ok], np.tile([1e5, 0, 1e5], ((~ok).sum(), 1)))
    g.close()


def test_cut_zero_threshold_equals_exact(gpu_ctx):
    """Pixels on a floor under every AVPL: no node is culled, so a zero
    threshold refines to every leaf and equals the exact gather."""
    from tests.test_oracle_gather import floor_scene

    sc = floor_scene(3000, 8, W=40, H=30)
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4))
    av = DeviceAvpls.from_scene(sc)
    g.run(av, sc.cam, sc.extent, default_params(useLightCuts=0))
    gpu_ctx.sync()
    exact = g.radiance[:, :3].double().cpu().numpy()
    st = g.run(av, sc.cam, sc.extent, default_params(useLightCuts=1, clusterRefinementThreshold=0.0,
                                                     stackSize=64), ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    assert st.errorPixels == 0
    assert st.evaluatedNodes == sc.npix * 3000
    e = rel_l2(g.radiance[:, :3].double().cpu().numpy(), exact)
    assert e < 1e-5
    g.close()


@pytest.mark.parametrize("model", [ffi.ARC_ANTIRAD_DISK, ffi.ARC_ANTIRAD_CONE])
@pytest.mark.parametrize("which", ["cornell", "clutter"])
def test_clustered_antiradiance_equals_exact(gpu_ctx, cornell, model, which):
    """Per-cluster AVPL lists (the reference's clustered-deferred path) sum the
    same pairs as the exact gather: only the summation order differs."""
    sc = cornell if which == "cornell" else S.clutter_1080p(num_avpls=200_000)
    print(f"[gather] clustered vs exact {which} model={model}", flush=True)
    g = make_gather(gpu_ctx, sc)
    av = DeviceAvpls.from_scene(sc)
    base = dict(useLightCuts=0, antiradianceModel=model, coneHalfAngle=math.pi / 18)
    g.run(av, sc.cam, sc.extent, default_params(antiradianceMode=ffi.ARC_ANTI_EXACT, **base))
    gpu_ctx.sync()
    exact = g.antiradiance[:, :3].double().cpu().numpy()
    st = g.run(av, sc.cam, sc.extent, default_params(antiradianceMode=ffi.ARC_ANTI_CLUSTERED, **base),
               ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    clus = g.antiradiance[:, :3].double().cpu().numpy()
    e = rel_l2(clus, exact)
    tested = st.antiEvaluatedNodes / (sc.npix * sc.num_avpls)
    print(f"[gather] clustered rel-L2 {e:.3e}, pairs tested {tested:.4%} of exact", flush=True)
    assert e < 1e-6
    np.testing.assert_allclose(clus, exact, rtol=1e-5, atol=1e-5 * np.abs(exact).max())
    g.close()


@pytest.mark.parametrize("half_angle", [math.pi / 4, 0.6 * math.pi])
def test_clustered_antiradiance_wide_cones(gpu_ctx, half_angle):
    """Cones of 45 and 108 degrees on a cluttered view (tiles that mix
    foreground triangles and walls, normals spread over whole tiles): the
    cluster and tile tests (the cone against a box's bounding sphere, and
    the tile's normal cone against the AVPL's side of every pixel's tangent
    plane) keep every pair the exact gather sums."""
    sc = S.clutter_1080p(width=480, height=272, num_avpls=60_000)
    print(f"[gather] clustered vs exact, cone half-angle {half_angle:.3f}", flush=True)
    g = make_gather(gpu_ctx, sc)
    av = DeviceAvpls.from_scene(sc)
    base = dict(useLightCuts=0, antiradianceModel=ffi.ARC_ANTIRAD_CONE, coneHalfAngle=half_angle)
    g.run(av, sc.cam, sc.extent, default_params(antiradianceMode=ffi.ARC_ANTI_EXACT, **base))
    gpu_ctx.sync()
    exact = g.antiradiance[:, :3].double().cpu().numpy()
    g.run(av, sc.cam, sc.extent, default_params(antiradianceMode=ffi.ARC_ANTI_CLUSTERED, **base))
    gpu_ctx.sync()
    clus = g.antiradiance[:, :3].double().cpu().numpy()
    assert np.abs(exact).max() > 0
    np.testing.assert_allclose(clus, exact, rtol=1e-5, atol=1e-5 * np.abs(exact).max())
    g.close()


def test_clustered_antiradiance_wide_disks(gpu_ctx, cornell):
    """Disks half the scene wide: most rays reach most clusters, so a block's
    appends overflow its LDS buffer into direct appends, and the sums still
    equal the exact gather's."""
    sc = cornell
    print("[gather] clustered vs exact, wide disks", flush=True)
    g = make_gather(gpu_ctx, sc)
    av = DeviceAvpls.from_scene(sc)
    base = dict(useLightCuts=0, antiradianceModel=ffi.ARC_ANTIRAD_DISK, photonRadiusScale=50.0)
    g.run(av, sc.cam, sc.extent, default_params(antiradianceMode=ffi.ARC_ANTI_EXACT, **base))
    gpu_ctx.sync()
    exact = g.antiradiance[:, :3].double().cpu().numpy()
    st = g.run(av, sc.cam, sc.extent, default_params(antiradianceMode=ffi.ARC_ANTI_CLUSTERED, **base),
               ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    clus = g.antiradiance[:, :3].double().cpu().numpy()
    n_anti = int((np.abs(sc.avpl_anti).sum(axis=1) > 0).sum())
    lists = st.antiVisitedNodes / 16  # each list entry is filtered against the 16 tiles of its cluster
    print(f"[gather] wide disks: {lists / max(n_anti, 1):.1f} clusters per antiradiance AVPL, "
          f"rel-L2 {rel_l2(clus, exact):.3e}", flush=True)
    assert lists > 16 * n_anti  # more than 4096 appends per 256-AVPL block on average
    assert np.abs(exact).max() > 0
    np.testing.assert_allclose(clus, exact, rtol=1e-5, atol=1e-5 * np.abs(exact).max())
    g.close()
