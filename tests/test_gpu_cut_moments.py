"""The moment estimator of the error-bound cut (ARC_ESTIMATOR_MOMENTS)
through the C ABI.

What fixes it without a reference implementation (the reference's cut sums a
random representative, cudaGather.cu:416-420, bvh.cu:726-744):
  * its node records equal FP64 sums over each node's leaf range (count,
    first leaf, per-channel I, I n and I n n^T, the intensity-weighted
    position mean, and four points whose mean and covariance are the node's);
    the centroid records of the same tree are checked the same way, over
    several 32K-leaf refit spans and at 1M AVPLs;
  * errorBound = 0 refines every node that can contribute down to nodes of
    at most cutExactCount AVPLs, summed exactly: the exact FP64 radiance
    (relative L2 <= 1e-5) for exact counts 0, 1, 16 and 128;
  * a node whose AVPLs share one position is summed exactly by the four
    points, so the estimate equals sum_c max(sum_j I_jc f2(t_j), 0) G_p with
    f2(t) = 3/32 + t/2 + 15/32 t^2 evaluated per AVPL (the addition theorem
    that lets the node keep three moments instead of its AVPLs);
  * on the 1080p x 1M frame the Lightcuts bound at 1 % stays within the
    measured error, with cone antiradiance.
"""
import math

import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import DeviceAvpls, Gather, default_params, device_array
from tests import moment_checks as M

pytestmark = pytest.mark.gpu


def mom_params(**kw):
    base = dict(useLightCuts=1, cutMode=ffi.ARC_CUT_ERROR_BOUND, cutPerTile=1,
                cutEstimator=ffi.ARC_ESTIMATOR_MOMENTS, cutCriterion=ffi.ARC_CRITERION_BOUND,
                cutImageFloor=0.0, errorBound=0.01, cutExactCount=16)
    base.update(kw)
    return default_params(**base)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_rig(gpu_ctx, sc):
    g = Gather(gpu_ctx, sc.width, sc.height, sc.materials)
    g.set_gbuffer(device_array(sc.pos4), device_array(sc.nrm4), device_array(sc.albedo4))
    return g, DeviceAvpls.from_scene(sc)


@pytest.mark.parametrize("num_avpls", [100_000, 1_000_000])
def test_node_records_are_range_sums(gpu_ctx, num_avpls):
    sc = S.clutter_1080p(width=64, height=48, num_avpls=num_avpls)
    print(f"[mom] records n={num_avpls}", flush=True)
    g, av = make_rig(gpu_ctx, sc)
    try:
        g.run(av, sc.cam, sc.extent, mom_params())
        gpu_ctx.sync()
        tree = g.bvh_to_host()
        crec = g.cut_records(ffi.ARC_RECORDS_CENTROID)
        mom = g.cut_records(ffi.ARC_RECORDS_MOMENTS)
    finally:
        g.close()
    n = tree["codes"].shape[0]
    assert n == num_avpls
    ranges = np.asarray(tree["ranges"]).reshape(-1, 2).astype(np.int64)
    # the leaves' records hold the sorted AVPLs
    perm = tree["perm"]
    # a leaf's centroid is (w x) / w in FP32: its position within an ulp or two
    np.testing.assert_allclose(crec[:n, 0, :3], sc.avpl_pos[perm], rtol=3e-7, atol=1e-7)
    diff = sc.materials[sc.avpl_mat[perm], 3:6].astype(np.float32)
    want_I = sc.avpl_rad[perm] * (np.float32(0.318309886) * diff)
    np.testing.assert_allclose(crec[:n, 2, :3], want_I, rtol=1e-6)
    nrm = sc.avpl_nrm[perm].astype(np.float64)
    np.testing.assert_allclose(crec[:n, 1, :3], nrm / np.linalg.norm(nrm, axis=1, keepdims=True), atol=1e-6)
    live_c = M.check_centroid_records(crec, ranges, n)
    live_m = M.check_moment_records(mom, crec, ranges, n, tree["nodes"])
    # nodes with sum(w) above 1e-6 of the root: 98 % at 100K AVPLs, 81 % at 1M
    assert live_c > 0.75 * (n - 1) and live_m > 0.75 * (n - 1)


@pytest.fixture(scope="module")
def small(gpu_ctx):
    sc = S.clutter_1080p(width=96, height=64, num_avpls=20_000, num_tris=20_000)
    g, av = make_rig(gpu_ctx, sc)
    p = mom_params(antiradianceModel=ffi.ARC_ANTIRAD_CONE)
    r64, a64 = g.exact_f64(av, sc.cam, sc.extent, p)
    yield sc, g, av, r64.cpu().numpy(), a64.cpu().numpy()
    g.close()


def run(gpu_ctx, g, av, sc, p):
    st = g.run(av, sc.cam, sc.extent, p, ffi.ARC_RUN_COUNTERS)
    gpu_ctx.sync()
    return g.radiance[:, :3].double().cpu().numpy(), g.result[:, :3].double().cpu().numpy(), st


@pytest.mark.parametrize("criterion", [ffi.ARC_CRITERION_BOUND, ffi.ARC_CRITERION_EXPANSION])
def test_zero_bound_is_exact(gpu_ctx, small, criterion):
    sc, g, av, r64, a64 = small
    print(f"[mom] zero bound criterion={criterion}", flush=True)
    rad, res, st = run(gpu_ctx, g, av, sc, mom_params(errorBound=0.0, cutCriterion=criterion, cutExactCount=0,
                                                      antiradianceModel=ffi.ARC_ANTIRAD_CONE))
    assert st.errorPixels == 0 and st.budgetHits == 0
    assert rel_l2(rad, r64) <= 1e-5
    assert rel_l2(res, r64 - a64) <= 1e-5


def test_exact_count_sums_small_nodes_exactly(gpu_ctx, small):
    """errorBound = 0 refines down to nodes of at most cutExactCount AVPLs,
    which are summed exactly: the exact radiance for every exact count."""
    sc, g, av, r64, a64 = small
    for ec in (1, 16, 128):
        print(f"[mom] exact count {ec}", flush=True)
        rad, _, st = run(gpu_ctx, g, av, sc, mom_params(errorBound=0.0, cutExactCount=ec,
                                                        antiradianceModel=ffi.ARC_ANTIRAD_CONE))
        assert st.exactLeaves > 0 and st.exactLeaves <= sc.num_avpls * sc.npix
        assert rel_l2(rad, r64) <= 1e-5
    with pytest.raises(Exception):
        run(gpu_ctx, g, av, sc, mom_params(cutExactCount=129))


def test_bound_tightens_error(gpu_ctx, small):
    sc, g, av, r64, a64 = small
    errs, work = [], []
    for eb in (0.1, 0.01, 0.001):
        rad, _, st = run(gpu_ctx, g, av, sc, mom_params(errorBound=eb, antiradianceModel=ffi.ARC_ANTIRAD_CONE))
        errs.append(rel_l2(rad, r64))
        work.append(st.evaluatedNodes)
        print(f"[mom] bound {eb}: rel-L2 {errs[-1]:.2e}, evaluated {st.evaluatedNodes}", flush=True)
    assert errs[0] > errs[1] > errs[2] and work[0] < work[1] < work[2]
    assert errs[1] < 1e-2


def test_tile_order_keeps_the_image(gpu_ctx):
    """A frame's tiles launch in the order of the last frame's tile costs,
    heaviest first; a tile's sums do not depend on when it runs, so the
    second frame (ordered) equals the first (launch order) bit for bit."""
    sc = S.clutter_1080p(width=200, height=120, num_avpls=60_000)
    print("[mom] ordered launch", flush=True)
    g, av = make_rig(gpu_ctx, sc)
    try:
        p = mom_params(antiradianceModel=ffi.ARC_ANTIRAD_CONE)
        first, _, st1 = run(gpu_ctx, g, av, sc, p)
        second, _, st2 = run(gpu_ctx, g, av, sc, p)
        assert np.array_equal(first, second)
        assert st1.evaluatedNodes == st2.evaluatedNodes and st1.testedNodes == st2.testedNodes
    finally:
        g.close()


@pytest.mark.parametrize("w,h,n", [(200, 120, 60_000), (96, 64, 300_000), (203, 117, 50_000)])
def test_estimate_pool_matches_traversal_estimates(gpu_ctx, w, h, n):
    """Kept nodes estimated on the matrix cores (the default pool), in the
    traversal (pool 0) and mixed (one chunk per tile: the pool fills and the
    later tiles estimate in place) give one cut and one image up to
    rounding; the matrix-core frame repeats bit for bit."""
    sc = S.clutter_1080p(width=w, height=h, num_avpls=n)
    print(f"[mom] estimate pool {w}x{h} n={n}", flush=True)
    g, av = make_rig(gpu_ctx, sc)
    try:
        p = mom_params(antiradianceModel=ffi.ARC_ANTIRAD_CONE)
        g.set_estimate_pool(0)
        ref, _, st0 = run(gpu_ctx, g, av, sc, p)
        assert np.isfinite(ref).all() and ref.max() > 0
        for pool in (32, 1):
            g.set_estimate_pool(pool)
            rad, _, st = run(gpu_ctx, g, av, sc, p)
            e = rel_l2(rad, ref)
            worst = float(np.max(np.abs(rad - ref)) / np.sqrt(np.mean(ref ** 2)))
            print(f"[mom] pool {pool}: rel-L2 {e:.2e} vs estimates in the traversal, max {worst:.2e} x rms",
                  flush=True)
            assert st.evaluatedNodes == st0.evaluatedNodes and st.testedNodes == st0.testedNodes
            assert e <= 1e-5 and worst <= 1e-4
        g.set_estimate_pool(32)
        first, _, _ = run(gpu_ctx, g, av, sc, p)
        second, _, _ = run(gpu_ctx, g, av, sc, p)
        assert np.array_equal(first, second)
        with pytest.raises(ffi.ArcError):
            g.set_estimate_pool(-1)
    finally:
        g.close()


def test_launch_shapes_keep_the_image(gpu_ctx):
    """The traversal stack (64, 96, 128 nodes) and the antiradiance's list
    entries per pass (768, 1536, 3072) change the launch, not the image: the
    same cut bit for bit while no tile overflows its stack, the antiradiance
    to rounding (its entries are summed in the order of the LDS appends).
    The tuner keeps a schedule without error pixels and the same image."""
    from antiradiancecuts_amd import autotune

    sc = S.clutter_1080p(width=200, height=120, num_avpls=60_000)
    print("[mom] launch shapes", flush=True)
    g, av = make_rig(gpu_ctx, sc)
    try:
        p = mom_params(antiradianceModel=ffi.ARC_ANTIRAD_CONE, antiradianceMode=ffi.ARC_ANTI_CLUSTERED)
        rad0, res0, st0 = run(gpu_ctx, g, av, sc, p)
        assert st0.errorPixels == 0
        for stack in (64, 128):
            g.set_launch_shape(stack, 1536)
            rad, _, st = run(gpu_ctx, g, av, sc, p)
            print(f"[mom] stack {stack}: error pixels {st.errorPixels}", flush=True)
            if st.errorPixels == 0:
                assert np.array_equal(rad, rad0) and st.evaluatedNodes == st0.evaluatedNodes
        for chunk in (768, 3072):
            g.set_launch_shape(96, chunk)
            rad, res, st = run(gpu_ctx, g, av, sc, p)
            e = rel_l2(res, res0)
            print(f"[mom] antiradiance chunk {chunk}: result rel-L2 {e:.2e}", flush=True)
            assert np.array_equal(rad, rad0) and e <= 1e-6
        for bad in ((32, 1536), (96, 1000)):
            with pytest.raises(ffi.ArcError):
                g.set_launch_shape(*bad)
        g.set_launch_shape(96, 1536)
        best, sched, table = autotune.tune(g, av, sc.cam, sc.extent, p, reps=2)
        print(f"[mom] tuned {sched} from {table}", flush=True)
        rad, res, st = run(gpu_ctx, g, av, sc, best)
        assert st.errorPixels == 0 and rel_l2(res, res0) <= 1e-5
    finally:
        g.close()


@pytest.mark.parametrize("exact_count", [0, 16])
def test_nodes_facing_away_add_nothing(gpu_ctx, exact_count):
    """AVPLs on the plane z = 1 with normals +z, above pixels on z = 0 that
    face +z: every AVPL faces away from every pixel, so the exact radiance is
    0.  Kept nodes (one, the root, at a huge bound) and exact leaves are both
    recognised as facing away from the tile (normal box against the node and
    tile boxes) and add nothing, where the second-order estimate would leave
    a positive residue (its clamped cosine is 2/32 at t = -1).  Flipped to
    face the pixels, the same AVPLs light them."""
    rng = np.random.default_rng(5)
    P, N = 64, 4096
    pos = np.zeros((P, 4), np.float32)
    pos[:, 0] = rng.uniform(-0.5, 0.5, P)
    pos[:, 1] = rng.uniform(-0.5, 0.5, P)
    pos[:, 3] = 1
    nrm = np.zeros((P, 4), np.float32)
    nrm[:, 2] = 1
    nrm[:, 3] = 1
    alb = np.zeros((P, 4), np.float32)
    alb[:, :3] = 0.5
    a = np.stack([rng.uniform(-2, 2, N), rng.uniform(-2, 2, N), np.ones(N)], 1).astype(np.float32)
    L = rng.uniform(0.5, 2.0, (N, 3)).astype(np.float32)
    mats = np.array([[0] * 10, [0, 0, 0, 0.5, 0.6, 0.7, 0, 0, 0, 0]], np.float32)
    out = {}
    for name, nz in (("away", 1.0), ("facing", -1.0)):
        an = np.zeros((N, 3), np.float32)
        an[:, 2] = nz
        sc = S.Scene(8, 8, pos, nrm, alb, mats, np.array([0, 0, 10], np.float32), 4.0, a, an, -an, L,
                     np.zeros_like(L), np.ones(N, np.int32), np.ones(N, np.int32))
        g, av = make_rig(gpu_ctx, sc)
        try:
            rad, _, st = run(gpu_ctx, g, av, sc, mom_params(errorBound=1e30 if exact_count == 0 else 0.01,
                                                            cutExactCount=exact_count))
        finally:
            g.close()
        out[name] = rad
        print(f"[mom] AVPLs {name}, exact <= {exact_count}: max radiance {rad.max():.3e}", flush=True)
    assert np.all(out["away"] == 0.0)
    assert out["facing"].min() > 0.0


def test_colocated_node_is_the_sh2_sum(gpu_ctx):
    """Three AVPLs at one point with different normals and colours, seen from
    pixels far away: the root is kept (huge bound) and estimated; its four
    points coincide with the AVPLs' position, so the estimate is
    sum_c max(sum_j I_jc f2(-n_j . d^), 0) * max(n_p . d, 0) / |d|^3."""
    rng = np.random.default_rng(4)
    P = 64
    pos = np.zeros((P, 4), np.float32)
    pos[:, 0] = rng.uniform(-0.5, 0.5, P)
    pos[:, 1] = rng.uniform(-0.5, 0.5, P)
    pos[:, 3] = 1
    nrm = np.zeros((P, 4), np.float32)
    nrm[:, 2] = 1
    nrm[:, 3] = 1
    alb = np.zeros((P, 4), np.float32)
    alb[:, :3] = 0.5
    a = np.array([0.1, -0.2, 2.0], np.float32)
    an = np.array([[0, 0, -1], [0.6, 0, -0.8], [0, 0.8, 0.6]], np.float32)
    L = np.array([[1, 2, 3], [4, 0.5, 0.1], [0.3, 3, 1]], np.float32)
    mats = np.array([[0] * 10, [0, 0, 0, 0.5, 0.6, 0.7, 0, 0, 0, 0]], np.float32)
    sc = S.Scene(8, 8, pos, nrm, alb, mats, np.array([0, 0, 10], np.float32), 1.0, np.tile(a, (3, 1)), an,
                 -an, L, np.zeros_like(L), np.ones(3, np.int32), np.ones(3, np.int32))
    g, av = make_rig(gpu_ctx, sc)
    try:
        rad, _, st = run(gpu_ctx, g, av, sc, mom_params(errorBound=1e30, cutExactCount=0))
    finally:
        g.close()
    assert st.evaluatedNodes == P and st.exactLeaves == 0
    I = L.astype(np.float64) * (np.float32(0.318309886) * mats[1, 3:6]).astype(np.float64)
    want = np.zeros((P, 3))
    for q in range(P):
        d = a.astype(np.float64) - pos[q, :3]
        dl = np.linalg.norm(d)
        t = -(an.astype(np.float64) @ d) / dl
        f2 = 3 / 32 + t / 2 + 15 / 32 * t * t
        A = np.maximum((I * f2[:, None]).sum(0), 0)
        want[q] = A * max(d[2], 0) / dl ** 3 * (0.5 / math.pi)
    np.testing.assert_allclose(rad, want, rtol=2e-5)
