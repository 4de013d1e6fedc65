"""GPU LBVH vs the oracle, bit for bit: codes, order, topology, leaf ranges,
boxes, parents and AvplBvh node data (through the C ABI)."""
import sys

import numpy as np
import pytest

from antiradiancecuts_amd import ffi
from antiradiancecuts_amd import scene as S
from antiradiancecuts_amd.runtime import Bvh, DeviceAvpls, device_array
from tests import lbvh_checks as K
from tests.test_oracle_lbvh import cases

pytestmark = pytest.mark.gpu

FIELDS = ("codes", "perm", "parents", "ranges", "positions")


def compare(gpu: dict, ref: dict, node_data=False):
    for f in FIELDS:
        np.testing.assert_array_equal(gpu[f], np.asarray(ref[f]).reshape(gpu[f].shape), err_msg=f)
    for f in ("left", "right", "bbMin", "bbMax", "visited"):
        np.testing.assert_array_equal(gpu["nodes"][f], ref["nodes"][f], err_msg=f"nodes.{f}")
    if node_data:
        for f in ("nd_size", "nd_mat", "nd_pos", "nd_nrm", "nd_rad", "nd_dir"):
            np.testing.assert_array_equal(gpu[f], ref[f], err_msg=f)


def build_points(ctx, pos, nrm=None, consider_normals=False):
    b = Bvh(ctx, consider_normals)
    try:
        b.build_points(device_array(pos), device_array(nrm) if nrm is not None else None)
        out = b.to_host()
        out["max_stack"] = b.check_integrity()
        return out
    finally:
        b.close()


@pytest.mark.parametrize("name,pos", list(cases()), ids=[c[0] for c in cases()])
def test_points_bit_exact(gpu_ctx, orc, name, pos):
    print(f"[lbvh] case {name} n={len(pos)}", flush=True)
    g = build_points(gpu_ctx, pos)
    compare(g, orc.lbvh_build(pos))


def test_points_1m_mixture(gpu_ctx, orc):
    """1,000,000 points: half uniform, half 32 Gaussians (sigma 0.01)."""
    pos = S.points(1_000_000, seed=1)
    print("[lbvh] 1M mixture", flush=True)
    g = build_points(gpu_ctx, pos)
    ref = orc.lbvh_build(pos)
    compare(g, ref)
    K.check_karras_ranges(g["ranges"])
    # sortedness and permutation properties at full size
    assert np.all(np.diff(g["codes"].astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(g["perm"]), np.arange(len(pos)))


def test_points_6d(gpu_ctx, orc):
    rng = np.random.default_rng(9)
    pos = rng.random((5000, 3), dtype=np.float32)
    nrm = S.unit_normals(5000, seed=9)
    g = build_points(gpu_ctx, pos, nrm, consider_normals=True)
    compare(g, orc.lbvh_build(pos, nrm, consider_normals=True))


def test_too_few_points(gpu_ctx):
    b = Bvh(gpu_ctx)
    with pytest.raises(ffi.ArcError) as e:
        b.build_points(device_array(np.zeros((1, 3), np.float32)))
    assert e.value.status == ffi.ARC_E_TOO_FEW_POINTS
    b.close()


def avpl_tree(ctx, sc, seed=77):
    b = Bvh(ctx)
    try:
        b.build_avpls(DeviceAvpls.from_scene(sc), seed=seed)
        return b.to_host()
    finally:
        b.close()


@pytest.mark.parametrize("n", [2, 3, 4096, 100_000])
def test_avpl_bvh_bit_exact(gpu_ctx, orc, n):
    sc = S.cornell_256(seed=n, width=4, height=4, num_avpls=n)
    print(f"[lbvh] avpl n={n}", flush=True)
    g = avpl_tree(gpu_ctx, sc)
    ref = orc.avpl_bvh_build(sc, g["nd_rand"])
    compare(g, ref, node_data=True)
    if n <= 5000:
        K.check_topology(g["codes"], g["nodes"], g["parents"], g["ranges"])
        K.check_cluster(g, sc.avpl_rad[g["perm"]])


def test_avpl_bvh_1m(gpu_ctx, orc):
    sc = S.clutter_1080p(width=8, height=8)  # 1M AVPLs, tiny G-buffer
    print("[lbvh] avpl 1M", flush=True)
    g = avpl_tree(gpu_ctx, sc)
    compare(g, orc.avpl_bvh_build(sc, g["nd_rand"]), node_data=True)


def test_avpl_bvh_rejects_bounce0(gpu_ctx):
    sc = S.cornell_256(seed=1, width=4, height=4, num_avpls=100)
    sc.avpl_bounce[17] = 0
    b = Bvh(gpu_ctx)
    with pytest.raises(ffi.ArcError) as e:
        b.build_avpls(DeviceAvpls.from_scene(sc))
    assert e.value.status == ffi.ARC_E_INVALID
    b.close()


def test_rebuild_reuses_object(gpu_ctx, orc):
    """Two builds of different sizes on one object (buffers grow, results stay exact)."""
    b = Bvh(gpu_ctx)
    rng = np.random.default_rng(3)
    for n in (3000, 700, 20000):
        pos = rng.random((n, 3), dtype=np.float32)
        b.build_points(device_array(pos))
        compare(b.to_host(), orc.lbvh_build(pos))
    b.close()
