"""Morton codes against the fixture written from the reference's own morton.h
(tests/golden/make_morton_golden.py): the oracle on the CPU and the GPU
encode through arc_bvh_build_points.

The fixture's point sets contain the corners 0 and 1 on every axis, so the
reference normalisation (v - min) / (max - min) (bvh.cu:73-81) is the
identity on them and the sorted codes of a build are the fixture's codes."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "morton_ref.npz"


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_expand_bits_by2(orc, golden):
    L = orc.lib()
    got = np.array([L.orc_expand_bits_by2(x) for x in range(1024)], np.uint32)
    np.testing.assert_array_equal(got, golden["e2"])


def test_oracle_morton3d(orc, golden):
    L = orc.lib()
    got = np.array([L.orc_morton3d(*map(float, p)) for p in golden["p3"]], np.uint32)
    np.testing.assert_array_equal(got, golden["c3"])


def test_oracle_morton6d(orc, golden):
    L = orc.lib()
    got = np.array([L.orc_morton6d(*map(float, a)) for a in golden["p6"]], np.uint64)
    np.testing.assert_array_equal(got, golden["c6"])


def _gpu_codes(ctx, pos, nrm=None):
    from antiradiancecuts_amd.runtime import Bvh, device_array

    b = Bvh(ctx, nrm is not None)
    try:
        b.build_points(device_array(pos), device_array(nrm) if nrm is not None else None)
        return b.to_host()
    finally:
        b.close()


@pytest.mark.gpu
def test_gpu_codes_3d(gpu_ctx, golden):
    print("[golden] 3D codes", flush=True)
    out = _gpu_codes(gpu_ctx, golden["p3"])
    np.testing.assert_array_equal(out["codes"], golden["c3"][out["perm"]])
    assert np.all(np.diff(out["codes"].astype(np.int64)) >= 0)


@pytest.mark.gpu
def test_gpu_codes_6d_truncated(gpu_ctx, golden):
    """considerNormals: the 60-bit code kept to its low 32 bits (bvh.h:17)."""
    print("[golden] 6D codes", flush=True)
    p6 = golden["p6"]
    out = _gpu_codes(gpu_ctx, np.ascontiguousarray(p6[:, :3]), np.ascontiguousarray(p6[:, 3:]))
    want = (golden["c6"] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    np.testing.assert_array_equal(out["codes"], want[out["perm"]])
