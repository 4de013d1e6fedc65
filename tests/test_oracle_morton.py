"""Pins of the oracle's Morton restatement (morton.h:24-107).

Checks: exhaustive comparison of every 10-bit input against a naive bit loop,
closed-form known answers, quantisation special cases (clamp, negative, NaN),
and — where oracle/_ref was built from the reference's own morton.h — equality
with the reference on every 10-bit input and on random floats.
"""
import numpy as np
import pytest


def naive_spread(x: int, stride: int, bits: int) -> int:
    r = 0
    for i in range(bits):
        if (x >> i) & 1:
            r |= 1 << (stride * i)
    return r


def test_expand_bits_by2_exhaustive(orc):
    L = orc.lib()
    for x in range(1024):
        assert L.orc_expand_bits_by2(x) == naive_spread(x, 3, 10), x
    # bits above the low 10 are masked off
    assert L.orc_expand_bits_by2(0xFFFFFC00) == 0
    assert L.orc_expand_bits_by2(0xFFFFFFFF) == naive_spread(1023, 3, 10)


def test_expand_bits_by1_exhaustive(orc):
    L = orc.lib()
    for x in list(range(4096)) + [0xFFFF, 0x8000, 0x1234]:
        assert L.orc_expand_bits_by1(x) == naive_spread(x, 2, 16), x


def test_interleave3_known_answers(orc):
    L = orc.lib()
    assert L.orc_interleave3(1, 0, 0) == 1
    assert L.orc_interleave3(0, 1, 0) == 2
    assert L.orc_interleave3(0, 0, 1) == 4
    assert L.orc_interleave3(1023, 1023, 1023) == 0x3FFFFFFF
    assert L.orc_interleave3(2, 0, 0) == 8
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(500, 3)):
        want = naive_spread(int(x), 3, 10) | (naive_spread(int(y), 3, 10) << 1) | (naive_spread(int(z), 3, 10) << 2)
        assert L.orc_interleave3(int(x), int(y), int(z)) == want


def test_interleave6_bit_positions(orc):
    L = orc.lib()
    for k in range(6):
        for i in range(10):
            args = [0] * 6
            args[k] = 1 << i
            assert L.orc_interleave6(*args) == 1 << (6 * i + k)
    assert L.orc_interleave6(*([1023] * 6)) == (1 << 60) - 1


def test_morton3d_quantisation(orc):
    L = orc.lib()
    assert L.orc_morton3d(0.0, 0.0, 0.0) == 0
    assert L.orc_morton3d(1.0, 1.0, 1.0) == 0x3FFFFFFF          # 1024 clamps to 1023
    assert L.orc_morton3d(-0.5, -3.0, -1e-30) == 0               # negatives clamp to 0
    assert L.orc_morton3d(float("nan"), 0.0, 0.0) == 0           # fmaxf(NaN, 0) = 0
    assert L.orc_morton3d(2.0, 0.0, 0.0) == L.orc_interleave3(1023, 0, 0)
    # truncation, not rounding: 1023.9/1024 -> 1023, 0.99/1024 -> 0
    assert L.orc_morton3d(1023.9 / 1024, 0.0, 0.0) == L.orc_interleave3(1023, 0, 0)
    assert L.orc_morton3d(0.99 / 1024, 0.0, 0.0) == 0
    assert L.orc_morton3d(1.0 / 1024, 0.0, 0.0) == 1


def test_morton6d_truncation_low32(orc):
    L = orc.lib()
    full = L.orc_morton6d(1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
    assert full == (1 << 60) - 1
    assert (full & 0xFFFFFFFF) == 0xFFFFFFFF


@pytest.mark.skipif(not __import__("oracle.pyoracle", fromlist=["x"]).REF_LIB.exists(),
                    reason="oracle/_ref not built (reference tree absent)")
def test_against_reference_morton_h(orc):
    """The reference's own morton.h, compiled in place, on every 10-bit input
    and on 200k random floats including out-of-range and NaN values."""
    R = orc.ref_morton()
    L = orc.lib()
    for x in range(1024):
        assert R.ref_expand_bits_by2(x) == L.orc_expand_bits_by2(x)
    for x in range(0, 65536, 7):
        assert R.ref_expand_bits_by1(x) == L.orc_expand_bits_by1(x)
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, size=(2000, 3)):
        assert R.ref_interleave3(int(x), int(y), int(z)) == L.orc_interleave3(int(x), int(y), int(z))
    for v in rng.integers(0, 1024, size=(500, 6)):
        assert R.ref_interleave6(*map(int, v)) == L.orc_interleave6(*map(int, v))
    pts = rng.uniform(-0.1, 1.1, size=(200_000, 3)).astype(np.float32)
    pts[::97] = np.nan
    ref = np.zeros(len(pts), np.uint32)
    R.ref_morton3d_batch(pts.ctypes.data_as(orc.vp), len(pts), ref.ctypes.data_as(orc.vp))
    mine = np.array([L.orc_morton3d(*map(float, p)) for p in pts[:20000]], np.uint32)
    np.testing.assert_array_equal(mine, ref[:20000])
    for a in rng.uniform(-0.1, 1.1, size=(2000, 6)).astype(np.float32):
        assert R.ref_morton6d(*map(float, a)) == L.orc_morton6d(*map(float, a))
