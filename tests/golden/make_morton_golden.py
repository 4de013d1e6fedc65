"""Write tests/golden/morton_ref.npz from the reference's own morton.h.

Inputs and outputs only: the codes come from oracle/_ref/libref_morton.so,
which oracle/Makefile compiles from Framework/morton.h where the reference
tree lies (make -C oracle ref).  The fixture travels with the tests, so the
oracle and the GPU encode are pinned to the reference's codes where the
reference tree is absent.

    python tests/golden/make_morton_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def main() -> int:
    from oracle import pyoracle as orc

    R = orc.ref_morton()
    if R is None:
        print("oracle/_ref/libref_morton.so is not built (make -C oracle ref)", file=sys.stderr)
        return 1
    rng = np.random.default_rng(20261202)
    # 3D: points in [0,1]^3 with the corners 0 and 1 present, so the
    # reference normalisation (v - min) / (max - min) is the identity on them,
    # plus values on and next to the quantisation steps k / 1024.
    steps = (np.arange(0, 1025, 1, dtype=np.float64) / 1024.0).astype(np.float32)
    near = np.concatenate([steps, np.nextafter(steps, np.float32(2)), np.nextafter(steps, np.float32(-1))])
    near = np.clip(near, 0.0, 1.0).astype(np.float32)
    edge = np.stack([near, rng.permutation(near), rng.permutation(near)], 1)
    uni = rng.random((2048, 3), dtype=np.float32)
    p3 = np.concatenate([np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32), edge, uni]).astype(np.float32)
    c3 = np.zeros(len(p3), np.uint32)
    R.ref_morton3d_batch(p3.ctypes.data_as(orc.vp), len(p3), c3.ctypes.data_as(orc.vp))
    # 6D: position and normal-like components, full 64-bit code
    # (rows of 0 and 1 make the per-half normalisation the identity)
    p6 = np.concatenate([np.zeros((1, 6), np.float32), np.ones((1, 6), np.float32),
                         rng.uniform(0.0, 1.0, size=(2046, 6)).astype(np.float32)])
    c6 = np.array([R.ref_morton6d(*map(float, a)) for a in p6], np.uint64)
    # expandBitsBy2 on every 10-bit input
    e2 = np.array([R.ref_expand_bits_by2(x) for x in range(1024)], np.uint32)
    out = Path(__file__).resolve().parent / "morton_ref.npz"
    np.savez_compressed(out, p3=p3, c3=c3, p6=p6, c6=c6, e2=e2)
    print(f"wrote {out}: {len(p3)} 3D codes, {len(p6)} 6D codes, {len(e2)} expandBitsBy2 values")
    return 0


if __name__ == "__main__":
    sys.exit(main())
