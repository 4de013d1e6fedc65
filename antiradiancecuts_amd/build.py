"""Build the native libraries in-tree.

libarc.so        HIP kernels + C ABI (include/arc.h), gfx950 only
libarc_scene.so  host C++ synthetic scene / AVPL / point generators
                 (include/arc_scene.h)
_build/arc_facade_test
                 host program over the C++ facade (include/arc_cudagather.hpp),
                 run by the GPU tests

Run as ``python -m antiradiancecuts_amd.build`` or through
``__graft_entry__.build()``.  Every HIP translation unit is compiled with
packed FP32 instructions disabled (``-packed-fp32-ops``); the LBVH unit also
with ``-ffp-contract=off`` so its float results match the CPU restatement bit
for bit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OBJDIR = PKG / "_build"

HIP_SOURCES = {
    "capi.hip": [],
    "sort.hip": [],
    "lbvh.hip": ["-ffp-contract=off"],
    "gather.hip": [],
    "gather_anti.hip": [],
    "gather_cut.hip": [],
    "cut_moments.hip": [],
    "noalloc.hip": [],
    "output.hip": [],
    "trace.hip": ["-ffp-contract=off"],
    "batch.hip": [],
    "cutnet.hip": [],
}

HIPCC_FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")


def build_hip(verbose: bool = False) -> Path:
    """Compile every HIP source for gfx950 and link libarc.so."""
    hipcc = _hipcc()
    OBJDIR.mkdir(exist_ok=True)
    jobs = []
    for src, extra in HIP_SOURCES.items():
        obj = OBJDIR / (Path(src).stem + ".o")
        jobs.append([hipcc, *HIPCC_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        for _ in ex.map(_run, jobs):
            pass
    out = PKG / "libarc.so"
    tmp = PKG / "libarc.so.tmp"
    objs = [str(OBJDIR / (Path(s).stem + ".o")) for s in HIP_SOURCES]
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_scene(verbose: bool = False) -> Path:
    """Compile the host-only synthetic scene generator."""
    out = PKG / "libarc_scene.so"
    tmp = PKG / "libarc_scene.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", f"-I{INCLUDE}",
          "-o", str(tmp), str(CSRC / "arc_scene.cpp")])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_facade_test(verbose: bool = False) -> Path:
    """Compile tests/cpp/facade_test.cpp against the two libraries."""
    OBJDIR.mkdir(exist_ok=True)
    out = OBJDIR / "arc_facade_test"
    tmp = OBJDIR / "arc_facade_test.tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", "-o", str(tmp),
          str(ROOT / "tests" / "cpp" / "facade_test.cpp"), f"-L{PKG}", "-larc", "-larc_scene",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_all(verbose: bool = False) -> None:
    build_scene(verbose)
    build_hip(verbose)
    build_facade_test(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
