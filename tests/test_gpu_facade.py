"""The C++ facade (include/arc_cudagather.hpp) driven by a host program with
the reference renderer's call sequences (Render.cpp:64-81, 174-181,
297-301): SimpleBvh smoke points, CudaGather exact and light-cut frames,
AvplBvh getters.  The program is built in-tree by antiradiancecuts_amd/build.py."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "antiradiancecuts_amd" / "_build" / "arc_facade_test"


@pytest.mark.gpu
def test_cpp_facade_program():
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built (python -m antiradiancecuts_amd.build)")
    print("running", EXE, flush=True)
    proc = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=100)
    print(proc.stdout[-4000:], proc.stderr[-2000:])
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "facade ok" in proc.stdout
