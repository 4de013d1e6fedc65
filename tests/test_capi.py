"""The C ABI library loads and exports every symbol its headers declare.

CPU-only: no compute entry point is called (no GPU here); only the pure host
functions (version, status strings, parameter defaults) run.
"""
import ctypes
import math
import re
from pathlib import Path

import pytest

from antiradiancecuts_amd import ffi

ROOT = Path(__file__).resolve().parents[1]
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(arc_\w+)\s*\(", re.M)


def declared(header: str):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


def test_arc_h_symbols_exported():
    names = declared("arc.h")
    assert len(names) >= 25
    lib = ctypes.CDLL(str(ffi.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python bindings cover the whole header
    assert sorted(n for n, _, _ in ffi.ARC_SYMBOLS) == names


def test_arc_noalloc_h_symbols_exported():
    names = declared("arc_noalloc.h")
    assert names == ["arc_bvh_build_points_noalloc"]
    lib = ctypes.CDLL(str(ffi.LIB_PATH))
    assert hasattr(lib, names[0])
    assert sorted(n for n, _, _ in ffi.NOALLOC_SYMBOLS) == names
    assert ctypes.sizeof(ffi.arc_lbvh_outputs) == 7 * ctypes.sizeof(ctypes.c_void_p)


def test_arc_scene_h_symbols_exported():
    names = declared("arc_scene.h")
    lib = ctypes.CDLL(str(ffi.SCENE_LIB_PATH))
    assert not [n for n in names if not hasattr(lib, n)]
    assert sorted(n for n, _, _ in ffi.SCENE_SYMBOLS) == names


def test_host_only_entry_points():
    L = ffi.lib()
    assert L.arc_version().decode().startswith("arc ")
    assert L.arc_status_string(ffi.ARC_E_TOO_FEW_POINTS).decode() == "not enough points for bvh construction"
    assert L.arc_status_string(12345).decode() == "unknown status"
    p = ffi.arc_gather_params()
    L.arc_gather_params_default(ctypes.byref(p))
    # CConfigManager.cpp:15,26,37 and STACK_SIZE (cudaGather.cu:19)
    assert p.useLightCuts == 1
    assert p.photonRadiusScale == pytest.approx(2.0)
    assert p.clusterRefinementThreshold == pytest.approx(math.pi / 10, rel=1e-6)
    assert p.stackSize == 32
    assert p.antiradianceModel == ffi.ARC_ANTIRAD_DISK


def test_struct_layouts_match_header():
    assert ctypes.sizeof(ffi.arc_bvh_node) == 36
    assert ctypes.sizeof(ffi.arc_material) == 40
    assert ctypes.sizeof(ffi.arc_float3) == 12
    assert ctypes.sizeof(ffi.arc_avpls) == 8 * 7 + 8


def _c_layout(struct: str, fields):
    """sizeof and offsetof as the C compiler lays the header's struct out."""
    import subprocess
    import tempfile

    body = "".join(f'    printf("%zu\\n", offsetof({struct}, {f}));\n' for f in fields)
    src = ("#include <stddef.h>\n#include <stdio.h>\n#include \"arc.h\"\nint main(void) {\n"
           f'    printf("%zu\\n", sizeof({struct}));\n{body}    return 0;\n}}\n')
    with tempfile.TemporaryDirectory() as td:
        c, exe = Path(td) / "lay.c", Path(td) / "lay"
        c.write_text(src)
        subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", "-o", str(exe), str(c)], check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return [int(v) for v in out]


@pytest.mark.parametrize("struct", ["arc_gather_params", "arc_gather_stats", "arc_image_error_t", "arc_batch_view", "arc_area_light", "arc_avpls", "arc_bvh_param",
                                    "arc_bvh_data", "arc_avpl_node_data", "arc_bvh_node", "arc_material"])
def test_ctypes_layout_equals_c(struct):
    ct = getattr(ffi, struct)
    fields = [f[0] for f in ct._fields_]
    got = _c_layout(struct, fields)
    assert got[0] == ctypes.sizeof(ct)
    assert got[1:] == [getattr(ct, f).offset for f in fields]


def test_cpp_facade_compiles():
    """include/arc_cudagather.hpp is self-contained C++17."""
    import subprocess

    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "cpp" / "facade_test.cpp")], check=True)
