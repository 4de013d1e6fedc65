"""ctypes bindings of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the timed CPU baseline.  ``ref_morton()``
loads oracle/_ref/libref_morton.so (the reference's own morton.h compiled in
place) when it was built.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_morton.so"

NODE_DTYPE = np.dtype([("left", "<i4"), ("right", "<i4"), ("bbMin", "<f4", 3), ("bbMax", "<f4", 3),
                       ("visited", "<i4")])

vp = C.c_void_p
_lib = None
_ref = None


class orc_scene(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("pos4", vp), ("nrm4", vp), ("albedo4", vp),
                ("mats", vp), ("num_mats", C.c_int), ("num_avpls", C.c_int), ("apos3", vp), ("anrm3", vp),
                ("adir3", vp), ("arad3", vp), ("aanti3", vp), ("amat", vp), ("cam", C.c_float * 3),
                ("photon_radius", C.c_float), ("antirad_model", C.c_int), ("cone_half_angle", C.c_float)]


def build() -> None:
    """make -C oracle (liboracle.so, and _ref/ when the reference is present)."""
    proc = subprocess.run(["make", "-s", "-C", str(HERE)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{proc.stdout}")


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        u32, u64, i32, f32 = C.c_uint32, C.c_uint64, C.c_int, C.c_float
        sig = {
            "orc_expand_bits_by1": (u32, [u32]),
            "orc_expand_bits_by2": (u32, [u32]),
            "orc_interleave3": (u32, [u32, u32, u32]),
            "orc_interleave6": (u64, [u64] * 6),
            "orc_morton3d": (u32, [f32] * 3),
            "orc_morton6d": (u64, [f32] * 6),
            "orc_bounds": (None, [vp, i32, vp, vp]),
            "orc_normalize": (None, [vp, i32, vp, vp, vp]),
            "orc_delta": (i32, [i32, i32, vp, i32]),
            "orc_radix_tree": (i32, [vp, i32, vp, vp, vp]),
            "orc_refit": (i32, [vp, vp, vp, i32]),
            "orc_check_tree_integrity": (i32, [vp, vp, i32, C.POINTER(i32)]),
            "orc_lbvh_build": (i32, [vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp]),
            "orc_avpl_bvh_build": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                         vp, vp, vp, vp, vp, vp]),
            "orc_pair": (i32, [C.POINTER(orc_scene), i32, i32, i32, vp, vp]),
            "orc_gather_exact": (i32, [C.POINTER(orc_scene), i32, vp, i32, vp, vp]),
            "orc_image_accumulate": (None, [vp, vp, i32]),
            "orc_tri_intersect": (i32, [vp, vp, vp, vp, vp, f32, vp]),
            "orc_trace_closest": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, f32, vp]),
            "orc_trace_shoot": (i32, [i32, vp, vp, vp, vp, vp, i32, vp, u64, i32, i32, i32, vp, vp, vp, vp, vp,
                                      vp, vp, vp, i32]),
            "orc_image_normalize": (None, [vp, vp, i32, i32]),
            "orc_image_error": (f32, [vp, vp, i32, vp]),
            "orc_set_threads": (i32, [i32]),
            "orc_gather_exact_parts": (i32, [C.POINTER(orc_scene), i32, vp, i32, vp, vp, i32]),
            "orc_gather_cut_solid_angle": (i32, [C.POINTER(orc_scene), i32, i32, vp, vp, vp, vp, vp, vp, f32,
                                                 vp, i32, vp, vp, C.POINTER(C.c_longlong),
                                                 C.POINTER(C.c_longlong)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def ref_morton():
    """The reference's morton.h compiled in place, or None where unavailable."""
    global _ref
    if _ref is None and REF_LIB.exists():
        L = C.CDLL(str(REF_LIB))
        u32, u64, f32 = C.c_uint32, C.c_uint64, C.c_float
        for name, res, args in [("ref_expand_bits_by1", u32, [u32]), ("ref_expand_bits_by2", u32, [u32]),
                                ("ref_interleave3", u32, [u32] * 3), ("ref_interleave6", u64, [u64] * 6),
                                ("ref_morton3d", u32, [f32] * 3), ("ref_morton6d", u64, [f32] * 6),
                                ("ref_morton3d_batch", None, [vp, C.c_int, vp])]:
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _ref = L
    return _ref


def _p(a):
    return None if a is None else np.ascontiguousarray(a).ctypes.data_as(vp)


def bounds(p3: np.ndarray):
    p3 = np.ascontiguousarray(p3, np.float32)
    mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().orc_bounds(_p(p3), len(p3), _p(mn), _p(mx))
    return mn, mx


def lbvh_build(pos: np.ndarray, nrm: np.ndarray | None = None, consider_normals: bool = False) -> dict:
    pos = np.ascontiguousarray(pos, np.float32)
    n = len(pos)
    nrm_c = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
    out = dict(codes=np.zeros(n, np.uint32), perm=np.zeros(n, np.int32),
               nodes=np.zeros(max(n - 1, 0), NODE_DTYPE), parents=np.zeros(max(2 * n - 1, 0), np.int32),
               ranges=np.zeros((max(n - 1, 0), 2), np.int32), positions=np.zeros((n, 3), np.float32),
               normals=np.zeros((n, 3), np.float32))
    st = lib().orc_lbvh_build(_p(pos), _p(nrm_c), n, int(consider_normals), _p(out["codes"]), _p(out["perm"]),
                              _p(out["nodes"]), _p(out["parents"]), _p(out["ranges"]), _p(out["positions"]),
                              _p(out["normals"]) if nrm_c is not None else None)
    out["status"] = st
    return out


def avpl_bvh_build(scene_like, rand: np.ndarray) -> dict:
    s = scene_like
    n = len(s.avpl_pos)
    slots = 2 * n - 1
    out = dict(codes=np.zeros(n, np.uint32), perm=np.zeros(n, np.int32), nodes=np.zeros(n - 1, NODE_DTYPE),
               parents=np.zeros(slots, np.int32), ranges=np.zeros((n - 1, 2), np.int32),
               positions=np.zeros((n, 3), np.float32), normals=np.zeros((n, 3), np.float32),
               nd_size=np.zeros(slots, np.int32), nd_mat=np.zeros(slots, np.int32),
               nd_pos=np.zeros((slots, 3), np.float32), nd_nrm=np.zeros((slots, 3), np.float32),
               nd_rad=np.zeros((slots, 3), np.float32), nd_dir=np.zeros((slots, 3), np.float32))
    rand = np.ascontiguousarray(rand, np.float32)
    st = lib().orc_avpl_bvh_build(n, _p(s.avpl_pos), _p(s.avpl_nrm), _p(s.avpl_dir), _p(s.avpl_rad),
                                  _p(s.avpl_mat.astype(np.int32)), _p(s.avpl_bounce.astype(np.int32)), _p(rand),
                                  _p(out["codes"]), _p(out["perm"]), _p(out["nodes"]), _p(out["parents"]),
                                  _p(out["ranges"]), _p(out["positions"]), _p(out["normals"]),
                                  _p(out["nd_size"]), _p(out["nd_mat"]), _p(out["nd_pos"]), _p(out["nd_nrm"]),
                                  _p(out["nd_rad"]), _p(out["nd_dir"]))
    out["status"] = st
    return out


class SceneRef:
    """Keeps the numpy arrays alive while the oracle reads them."""

    def __init__(self, scene, photon_radius: float, model: int = 0, cone_half_angle: float = 0.0,
                 avpl_subset=None):
        self.keep = []

        def keep(a, dt):
            a = np.ascontiguousarray(a, dt)
            self.keep.append(a)
            return a.ctypes.data_as(vp)

        sel = slice(None) if avpl_subset is None else avpl_subset
        s = orc_scene()
        s.width, s.height = scene.width, scene.height
        s.pos4, s.nrm4 = keep(scene.pos4, np.float32), keep(scene.nrm4, np.float32)
        s.albedo4 = keep(scene.albedo4, np.float32) if scene.albedo4 is not None else None
        s.mats = keep(scene.materials, np.float32)
        s.num_mats = len(scene.materials)
        ap = np.asarray(scene.avpl_pos)[sel]
        s.num_avpls = len(ap)
        s.apos3 = keep(ap, np.float32)
        s.anrm3 = keep(np.asarray(scene.avpl_nrm)[sel], np.float32)
        s.adir3 = keep(np.asarray(scene.avpl_dir)[sel], np.float32)
        s.arad3 = keep(np.asarray(scene.avpl_rad)[sel], np.float32)
        s.aanti3 = keep(np.asarray(scene.avpl_anti)[sel], np.float32)
        s.amat = keep(np.asarray(scene.avpl_mat)[sel], np.int32)
        s.cam[0], s.cam[1], s.cam[2] = (float(v) for v in scene.cam)
        s.photon_radius = photon_radius
        s.antirad_model = model
        s.cone_half_angle = cone_half_angle
        self.s = s


def gather_exact(ref: SceneRef, precision: int = 64, pixels: np.ndarray | None = None, what: int = 3):
    npix = ref.s.width * ref.s.height if pixels is None else len(pixels)
    rad = np.zeros((npix, 3), np.float64)
    anti = np.zeros((npix, 3), np.float64)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    st = lib().orc_gather_exact_parts(C.byref(ref.s), precision, _p(px), npix, _p(rad), _p(anti), what)
    assert st == 0
    return rad, anti


def pair(ref: SceneRef, pixel: int, avpl: int, precision: int = 64):
    rad = np.zeros(3, np.float64)
    anti = np.zeros(3, np.float64)
    lib().orc_pair(C.byref(ref.s), precision, pixel, avpl, _p(rad), _p(anti))
    return rad, anti


def gather_cut_solid_angle(ref: SceneRef, tree: dict, ref_thresh: float, precision: int = 32,
                           pixels: np.ndarray | None = None):
    npix = ref.s.width * ref.s.height if pixels is None else len(pixels)
    rad = np.zeros((npix, 3), np.float64)
    err = np.zeros(npix, np.int32)
    vis, ev = C.c_longlong(0), C.c_longlong(0)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    st = lib().orc_gather_cut_solid_angle(
        C.byref(ref.s), precision, int(len(tree["codes"])), _p(tree["nodes"]), _p(tree["nd_mat"]),
        _p(tree["nd_pos"]), _p(tree["nd_nrm"]), _p(tree["nd_rad"]), _p(tree["nd_dir"]), ref_thresh, _p(px), npix,
        _p(rad), _p(err), C.byref(vis), C.byref(ev))
    assert st == 0
    return rad, err, vis.value, ev.value
