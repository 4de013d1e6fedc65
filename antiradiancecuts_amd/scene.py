"""Synthetic workloads (host numpy arrays) from libarc_scene.so.

Named workloads follow the sizes the gather and LBVH are measured on:

* ``cornell_256``  — 256x256 G-buffer of the unit Cornell box with its two
  rotated boxes, 4096 AVPLs (bounce 1-3, L ~ U[0,1]^3, 30 % antiradiance),
  seed 0: the exact-gather case.
* ``points_1m``    — 1,000,000 LBVH points: half uniform in [0,1)^3, half
  from 32 Gaussians (sigma 0.01), seed 1.
* ``clutter_1080p`` — 1920x1080 G-buffer of the box walls plus 200,000 random
  triangles (area U[1e-4, 1e-2]), 1,000,000 AVPLs over bounces 1-4 with
  per-channel radiance log-uniform in [1e-3, 10] and 30 % antiradiance,
  seed 2: the cut-gather frame.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import ffi


@dataclass
class Scene:
    width: int
    height: int
    pos4: np.ndarray      # (H*W, 4) float32
    nrm4: np.ndarray      # (H*W, 4) float32, w = material index
    albedo4: np.ndarray   # (H*W, 4) float32
    materials: np.ndarray  # (M, 10) float32: emissive, diffuse, specular, exponent
    cam: np.ndarray       # (3,) float32
    extent: float
    avpl_pos: np.ndarray  # (N, 3) float32
    avpl_nrm: np.ndarray
    avpl_dir: np.ndarray
    avpl_rad: np.ndarray
    avpl_anti: np.ndarray
    avpl_mat: np.ndarray  # (N,) int32
    avpl_bounce: np.ndarray
    misses: int = 0

    @property
    def num_avpls(self) -> int:
        return int(self.avpl_pos.shape[0])

    @property
    def npix(self) -> int:
        return self.width * self.height


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def desc(seed: int, width: int, height: int, num_avpls: int, min_bounce: int = 1, max_bounce: int = 3,
         radiance_lo: float = 0.0, radiance_hi: float = 1.0, log_uniform: bool = False,
         antiradiance_prob: float = 0.3, with_boxes: bool = True, num_random_tris: int = 0,
         tri_area_lo: float = 1e-4, tri_area_hi: float = 1e-2) -> ffi.arc_scene_desc:
    d = ffi.arc_scene_desc()
    d.seed = seed
    d.width, d.height, d.num_avpls = width, height, num_avpls
    d.min_bounce, d.max_bounce = min_bounce, max_bounce
    d.radiance_lo, d.radiance_hi = radiance_lo, radiance_hi
    d.radiance_log_uniform = int(log_uniform)
    d.antiradiance_prob = antiradiance_prob
    d.with_boxes = int(with_boxes)
    d.num_random_tris = num_random_tris
    d.tri_area_lo, d.tri_area_hi = tri_area_lo, tri_area_hi
    return d


def make_scene(d: ffi.arc_scene_desc, gbuffer: bool = True) -> Scene:
    lib = ffi.scene_lib()
    W, H, N = d.width, d.height, d.num_avpls
    pos4 = np.zeros((W * H, 4), np.float32)
    nrm4 = np.zeros((W * H, 4), np.float32)
    alb4 = np.zeros((W * H, 4), np.float32)
    misses = 0
    if gbuffer:
        misses = lib.arc_scene_gbuffer(C.byref(d), _ptr(pos4), _ptr(nrm4), _ptr(alb4))
        if misses < 0:
            raise ValueError("arc_scene_gbuffer rejected the description")
    nm = lib.arc_scene_num_materials()
    mats = np.zeros((nm, 10), np.float32)
    lib.arc_scene_materials(_ptr(mats), nm)
    cam = np.zeros(3, np.float32)
    ext = np.zeros(1, np.float32)
    lib.arc_scene_camera(C.byref(d), _ptr(cam), _ptr(ext))
    f3 = lambda: np.zeros((N, 3), np.float32)  # noqa: E731
    ap, an, ad, ar, aa = f3(), f3(), f3(), f3(), f3()
    am = np.zeros(N, np.int32)
    ab = np.zeros(N, np.int32)
    if lib.arc_scene_avpls(C.byref(d), _ptr(ap), _ptr(an), _ptr(ad), _ptr(ar), _ptr(aa), _ptr(am), _ptr(ab)):
        raise ValueError("arc_scene_avpls rejected the description")
    return Scene(W, H, pos4, nrm4, alb4, mats, cam, float(ext[0]), ap, an, ad, ar, aa, am, ab, misses)


def cornell_256(seed: int = 0, width: int = 256, height: int = 256, num_avpls: int = 4096) -> Scene:
    return make_scene(desc(seed, width, height, num_avpls, 1, 3, 0.0, 1.0, False, 0.3, True, 0))


def clutter_1080p(seed: int = 2, width: int = 1920, height: int = 1080, num_avpls: int = 1_000_000,
                  num_tris: int = 200_000) -> Scene:
    return make_scene(desc(seed, width, height, num_avpls, 1, 4, 1e-3, 10.0, True, 0.3, False, num_tris,
                           1e-4, 1e-2))


def points(n: int = 1_000_000, seed: int = 1, num_gaussians: int = 32, sigma: float = 0.01) -> np.ndarray:
    out = np.zeros((n, 3), np.float32)
    if ffi.scene_lib().arc_scene_points(seed, n, num_gaussians, sigma, _ptr(out)):
        raise ValueError("arc_scene_points rejected the arguments")
    return out


def unit_normals(n: int, seed: int = 7) -> np.ndarray:
    out = np.zeros((n, 3), np.float32)
    if ffi.scene_lib().arc_scene_normals(seed, n, _ptr(out)):
        raise ValueError("arc_scene_normals rejected the arguments")
    return out


# Render.cpp:64-81: the eight Cornell-ceiling points of the reference's
# SimpleBvh smoke build (double literals narrowed to float by make_float3).
SMOKE_POINTS = np.array([
    [536.419, 548.9, 431.274], [540.488, 548.9, 465.94], [419.977, 469.29, 559.3],
    [555.476, 471.466, 510.595], [438.946, 548.9, 506.097], [434.071, 501.603, 559.3],
    [452.146, 548.9, 529.52], [504.6, 509.978, 559.3]], dtype=np.float64).astype(np.float32)
SMOKE_NORMALS = np.tile(np.array([0.0, 1.0, 0.0], np.float32), (8, 1))


@dataclass
class TriangleScene:
    """The scene's triangles for the shooter and the G-buffer ray caster
    (Triangle convention: front-face normal normalize(cross(p2 - p0, p1 - p0))),
    the ceiling area light, and the material table with the light's emissive
    material appended."""
    p0: np.ndarray
    p1: np.ndarray
    p2: np.ndarray
    material: np.ndarray
    materials: np.ndarray      # (num_materials + 1, 10); the last one is the light's
    light: dict                # center, front, up, width, height, radiance, material
    light_axes: tuple          # unit side axes (s, u) of the light quad


def triangle_scene(d: ffi.arc_scene_desc, light_quad: bool = True) -> TriangleScene:
    lib = ffi.scene_lib()
    n = lib.arc_scene_num_triangles(C.byref(d))
    if n < 0:
        raise ValueError("arc_scene_num_triangles rejected the description")
    p0, p1, p2 = (np.zeros((n, 3), np.float32) for _ in range(3))
    mat = np.zeros(n, np.int32)
    got = lib.arc_scene_triangles(C.byref(d), _ptr(p0), _ptr(p1), _ptr(p2), _ptr(mat), n)
    assert got == n
    nm = lib.arc_scene_num_materials()
    mats = np.zeros((nm + 1, 10), np.float32)
    lib.arc_scene_materials(_ptr(mats), nm)
    c, f, u, size, rad = (np.zeros(k, np.float32) for k in (3, 3, 3, 2, 3))
    lmat = np.zeros(10, np.float32)
    lib.arc_scene_light(C.byref(d), _ptr(c), _ptr(f), _ptr(u), _ptr(size), _ptr(rad), _ptr(lmat))
    mats[nm] = lmat
    # the quad's axes as the shooter forms them (lookAt towards -front)
    fw = -f / np.linalg.norm(f)
    s = np.cross(fw, u)
    s = (s / np.linalg.norm(s)).astype(np.float32)
    uu = np.cross(s, fw).astype(np.float32)
    light = dict(center=c, front=f, up=u, width=float(size[0]), height=float(size[1]), radiance=rad, material=nm)
    if light_quad:  # two emissive triangles facing along front
        hw, hh = 0.5 * size[0], 0.5 * size[1]
        q = [c - hw * s - hh * uu, c + hw * s - hh * uu, c + hw * s + hh * uu, c - hw * s + hh * uu]
        tri = [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
        lp = []
        for a, b, e in tri:
            if np.dot(np.cross(e - a, b - a), f) < 0:  # orient the front face along front
                b, e = e, b
            lp.append((a, b, e))
        p0 = np.concatenate([p0, np.array([t[0] for t in lp], np.float32)])
        p1 = np.concatenate([p1, np.array([t[1] for t in lp], np.float32)])
        p2 = np.concatenate([p2, np.array([t[2] for t in lp], np.float32)])
        mat = np.concatenate([mat, np.full(2, nm, np.int32)])
    return TriangleScene(p0, p1, p2, mat, mats, light, (s, uu))
