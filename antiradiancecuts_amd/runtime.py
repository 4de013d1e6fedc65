"""Python host objects over the C ABI (device memory held in torch tensors).

Mirrors the reference's host classes for this path: ``Bvh`` covers
``SimpleBvh`` / ``AvplBvh`` (bvh.h:53-161) and ``Gather`` covers
``CudaGather`` (cudaGather.h:25-70).  All compute runs in libarc.so; torch
only provides device buffers and the stream.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import ffi
from .ffi import check

NODE_DTYPE = np.dtype([("left", "<i4"), ("right", "<i4"), ("bbMin", "<f4", 3), ("bbMax", "<f4", 3),
                       ("visited", "<i4")])
assert NODE_DTYPE.itemsize == 36


def _torch():
    import torch  # local import: CPU-only tests never need it

    return torch


class Context:
    """arc_ctx bound to torch's current stream on ``cuda:device``."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the HIP path has no CPU fallback")
        self.lib = ffi.lib()
        self.device = device
        h = C.c_void_p()
        check(self.lib.arc_ctx_create(device, C.byref(h)), "arc_ctx_create")
        self.h = h
        if use_torch_stream:
            s = torch.cuda.current_stream(device).cuda_stream
            check(self.lib.arc_ctx_set_stream(self.h, C.c_void_p(s)), "arc_ctx_set_stream", self.h)

    def sync(self) -> None:
        check(self.lib.arc_ctx_synchronize(self.h), "arc_ctx_synchronize", self.h)

    def close(self) -> None:
        if self.h:
            self.lib.arc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def to_host(self, ptr: int, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if count:
            check(self.lib.arc_copy_to_host(self.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes),
                  "arc_copy_to_host", self.h)
        return out


def device_array(a: np.ndarray, device: int = 0):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to(f"cuda:{device}")


@dataclass
class DeviceAvpls:
    """AVPL SoA in device memory (AvplsGpuParam layout)."""

    tensors: dict
    struct: ffi.arc_avpls

    @classmethod
    def from_scene(cls, scene, device: int = 0) -> "DeviceAvpls":
        return cls.from_arrays(scene.avpl_pos, scene.avpl_nrm, scene.avpl_dir, scene.avpl_rad, scene.avpl_anti,
                               scene.avpl_mat, scene.avpl_bounce, device)

    @classmethod
    def from_arrays(cls, pos, nrm, dirs, rad, anti, mat, bounce, device: int = 0) -> "DeviceAvpls":
        t = {k: device_array(v, device) for k, v in
             dict(pos=pos, nrm=nrm, dir=dirs, rad=rad, anti=anti, mat=mat.astype(np.int32),
                  bounce=bounce.astype(np.int32)).items()}
        s = ffi.arc_avpls()
        s.position, s.normal, s.incDirection = t["pos"].data_ptr(), t["nrm"].data_ptr(), t["dir"].data_ptr()
        s.incRadiance, s.antiradiance = t["rad"].data_ptr(), t["anti"].data_ptr()
        s.materialIndex, s.bounce = t["mat"].data_ptr(), t["bounce"].data_ptr()
        s.numAvpls = int(pos.shape[0])
        return cls(t, s)


class Bvh:
    """Bvh / SimpleBvh / AvplBvh over the C ABI."""

    def __init__(self, ctx: Context, consider_normals: bool = False):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        check(self.lib.arc_bvh_create(ctx.h, int(consider_normals), C.byref(h)), "arc_bvh_create", ctx.h)
        self.h = h
        self.avpl = False

    def close(self) -> None:
        if self.h:
            self.lib.arc_bvh_destroy(self.h)
            self.h = None

    def build_points(self, pos, nrm=None) -> None:
        """SimpleBvh(positions, normals) on device tensors (N, 3) float32."""
        n = int(pos.shape[0])
        st = self.lib.arc_bvh_build_points(self.h, C.c_void_p(pos.data_ptr()),
                                           C.c_void_p(nrm.data_ptr()) if nrm is not None else None, n, 0)
        check(st, "arc_bvh_build_points", self.ctx.h)
        self.avpl = False

    def build_avpls(self, avpls: DeviceAvpls, rand=None, seed: int = 5489, trust: bool = False) -> None:
        st = self.lib.arc_bvh_build_avpls(self.h, C.byref(avpls.struct),
                                          C.c_void_p(rand.data_ptr()) if rand is not None else None, seed,
                                          ffi.ARC_TRUST_INPUT if trust else 0)
        check(st, "arc_bvh_build_avpls", self.ctx.h)
        self.avpl = True

    def to_host(self) -> dict:
        """Copies every output array to numpy (synchronous)."""
        p, d, nd = ffi.arc_bvh_param(), ffi.arc_bvh_data(), ffi.arc_avpl_node_data()
        check(self.lib.arc_bvh_get_param(self.h, C.byref(p)), "arc_bvh_get_param", self.ctx.h)
        check(self.lib.arc_bvh_get_data(self.h, C.byref(d)), "arc_bvh_get_data", self.ctx.h)
        check(self.lib.arc_bvh_get_node_data(self.h, C.byref(nd)), "arc_bvh_get_node_data", self.ctx.h)
        n = p.numLeafs
        h = self.ctx.to_host
        out = dict(
            n=n,
            codes=h(d.morton, np.uint32, n),
            perm=h(d.perm, np.int32, n),
            nodes=h(p.nodes, NODE_DTYPE, n - 1),
            parents=h(d.parents, np.int32, 2 * n - 1),
            ranges=h(d.ranges, np.int32, 2 * (n - 1)).reshape(-1, 2),
            positions=h(p.positions, np.float32, 3 * n).reshape(n, 3),
            normals=h(p.normals, np.float32, 3 * n).reshape(n, 3),
        )
        if nd.size:
            s = 2 * n - 1
            out.update(
                nd_size=h(nd.size, np.int32, s), nd_mat=h(nd.materialIndex, np.int32, s),
                nd_rand=h(nd.randomNumbers, np.float32, s),
                nd_pos=h(nd.position, np.float32, 3 * s).reshape(s, 3),
                nd_nrm=h(nd.normal, np.float32, 3 * s).reshape(s, 3),
                nd_rad=h(nd.incRadiance, np.float32, 3 * s).reshape(s, 3),
                nd_dir=h(nd.incDirection, np.float32, 3 * s).reshape(s, 3),
            )
        return out

    def check_integrity(self) -> int:
        ms = C.c_int(0)
        check(self.lib.arc_bvh_check_integrity(self.h, C.byref(ms)), "arc_bvh_check_integrity", self.ctx.h)
        return ms.value


def default_params(**kw) -> ffi.arc_gather_params:
    p = ffi.arc_gather_params()
    ffi.lib().arc_gather_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def materials_struct(mats: np.ndarray):
    arr = (ffi.arc_material * len(mats))()
    for i, m in enumerate(np.asarray(mats, np.float32)):
        arr[i].emissive = ffi.arc_float3(*m[0:3])
        arr[i].diffuse = ffi.arc_float3(*m[3:6])
        arr[i].specular = ffi.arc_float3(*m[6:9])
        arr[i].exponent = float(m[9])
    return arr


class Gather:
    """CudaGather over the C ABI, with the G-buffer and targets as device tensors."""

    def __init__(self, ctx: Context, width: int, height: int, materials: np.ndarray):
        torch = _torch()
        self.ctx, self.lib = ctx, ctx.lib
        self.width, self.height = width, height
        mats = materials_struct(materials)
        h = C.c_void_p()
        check(self.lib.arc_gather_create(ctx.h, width, height, mats, len(materials), C.byref(h)),
              "arc_gather_create", ctx.h)
        self.h = h
        dev = f"cuda:{ctx.device}"
        npix = width * height
        self.result = torch.zeros((npix, 4), dtype=torch.float32, device=dev)
        self.radiance = torch.zeros((npix, 4), dtype=torch.float32, device=dev)
        self.antiradiance = torch.zeros((npix, 4), dtype=torch.float32, device=dev)
        check(self.lib.arc_gather_set_outputs(self.h, C.c_void_p(self.result.data_ptr()),
                                              C.c_void_p(self.radiance.data_ptr()),
                                              C.c_void_p(self.antiradiance.data_ptr())),
              "arc_gather_set_outputs", ctx.h)
        self._gb = None

    def close(self) -> None:
        if self.h:
            self.lib.arc_gather_destroy(self.h)
            self.h = None

    def set_gbuffer(self, pos4, nrm4, albedo4=None) -> None:
        self._gb = (pos4, nrm4, albedo4)
        check(self.lib.arc_gather_set_gbuffer(self.h, C.c_void_p(pos4.data_ptr()), C.c_void_p(nrm4.data_ptr()),
                                              C.c_void_p(albedo4.data_ptr()) if albedo4 is not None else None),
              "arc_gather_set_gbuffer", self.ctx.h)

    def run(self, avpls: DeviceAvpls, cam, extent: float, params=None, flags: int = 0) -> ffi.arc_gather_stats:
        params = params or default_params()
        stats = ffi.arc_gather_stats()
        st = self.lib.arc_gather_run(self.h, C.byref(avpls.struct), ffi.arc_float3(*map(float, cam)),
                                     float(extent), C.byref(params), flags, C.byref(stats))
        check(st, "arc_gather_run", self.ctx.h)
        return stats

    def exact_f64(self, avpls: DeviceAvpls, cam, extent: float, params=None):
        torch = _torch()
        params = params or default_params()
        dev = f"cuda:{self.ctx.device}"
        npix = self.width * self.height
        rad = torch.zeros((npix, 3), dtype=torch.float64, device=dev)
        anti = torch.zeros((npix, 3), dtype=torch.float64, device=dev)
        st = self.lib.arc_gather_exact_f64(self.h, C.byref(avpls.struct), ffi.arc_float3(*map(float, cam)),
                                           float(extent), C.byref(params), 0, C.c_void_p(rad.data_ptr()),
                                           C.c_void_p(anti.data_ptr()))
        check(st, "arc_gather_exact_f64", self.ctx.h)
        return rad, anti

    def set_cutnet(self, net: "CutNet | None") -> None:
        check(self.lib.arc_gather_set_cutnet(self.h, net.h if net is not None else None), "arc_gather_set_cutnet",
              self.ctx.h)
        self._net = net

    def set_estimate_pool(self, chunks_per_tile: int) -> None:
        """The moment cut's pool of kept nodes for the matrix-core pass, in
        chunks of 128 per tile (0: every node estimated in the traversal)."""
        check(self.lib.arc_gather_set_estimate_pool(self.h, int(chunks_per_tile)), "arc_gather_set_estimate_pool",
              self.ctx.h)

    def set_launch_shape(self, cut_stack: int = 96, anti_chunk: int = 1536) -> None:
        """The moment cut's traversal stack (64, 96, 128 nodes) and the
        clustered antiradiance's list entries filed per pass (768, 1536,
        3072); neither changes the image beyond rounding."""
        check(self.lib.arc_gather_set_launch_shape(self.h, int(cut_stack), int(anti_chunk)),
              "arc_gather_set_launch_shape", self.ctx.h)

    def cut_samples(self, avpls: DeviceAvpls, cam, extent: float, params, capacity: int, seed: int = 1,
                    keep: float = 1.0, flags: int = 0):
        """Training rows of this frame (arc_gather_cut_samples): (X, y, rows met)."""
        torch = _torch()
        dev = f"cuda:{self.ctx.device}"
        X = torch.zeros((max(capacity, 1), ffi.ARC_CUTNET_FEATURES), dtype=torch.float32, device=dev)
        y = torch.zeros(max(capacity, 1), dtype=torch.float32, device=dev)
        cnt = C.c_int(0)
        check(self.lib.arc_gather_cut_samples(self.h, C.byref(avpls.struct), ffi.arc_float3(*map(float, cam)),
                                              float(extent), C.byref(params), flags, C.c_void_p(X.data_ptr()),
                                              C.c_void_p(y.data_ptr()), int(capacity), int(seed), float(keep),
                                              C.byref(cnt)), "arc_gather_cut_samples", self.ctx.h)
        m = min(cnt.value, capacity)
        return X[:m], y[:m], cnt.value

    def cut_records(self, which: int) -> np.ndarray:
        """The last cut frame's node records (arc_gather_copy_cut_records):
        (2n - 1, 4, 4) centroid records or (n - 1, 13, 4) moment records."""
        cnt = C.c_longlong(0)
        check(self.lib.arc_gather_copy_cut_records(self.h, which, None, 0, C.byref(cnt)),
              "arc_gather_copy_cut_records", self.ctx.h)
        out = np.zeros((cnt.value, 4), np.float32)
        check(self.lib.arc_gather_copy_cut_records(self.h, which, out.ctypes.data, cnt.value, C.byref(cnt)),
              "arc_gather_copy_cut_records", self.ctx.h)
        per = 4 if which == ffi.ARC_RECORDS_CENTROID else 13
        return out.reshape(-1, per, 4)

    def bvh_to_host(self) -> dict:
        h = self.lib.arc_gather_get_bvh(self.h)
        if not h:
            raise RuntimeError("no cut frame has run")
        b = Bvh.__new__(Bvh)
        b.ctx, b.lib, b.h, b.avpl = self.ctx, self.lib, C.c_void_p(h), True
        return b.to_host()


def photon_radius(params, extent: float) -> float:
    return float(np.float32(params.photonRadiusScale) * np.float32(extent) / np.float32(100.0))


CONE_10DEG = math.pi / 18.0


class TraceScene:
    """arc_trace_scene: triangle BVH, AVPL shooter and G-buffer ray caster."""

    def __init__(self, ctx: Context, ts):
        self.ctx, self.lib, self.ts = ctx, ctx.lib, ts
        mats = materials_struct(ts.materials)
        self._mats = mats
        h = C.c_void_p()
        st = self.lib.arc_trace_scene_create(ctx.h, ts.p0.ctypes.data, ts.p1.ctypes.data, ts.p2.ctypes.data,
                                             ts.material.ctypes.data, int(len(ts.material)), mats,
                                             int(len(ts.materials)), ffi.ARC_INPUT_HOST, C.byref(h))
        check(st, "arc_trace_scene_create", ctx.h)
        self.h = h
        L = ts.light
        self.light = ffi.arc_area_light(ffi.arc_float3(*map(float, L["center"])), ffi.arc_float3(*map(float, L["front"])),
                                        ffi.arc_float3(*map(float, L["up"])), L["width"], L["height"],
                                        ffi.arc_float3(*map(float, L["radiance"])), int(L["material"]))

    def close(self) -> None:
        if self.h:
            self.lib.arc_trace_scene_destroy(self.h)
            self.h = None

    def info(self):
        n, d = C.c_int(0), C.c_int(0)
        check(self.lib.arc_trace_scene_info(self.h, C.byref(n), C.byref(d)), "arc_trace_scene_info", self.ctx.h)
        return n.value, d.value

    def intersect(self, org, dirs):
        torch = _torch()
        o, d = device_array(org, self.ctx.device), device_array(dirs, self.ctx.device)
        n = int(org.shape[0])
        t = torch.empty(n, dtype=torch.float32, device=o.device)
        tri = torch.empty(n, dtype=torch.int32, device=o.device)
        check(self.lib.arc_trace_intersect(self.h, C.c_void_p(o.data_ptr()), C.c_void_p(d.data_ptr()), n,
                                           C.c_void_p(t.data_ptr()), C.c_void_p(tri.data_ptr())),
              "arc_trace_intersect", self.ctx.h)
        return t, tri

    def shoot(self, num_paths: int, seed: int, flags: int = 0) -> DeviceAvpls:
        """AVPLs of num_paths paths; the arrays are borrowed from the scene
        object (valid until the next shoot)."""
        s = ffi.arc_avpls()
        n = C.c_int(0)
        check(self.lib.arc_trace_shoot(self.h, C.byref(self.light), int(num_paths), int(seed), int(flags), C.byref(s),
                                       C.byref(n)), "arc_trace_shoot", self.ctx.h)
        return DeviceAvpls({}, s)

    def gbuffer(self, cam, forward, up, fovy: float, width: int, height: int):
        torch = _torch()
        dev = f"cuda:{self.ctx.device}"
        pos4 = torch.empty((width * height, 4), dtype=torch.float32, device=dev)
        nrm4 = torch.empty((width * height, 4), dtype=torch.float32, device=dev)
        check(self.lib.arc_trace_gbuffer(self.h, ffi.arc_float3(*map(float, cam)), ffi.arc_float3(*map(float, forward)),
                                         ffi.arc_float3(*map(float, up)), float(fovy), int(width), int(height),
                                         C.c_void_p(pos4.data_ptr()), C.c_void_p(nrm4.data_ptr())),
              "arc_trace_gbuffer", self.ctx.h)
        return pos4, nrm4


def avpls_to_host(avpls: DeviceAvpls, ctx: Context) -> dict:
    """Copies an AVPL SoA (device pointers) to numpy arrays."""
    s = avpls.struct
    n = s.numAvpls
    out = {}
    for name, ptr, cols, dt in (("pos", s.position, 3, np.float32), ("nrm", s.normal, 3, np.float32),
                                ("dir", s.incDirection, 3, np.float32), ("rad", s.incRadiance, 3, np.float32),
                                ("anti", s.antiradiance, 3, np.float32), ("mat", s.materialIndex, 1, np.int32),
                                ("bounce", s.bounce, 1, np.int32)):
        a = ctx.to_host(ptr or 0, dt, n * cols)
        out[name] = a.reshape(n, 3) if cols == 3 else a
    return out


class LaneContext:
    """A batch lane's arc_ctx (owned by its Batch); gathers of that lane are
    created on it."""

    def __init__(self, lib, h, device: int):
        self.lib, self.h, self.device = lib, h, device

    def sync(self) -> None:
        check(self.lib.arc_ctx_synchronize(self.h), "arc_ctx_synchronize", self.h)


class Batch:
    """arc_batch: views captured into one HIP graph, one launch per batch."""

    def __init__(self, ctx: Context, lanes: int = 4):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        check(self.lib.arc_batch_create(ctx.h, int(lanes), C.byref(h)), "arc_batch_create", ctx.h)
        self.h = h
        self.lanes = [LaneContext(self.lib, C.c_void_p(self.lib.arc_batch_lane(h, i)), ctx.device)
                      for i in range(lanes)]
        self._views = None

    def lane(self, i: int) -> LaneContext:
        return self.lanes[i % len(self.lanes)]

    def capture(self, views) -> None:
        """views: (Gather, DeviceAvpls, cam, extent, params) per view."""
        arr = (ffi.arc_batch_view * len(views))()
        for k, (g, av, cam, extent, params) in enumerate(views):
            arr[k].gather = g.h.value if isinstance(g.h, C.c_void_p) else g.h
            arr[k].avpls = av.struct
            arr[k].camPos = ffi.arc_float3(*map(float, cam))
            arr[k].sceneExtent = float(extent)
            arr[k].params = params
        self._views = (arr, views)  # keep the inputs alive
        check(self.lib.arc_batch_capture(self.h, arr, len(views)), "arc_batch_capture", self.ctx.h)

    def set_one_cut_launch(self, on: bool) -> None:
        """All views' moment cuts in one launch, or one per view (the
        default); read by the next capture."""
        check(self.lib.arc_batch_set_one_cut_launch(self.h, int(bool(on))), "arc_batch_set_one_cut_launch",
              self.ctx.h)

    def launch(self) -> None:
        check(self.lib.arc_batch_launch(self.h), "arc_batch_launch", self.ctx.h)

    def close(self) -> None:
        if self.h:
            self.lib.arc_batch_destroy(self.h)
            self.h = None


class CutNet:
    """arc_cutnet: the learned cut's refine / keep predictor."""

    def __init__(self, ctx: Context):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        check(self.lib.arc_cutnet_create(ctx.h, C.byref(h)), "arc_cutnet_create", ctx.h)
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.arc_cutnet_destroy(self.h)
            self.h = None

    def init(self, seed: int, mean, std) -> None:
        m = np.ascontiguousarray(mean, np.float32)
        sd = np.ascontiguousarray(std, np.float32)
        assert m.shape == (ffi.ARC_CUTNET_FEATURES,) and sd.shape == m.shape
        check(self.lib.arc_cutnet_init(self.h, int(seed), m.ctypes.data_as(C.c_void_p),
                                       sd.ctypes.data_as(C.c_void_p)), "arc_cutnet_init", self.ctx.h)

    def set_weights(self, w) -> None:
        w = np.ascontiguousarray(w, np.float32)
        check(self.lib.arc_cutnet_set_weights(self.h, w.ctypes.data_as(C.c_void_p), int(w.size)),
              "arc_cutnet_set_weights", self.ctx.h)

    def get_weights(self) -> np.ndarray:
        w = np.zeros(ffi.ARC_CUTNET_PARAMS, np.float32)
        check(self.lib.arc_cutnet_get_weights(self.h, w.ctypes.data_as(C.c_void_p), int(w.size)),
              "arc_cutnet_get_weights", self.ctx.h)
        return w

    def forward(self, X):
        """Probabilities of the rows of a (n, 24) float32 device tensor."""
        torch = _torch()
        n = int(X.shape[0])
        p = torch.zeros(max(n, 1), dtype=torch.float32, device=X.device)
        check(self.lib.arc_cutnet_forward(self.h, C.c_void_p(X.data_ptr()), n, C.c_void_p(p.data_ptr())),
              "arc_cutnet_forward", self.ctx.h)
        return p[:n]

    def train(self, X, y, w=None, epochs: int = 10, batch: int = 4096, lr: float = 1e-3, seed: int = 1):
        loss = np.zeros(max(epochs, 1), np.float32)
        check(self.lib.arc_cutnet_train(self.h, C.c_void_p(X.data_ptr()), C.c_void_p(y.data_ptr()),
                                        C.c_void_p(w.data_ptr()) if w is not None else None, int(X.shape[0]),
                                        int(epochs), int(batch), float(lr), int(seed),
                                        loss.ctypes.data_as(C.c_void_p)), "arc_cutnet_train", self.ctx.h)
        return loss[:epochs]
