"""ctypes bindings of include/arc.h and include/arc_scene.h.

The HIP library ``libarc.so`` is the product; this module only declares its
C ABI for Python callers (tests, bench).  Loading fails loudly when the
library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libarc.so"
SCENE_LIB_PATH = PKG / "libarc_scene.so"

ARC_OK = 0
ARC_E_INVALID = -1
ARC_E_TOO_FEW_POINTS = -2
ARC_E_HIP = -3
ARC_E_NOMEM = -4
ARC_E_NO_DEVICE = -5
ARC_E_UNSUPPORTED = -6
ARC_E_STATE = -7

ARC_BVH_CONSIDER_NORMALS = 1
ARC_INPUT_HOST = 256
ARC_TRUST_INPUT = 512
ARC_CUT_SOLID_ANGLE = 0
ARC_CUT_ERROR_BOUND = 1
ARC_CUT_LEARNED = 2
ARC_CUTNET_FEATURES = 24
ARC_CUTNET_PARAMS = 6400
ARC_CRITERION_BOUND = 0
ARC_SHOOT_NO_EXPLICIT_DIRECT = 1
ARC_CRITERION_ESTIMATE = 1
ARC_CRITERION_EXPANSION = 2
ARC_ESTIMATOR_CENTROID = 0
ARC_ESTIMATOR_MOMENTS = 1
ARC_RECORDS_CENTROID = 0
ARC_RECORDS_MOMENTS = 1
ARC_ANTIRAD_DISK = 0
ARC_ANTIRAD_CONE = 1
ARC_ANTI_EXACT = 0
ARC_ANTI_CLUSTERED = 1
ARC_RUN_PROFILE = 1
ARC_RUN_COUNTERS = 2
ARC_RUN_TIMELINE = 4

vp = C.c_void_p


class arc_float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class arc_material(C.Structure):
    _fields_ = [("emissive", arc_float3), ("diffuse", arc_float3), ("specular", arc_float3),
                ("exponent", C.c_float)]


class arc_bvh_node(C.Structure):
    _fields_ = [("left", C.c_int), ("right", C.c_int), ("bbMin", arc_float3), ("bbMax", arc_float3),
                ("visited", C.c_int)]


class arc_avpls(C.Structure):
    _fields_ = [("position", vp), ("normal", vp), ("incDirection", vp), ("incRadiance", vp),
                ("antiradiance", vp), ("materialIndex", vp), ("bounce", vp), ("numAvpls", C.c_int)]


class arc_bvh_param(C.Structure):
    _fields_ = [("numLeafs", C.c_int), ("numNodes", C.c_int), ("nodes", vp), ("positions", vp),
                ("normals", vp)]


class arc_bvh_data(C.Structure):
    _fields_ = [("morton", vp), ("parents", vp), ("numLeafs", C.c_int), ("numNodes", C.c_int),
                ("perm", vp), ("ranges", vp)]


class arc_avpl_node_data(C.Structure):
    _fields_ = [("size", vp), ("materialIndex", vp), ("randomNumbers", vp), ("position", vp),
                ("normal", vp), ("incRadiance", vp), ("incDirection", vp)]


class arc_gather_params(C.Structure):
    _fields_ = [("useLightCuts", C.c_int), ("cutMode", C.c_int),
                ("clusterRefinementThreshold", C.c_float), ("errorBound", C.c_float),
                ("antiradianceModel", C.c_int), ("antiradianceMode", C.c_int),
                ("photonRadiusScale", C.c_float), ("coneHalfAngle", C.c_float),
                ("writeSentinel", C.c_int), ("stackSize", C.c_int), ("seed", C.c_uint64),
                ("cutEstimateDepth", C.c_int), ("cutBudget", C.c_int), ("cutPerTile", C.c_int),
                ("cutCriterion", C.c_int), ("cutImageFloor", C.c_float), ("cutConfidence", C.c_float),
                ("cutEvalDepth", C.c_int), ("cutEstimator", C.c_int), ("cutExactCount", C.c_int),
                ("cutOpening", C.c_float)]


class arc_gather_stats(C.Structure):
    _fields_ = [("msUpload", C.c_float), ("msBvh", C.c_float), ("msRadiance", C.c_float),
                ("msAntiradiance", C.c_float), ("msCombine", C.c_float), ("msTotal", C.c_float),
                ("visitedNodes", C.c_longlong), ("evaluatedNodes", C.c_longlong),
                ("antiVisitedNodes", C.c_longlong), ("antiEvaluatedNodes", C.c_longlong),
                ("errorPixels", C.c_longlong), ("numLeafs", C.c_int), ("budgetHits", C.c_longlong),
                ("fallbackNodes", C.c_longlong), ("testedNodes", C.c_longlong),
                ("exactLeaves", C.c_longlong)]


class arc_scene_desc(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("width", C.c_int), ("height", C.c_int), ("num_avpls", C.c_int),
                ("min_bounce", C.c_int), ("max_bounce", C.c_int), ("radiance_lo", C.c_float),
                ("radiance_hi", C.c_float), ("radiance_log_uniform", C.c_int),
                ("antiradiance_prob", C.c_float), ("with_boxes", C.c_int), ("num_random_tris", C.c_int),
                ("tri_area_lo", C.c_float), ("tri_area_hi", C.c_float)]


class arc_image_error_t(C.Structure):
    _fields_ = [("referenceError", C.c_float), ("rmse", C.c_float * 3), ("relL2", C.c_double),
                ("maxAbs", C.c_double)]


class arc_area_light(C.Structure):
    _fields_ = [("center", arc_float3), ("front", arc_float3), ("up", arc_float3), ("width", C.c_float),
                ("height", C.c_float), ("radiance", arc_float3), ("materialIndex", C.c_int)]


class arc_batch_view(C.Structure):
    _fields_ = [("gather", C.c_void_p), ("avpls", arc_avpls), ("camPos", arc_float3), ("sceneExtent", C.c_float),
                ("params", arc_gather_params)]


P = C.POINTER
# (name, restype, argtypes) for every symbol of include/arc.h
ARC_SYMBOLS = [
    ("arc_version", C.c_char_p, []),
    ("arc_status_string", C.c_char_p, [C.c_int]),
    ("arc_ctx_create", C.c_int, [C.c_int, P(vp)]),
    ("arc_ctx_destroy", None, [vp]),
    ("arc_ctx_set_stream", C.c_int, [vp, vp]),
    ("arc_ctx_use_own_stream", C.c_int, [vp]),
    ("arc_ctx_get_stream", vp, [vp]),
    ("arc_ctx_synchronize", C.c_int, [vp]),
    ("arc_ctx_last_error", C.c_char_p, [vp]),
    ("arc_device_alloc", C.c_int, [vp, C.c_size_t, P(vp)]),
    ("arc_device_free", C.c_int, [vp, vp]),
    ("arc_copy_to_device", C.c_int, [vp, vp, vp, C.c_size_t]),
    ("arc_copy_to_host", C.c_int, [vp, vp, vp, C.c_size_t]),
    ("arc_bvh_create", C.c_int, [vp, C.c_int, P(vp)]),
    ("arc_bvh_destroy", None, [vp]),
    ("arc_bvh_build_points", C.c_int, [vp, vp, vp, C.c_int, C.c_int]),
    ("arc_bvh_build_avpls", C.c_int, [vp, P(arc_avpls), vp, C.c_uint64, C.c_int]),
    ("arc_bvh_random_numbers", C.c_int, [vp, C.c_uint64, C.c_int, vp]),
    ("arc_bvh_get_param", C.c_int, [vp, P(arc_bvh_param)]),
    ("arc_bvh_get_data", C.c_int, [vp, P(arc_bvh_data)]),
    ("arc_bvh_get_node_data", C.c_int, [vp, P(arc_avpl_node_data)]),
    ("arc_bvh_get_node", C.c_int, [vp, C.c_int, P(arc_bvh_node)]),
    ("arc_bvh_check_integrity", C.c_int, [vp, P(C.c_int)]),
    ("arc_gather_params_default", None, [P(arc_gather_params)]),
    ("arc_gather_create", C.c_int, [vp, C.c_int, C.c_int, P(arc_material), C.c_int, P(vp)]),
    ("arc_gather_destroy", None, [vp]),
    ("arc_gather_set_gbuffer", C.c_int, [vp, vp, vp, vp]),
    ("arc_gather_set_outputs", C.c_int, [vp, vp, vp, vp]),
    ("arc_gather_run", C.c_int, [vp, P(arc_avpls), arc_float3, C.c_float, P(arc_gather_params), C.c_int,
                                 P(arc_gather_stats)]),
    ("arc_gather_exact_f64", C.c_int, [vp, P(arc_avpls), arc_float3, C.c_float, P(arc_gather_params),
                                       C.c_int, vp, vp]),
    ("arc_gather_profile_history", C.c_int, [vp, C.c_int, vp]),
    ("arc_gather_kernel_history", C.c_int, [vp, C.c_int, vp]),
    ("arc_gather_kernel_spans", C.c_int, [vp, C.c_int, vp]),
    ("arc_gather_get_bvh", vp, [vp]),
    ("arc_gather_copy_cut_records", C.c_int, [vp, C.c_int, vp, C.c_longlong, P(C.c_longlong)]),
    ("arc_gather_ctx", vp, [vp]),
    ("arc_cutnet_create", C.c_int, [vp, P(vp)]),
    ("arc_cutnet_destroy", None, [vp]),
    ("arc_cutnet_init", C.c_int, [vp, C.c_uint64, vp, vp]),
    ("arc_cutnet_set_weights", C.c_int, [vp, vp, C.c_int]),
    ("arc_cutnet_get_weights", C.c_int, [vp, vp, C.c_int]),
    ("arc_cutnet_forward", C.c_int, [vp, vp, C.c_int, vp]),
    ("arc_cutnet_train", C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, vp]),
    ("arc_gather_set_cutnet", C.c_int, [vp, vp]),
    ("arc_gather_set_estimate_pool", C.c_int, [vp, C.c_int]),
    ("arc_gather_set_launch_shape", C.c_int, [vp, C.c_int, C.c_int]),
    ("arc_gather_cut_samples", C.c_int, [vp, P(arc_avpls), arc_float3, C.c_float, P(arc_gather_params), C.c_int,
                                         vp, vp, C.c_int, C.c_uint64, C.c_float, P(C.c_int)]),
    ("arc_batch_create", C.c_int, [vp, C.c_int, P(vp)]),
    ("arc_batch_lane", vp, [vp, C.c_int]),
    ("arc_batch_capture", C.c_int, [vp, vp, C.c_int]),
    ("arc_batch_set_one_cut_launch", C.c_int, [vp, C.c_int]),
    ("arc_batch_launch", C.c_int, [vp]),
    ("arc_batch_destroy", None, [vp]),
    ("arc_trace_scene_create", C.c_int, [vp, vp, vp, vp, vp, C.c_int, vp, C.c_int, C.c_int, P(vp)]),
    ("arc_trace_scene_destroy", None, [vp]),
    ("arc_trace_scene_info", C.c_int, [vp, P(C.c_int), P(C.c_int)]),
    ("arc_trace_intersect", C.c_int, [vp, vp, vp, C.c_int, vp, vp]),
    ("arc_trace_shoot", C.c_int, [vp, P(arc_area_light), C.c_int, C.c_uint64, C.c_int, P(arc_avpls), P(C.c_int)]),
    ("arc_trace_gbuffer", C.c_int, [vp, arc_float3, arc_float3, arc_float3, C.c_float, C.c_int, C.c_int, vp, vp]),
    ("arc_image_accumulate", C.c_int, [vp, vp, vp, C.c_int]),
    ("arc_image_normalize", C.c_int, [vp, vp, vp, C.c_int, C.c_int]),
    ("arc_image_error", C.c_int, [vp, vp, vp, C.c_int, P(arc_image_error_t)]),
    ("arc_write_pfm", C.c_int, [C.c_char_p, vp, C.c_int, C.c_int]),
    ("arc_read_pfm", C.c_int, [C.c_char_p, vp, C.c_int, P(C.c_int), P(C.c_int)]),
]

SCENE_SYMBOLS = [
    ("arc_scene_num_materials", C.c_int, []),
    ("arc_scene_materials", C.c_int, [vp, C.c_int]),
    ("arc_scene_camera", C.c_int, [P(arc_scene_desc), vp, vp]),
    ("arc_scene_gbuffer", C.c_int, [P(arc_scene_desc), vp, vp, vp]),
    ("arc_scene_avpls", C.c_int, [P(arc_scene_desc), vp, vp, vp, vp, vp, vp, vp]),
    ("arc_scene_points", C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_float, vp]),
    ("arc_scene_normals", C.c_int, [C.c_uint64, C.c_int, vp]),
    ("arc_scene_num_triangles", C.c_int, [P(arc_scene_desc)]),
    ("arc_scene_triangles", C.c_int, [P(arc_scene_desc), vp, vp, vp, vp, C.c_int]),
    ("arc_scene_light", C.c_int, [P(arc_scene_desc), vp, vp, vp, vp, vp, vp]),
]

_lib = None
_scene = None


def _bind(lib, symbols):
    for name, res, args in symbols:
        fn = getattr(lib, name)  # AttributeError names a missing export
        fn.restype = res
        fn.argtypes = args
    return lib


# include/arc_noalloc.h: the allocation-free SimpleBvh build
ARC_NOALLOC_MAX_POINTS = 1 << 20


class arc_lbvh_outputs(C.Structure):
    _fields_ = [("codes", vp), ("perm", vp), ("nodes", vp), ("parents", vp), ("ranges", vp), ("positions", vp),
                ("normals", vp)]


NOALLOC_SYMBOLS = [
    ("arc_bvh_build_points_noalloc", C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.POINTER(arc_lbvh_outputs)]),
]


def lib() -> C.CDLL:
    """The HIP library.  Raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m antiradiancecuts_amd.build`")
        _lib = _bind(_bind(C.CDLL(str(LIB_PATH)), ARC_SYMBOLS), NOALLOC_SYMBOLS)
    return _lib


def scene_lib() -> C.CDLL:
    global _scene
    if _scene is None:
        if not SCENE_LIB_PATH.exists():
            raise RuntimeError(f"{SCENE_LIB_PATH} is missing: build it with `python -m antiradiancecuts_amd.build`")
        _scene = _bind(C.CDLL(str(SCENE_LIB_PATH)), SCENE_SYMBOLS)
    return _scene


class ArcError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = f"{where}: status {status}"
        try:
            msg += f" ({lib().arc_status_string(status).decode()})"
        except Exception:  # pragma: no cover - library already failing
            pass
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


def check(status: int, where: str, ctx=None) -> None:
    if status != ARC_OK:
        detail = lib().arc_ctx_last_error(ctx).decode() if ctx else ""
        raise ArcError(status, where, detail)
