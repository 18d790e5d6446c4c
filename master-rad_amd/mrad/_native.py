"""ctypes binding of libmrad.so (include/mrad.h).

The library is built ahead of time (master-rad_amd/build.py) and loaded from
this package directory.  There is no fallback: if the library is missing the
import raises, so nothing silently runs without the HIP code path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import os

# MRAD_LIBRARY selects another build of the same library (e.g. a tuned variant).
LIB_PATH = Path(os.environ.get("MRAD_LIBRARY", Path(__file__).resolve().parent / "libmrad.so"))

MRAD_OK = 0
MRAD_ERR_INVALID_ARGUMENT = -1
MRAD_ERR_ILLEGAL_STATE = -2
MRAD_ERR_UNAVAILABLE = -3
MRAD_ERR_HIP = -4
MRAD_ERR_OUT_OF_MEMORY = -5
MRAD_ERR_UNSUPPORTED = -6

MRAD_BVH_UNIFORM_SAH = 0
MRAD_BVH_WEIGHTED_SAH = 1
MRAD_PRIM_SPHERE, MRAD_PRIM_BOX, MRAD_PRIM_AFFINE_SPHERE, MRAD_PRIM_AFFINE_BOX = 0, 1, 2, 3
STAT_COUNT = 14
MRAD_RENDER_FORM_MEGAKERNEL, MRAD_RENDER_FORM_WAVEFRONT = 0, 1
STAT_NAMES = ("rays", "aabb_tests", "primitive_tests", "sphere_tests", "box_tests", "plane_tests",
              "affine_sphere_tests", "affine_box_tests", "stack_overflows", "max_stack_size",
              "internal_node_visits", "leaf_node_visits", "homogeneous_leaf_node_visits",
              "mixed_leaf_node_visits")

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
up = C.POINTER(C.c_uint32)


class BvhConfig(C.Structure):
    _fields_ = [("mode", C.c_int), ("leafSize", C.c_int), ("traversalWeight", C.c_double),
                ("sphereWeight", C.c_double), ("boxWeight", C.c_double), ("affineSphereWeight", C.c_double),
                ("affineBoxWeight", C.c_double), ("lambda_", C.c_double)]


class BvhStats(C.Structure):
    _fields_ = [("node_count", C.c_int), ("root_count", C.c_int), ("leaf_count", C.c_int), ("max_depth", C.c_int),
                ("primitive_ref_count", C.c_int), ("leaf_size", C.c_int), ("generalized_sah_cost", C.c_double),
                ("uniform_sah_cost", C.c_double), ("weighted_sah_cost", C.c_double), ("bytes", C.c_int64),
                ("min_leaf_occupancy", C.c_int), ("max_leaf_occupancy", C.c_int),
                ("mean_leaf_occupancy", C.c_double), ("build_nanos", C.c_int64)]


class SceneArrays(C.Structure):
    _fields_ = [("spheres", fp), ("sphere_count", C.c_int), ("boxes", fp), ("box_count", C.c_int),
                ("planes", fp), ("plane_count", C.c_int), ("affine_spheres", fp), ("affine_sphere_count", C.c_int),
                ("affine_boxes", fp), ("affine_box_count", C.c_int), ("node_bounds", fp), ("node_count", C.c_int),
                ("node_data", ip), ("primitive_refs", ip), ("primitive_ref_count", C.c_int),
                ("root_indices", ip), ("root_count", C.c_int), ("materials", fp), ("material_count", C.c_int),
                ("background", C.c_float * 3), ("revision", C.c_uint64)]


class TracerConfig(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("samples_per_frame", C.c_int),
                ("bvh_traversal", C.c_int), ("collect_metrics", C.c_int), ("bvh_stack_size", C.c_int),
                ("samples_per_launch", C.c_int), ("device", C.c_int), ("pixels_per_launch", C.c_int),
                ("replay_rays_per_launch", C.c_int), ("max_physical_kernel_nanos", C.c_int64)]


class FrameStats(C.Structure):
    _fields_ = [("total_nanos", C.c_int64), ("kernel_nanos", C.c_int64),
                ("maximum_physical_kernel_nanos", C.c_int64), ("copy_nanos", C.c_int64),
                ("upload_nanos", C.c_int64), ("counters", C.c_uint64 * STAT_COUNT),
                ("physical_launches", C.c_int)]


# Every symbol include/mrad.h declares, with its ctypes signature.
SIGNATURES = {
    "mrad_scene_create": (C.c_int, [C.POINTER(BvhConfig), C.POINTER(C.c_void_p)]),
    "mrad_scene_destroy": (None, [C.c_void_p]),
    "mrad_scene_set_background": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "mrad_scene_add_spheres": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    "mrad_scene_add_boxes": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    "mrad_scene_add_planes": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    "mrad_scene_add_affine": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, dp, dp]),
    "mrad_scene_add_affine_transforms": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, dp]),
    "mrad_scene_build": (C.c_int, [C.c_void_p]),
    "mrad_scene_view": (C.c_int, [C.c_void_p, C.POINTER(SceneArrays)]),
    "mrad_scene_bvh_stats": (C.c_int, [C.c_void_p, C.POINTER(BvhStats)]),
    "mrad_create": (C.c_int, [C.POINTER(TracerConfig), C.POINTER(C.c_void_p)]),
    "mrad_destroy": (None, [C.c_void_p]),
    "mrad_is_available": (C.c_int, [C.c_void_p]),
    "mrad_upload_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneArrays), C.c_uint64]),
    "mrad_set_russian_roulette": (C.c_int, [C.c_void_p, C.c_int]),
    "mrad_set_leaf_batch": (C.c_int, [C.c_void_p, C.c_int]),
    "mrad_set_render_form": (C.c_int, [C.c_void_p, C.c_int]),
    "mrad_render_sample": (C.c_int, [C.c_void_p, fp, fp, C.c_int, C.c_int, C.c_uint64]),
    "mrad_render_launch_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mrad_copy_frame": (C.c_int, [C.c_void_p, fp]),
    "mrad_trace_replay": (C.c_int, [C.c_void_p, fp, C.c_int, C.c_int, ip, fp, ip, ip, fp, up]),
    "mrad_trace_replay_noalloc": (C.c_int, [C.c_void_p, fp, C.c_int, C.c_int, ip, fp, ip, ip, fp, up]),
    "mrad_get_frame_stats": (C.c_int, [C.c_void_p, C.POINTER(FrameStats)]),
    "mrad_device_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
    "mrad_code_object_sha256": (C.c_int, [C.c_char_p]),
    "mrad_validate_full_path_depth": (C.c_int, [C.c_int]),
    "mrad_last_error": (C.c_char_p, []),
}


class View(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("samples", C.c_int), ("max_depth", C.c_int),
                ("rr_start", C.c_int), ("frame_index", C.c_int), ("frame_seed", C.c_uint64),
                ("camera", C.c_float * 12)]


class WavefrontConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("queue_capacity", C.c_int), ("waves_per_cu", C.c_int),
                ("bvh_stack_size", C.c_int), ("linear_scan", C.c_int), ("collect_metrics", C.c_int),
                ("use_graph", C.c_int), ("time_kernels", C.c_int), ("lds_stack_entries", C.c_int),
                ("leaf_regroup", C.c_int), ("result_chunk_paths", C.c_int), ("stack_storage", C.c_int),
                ("trace_passes", C.c_int), ("pass_pops", C.c_int), ("suspend_capacity", C.c_int),
                ("pass_dense_lanes", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("total_nanos", C.c_int64), ("kernel_nanos", C.c_int64), ("upload_nanos", C.c_int64),
                ("steps", C.c_int64), ("paths", C.c_uint64), ("counters", C.c_uint64 * STAT_COUNT),
                ("trace_nanos", C.c_int64), ("shade_nanos", C.c_int64), ("generate_nanos", C.c_int64)]


SIGNATURES.update({
    "mrad_batch_create": (C.c_int, [C.POINTER(WavefrontConfig), C.POINTER(C.c_void_p)]),
    "mrad_batch_destroy": (None, [C.c_void_p]),
    "mrad_batch_add_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneArrays), C.POINTER(View), C.POINTER(C.c_int)]),
    "mrad_batch_render": (C.c_int, [C.c_void_p]),
    "mrad_batch_configure": (C.c_int, [C.c_void_p, C.POINTER(WavefrontConfig)]),
    "mrad_batch_copy_image": (C.c_int, [C.c_void_p, C.c_int, fp]),
    "mrad_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    "mrad_batch_get_config": (C.c_int, [C.c_void_p, C.POINTER(WavefrontConfig)]),
    "mrad_backward": (C.c_int, [C.c_void_p, fp, fp, C.c_int, C.c_int, C.c_uint64, fp, fp, C.c_int]),
    "mrad_update_materials": (C.c_int, [C.c_void_p, fp, C.c_int]),
    "mrad_cost_create": (C.c_int, [C.c_int, fp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p)]),
    "mrad_cost_time": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp]),
    "mrad_cost_output": (C.c_int, [C.c_void_p, fp, C.c_int]),
    "mrad_cost_destroy": (None, [C.c_void_p]),
})


class MradError(RuntimeError):
    """Base of the status-code exceptions raised by the binding."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class InvalidArgument(MradError, ValueError):
    """MRAD_ERR_INVALID_ARGUMENT (Java IllegalArgumentException)."""


class IllegalState(MradError):
    """MRAD_ERR_ILLEGAL_STATE / UNAVAILABLE / HIP (Java IllegalStateException)."""


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with master-rad_amd/build.py "
                          "(or __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc: int) -> None:
    if rc == MRAD_OK:
        return
    msg = (lib.mrad_last_error() or b"").decode("utf-8", "replace")
    if rc == MRAD_ERR_INVALID_ARGUMENT:
        raise InvalidArgument(rc, msg)
    raise IllegalState(rc, msg)


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None
