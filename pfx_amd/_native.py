"""ctypes view of include/pfxhip.h and the loader for libpfxhip.so.

The structs below mirror the C header field for field; numpy structured
dtypes with the same layout let scene arrays pass straight through.
There is no fallback: if the library is missing, importing the tracer fails.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libpfxhip.so"

# ---- numpy layouts of the scene arrays ---------------------------------------
TRIANGLE = np.dtype([("vertices", "<f4", (9,)), ("material", "<u4")])
TRIANGLE_DETAIL = np.dtype([("vertices", "<f4", (9,)), ("normals", "<f4", (9,)), ("material", "<u4")])
MESH = np.dtype([("first", "<u4"), ("count", "<u4")])
INSTANCE = np.dtype([("mesh", "<u4"), ("word", "<u4"), ("transform", "<f4", (16,))])
MATERIAL = np.dtype(
    [
        ("base", "<f4", (3,)),
        ("roughness", "<f4"),
        ("metalness", "<f4"),
        ("specular", "<f4"),
        ("clearcoat", "<f4"),
        ("clearcoat_roughness", "<f4"),
        ("sheen", "<f4"),
        ("emission", "<f4", (3,)),
        ("fresnel_power", "<f4"),
    ]
)
LIGHT = np.dtype(
    [
        ("position", "<f4", (3,)),
        ("colour", "<f4", (3,)),
        ("intensity", "<f4"),
        ("radius", "<f4"),
        ("range", "<f4"),
        ("shadow", "<u4"),
        ("kind", "<u4"),
        ("direction", "<f4", (3,)),
        ("inner_deg", "<f4"),
        ("outer_deg", "<f4"),
        ("half_u", "<f4", 3),
        ("half_v", "<f4", 3),
        ("disc_radius", "<f4"),
        ("two_sided", "<u4"),
    ]
)
PROBE_HIT = np.dtype(
    [
        ("depth", "<f4"),
        ("distance", "<f4"),
        ("material_id", "<u4"),
        ("primitive", "<u4"),
        ("instance", "<u4"),
        ("normal", "<f4", (3,)),
    ]
)
BVH_NODE = np.dtype(
    [("lo", "<f4", (4,)), ("hi", "<f4", (4,)), ("first", "<u4"), ("count", "<u4"), ("pad", "<u4", (2,))]
)
assert TRIANGLE.itemsize == 40 and TRIANGLE_DETAIL.itemsize == 76 and INSTANCE.itemsize == 72
assert MATERIAL.itemsize == 52 and LIGHT.itemsize == 96 and PROBE_HIT.itemsize == 32 and BVH_NODE.itemsize == 48

LIGHT_POINT, LIGHT_SPOT, LIGHT_RECT, LIGHT_DISC = 0, 1, 2, 3
ACCUM_COMPENSATED, ACCUM_PLAIN, ACCUM_FP64 = 0, 1, 2
PROJ_PERSPECTIVE, PROJ_ORTHOGRAPHIC, PROJ_EQUIRECT = 0, 1, 2

OK, REFUSED, SETUP, INVALID, DEVICE = 0, 1, 2, 3, 4
MAX_ROUNDS = 64


class Environment(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("texels", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3),
        ("forward", C.c_float * 3),
        ("right", C.c_float * 3),
        ("up", C.c_float * 3),
    ]


class Sun(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float)]


class SceneC(C.Structure):
    _fields_ = [
        ("triangles", C.c_void_p),
        ("triangle_count", C.c_uint32),
        ("world_details", C.c_void_p),
        ("world_detail_count", C.c_uint32),
        ("mesh_triangles", C.c_void_p),
        ("mesh_triangle_count", C.c_uint32),
        ("meshes", C.c_void_p),
        ("mesh_count", C.c_uint32),
        ("instances", C.c_void_p),
        ("instance_count", C.c_uint32),
        ("materials", C.c_void_p),
        ("material_count", C.c_uint32),
        ("lights", C.c_void_p),
        ("light_count", C.c_uint32),
        ("sky", Environment),
        ("camera", Camera),
        ("sun", Sun),
    ]


class Bounces(C.Structure):
    _fields_ = [
        ("total", C.c_uint32),
        ("diffuse", C.c_uint32),
        ("glossy", C.c_uint32),
        ("transmission", C.c_uint32),
    ]


class DetailC(C.Structure):
    _fields_ = [
        ("bounces", Bounces),
        ("lens_shift", C.c_float * 2),
        ("aperture", C.c_float),
        ("focus_distance", C.c_float),
        ("sun_radius_deg", C.c_float),
        ("accumulate", C.c_uint32),
        ("wave_paths", C.c_uint32),
        ("projection", C.c_uint32),
        ("ortho_size", C.c_float * 2),
        ("no_emitter_sampling", C.c_uint32),
        ("triangle_test", C.c_uint32),
        ("walk_order", C.c_uint32),
        ("sampler", C.c_uint32),
        ("sobol_samples", C.c_uint32),
    ]


class DeviceLimitsC(C.Structure):
    _fields_ = [("max_buffer_size", C.c_uint64), ("max_memory", C.c_uint64)]


class AdaptiveC(C.Structure):
    _fields_ = [
        ("threshold", C.c_float),
        ("min_samples", C.c_uint32),
        ("max_samples", C.c_uint32),
        ("growth", C.c_float),
    ]


class ConvergedC(C.Structure):
    _fields_ = [
        ("round_count", C.c_uint32),
        ("samples", C.c_uint32 * MAX_ROUNDS),
        ("active", C.c_uint64 * MAX_ROUNDS),
        ("pixels", C.c_uint64),
    ]


class StatsC(C.Structure):
    _fields_ = [
        ("closest_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("paths", C.c_uint64),
        ("nodes_visited", C.c_uint64),
        ("prims_tested", C.c_uint64),
        ("shadow_nodes", C.c_uint64),
        ("shadow_prims", C.c_uint64),
        ("trace_ms", C.c_double),
        ("shadow_ms", C.c_double),
        ("shade_ms", C.c_double),
        ("launches", C.c_uint32),
        ("boxes_tested", C.c_uint64),
        ("rounds_inner", C.c_uint64),
        ("rounds_leaf", C.c_uint64),
        ("lanes_inner", C.c_uint64),
        ("lanes_leaf", C.c_uint64),
        ("lanes_live", C.c_uint64),
        ("leaf_trips", C.c_uint64),
    ]


# Every symbol include/pfxhip.h declares, with its ctypes signature.
class DenoiseTile(C.Structure):
    _fields_ = [(name, C.c_uint32 * 2) for name in ("src", "dst", "size", "out_src", "out_dst", "out_size")]


class CheckpointC(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("samples", C.c_uint32),
        ("round_count", C.c_uint32),
        ("accum", C.c_void_p),
        ("color", C.c_void_p),
        ("albedo", C.c_void_p),
        ("normal", C.c_void_p),
        ("round_samples", C.c_void_p),
        ("round_active", C.c_void_p),
        ("snaps", C.c_void_p),
    ]


NO_ROUNDS = 0xFFFFFFFF

class BatchStillC(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("seed", C.c_uint32)]


class BatchTuningC(C.Structure):
    _fields_ = [("wave_paths", C.c_uint32), ("refill", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("graph", C.c_uint32)]


LEARN_LAYERS = 16
LEARN_FP32 = 1


class LearnLayerC(C.Structure):
    _fields_ = [
        ("name", C.c_char * 16),
        ("in0", C.c_int32),
        ("in1", C.c_int32),
        ("in_channels", C.c_uint32),
        ("out_channels", C.c_uint32),
        ("level", C.c_uint32),
        ("upsample", C.c_uint32),
        ("pool", C.c_uint32),
        ("relu", C.c_uint32),
    ]

_P = C.c_void_p
_U = C.c_uint32
_S = C.c_char_p
_Z = C.c_size_t
SIGNATURES = {
    "pfxh_trace_new": (C.c_int, [_P, _P, _U, _U, C.POINTER(_P), _S, _Z]),
    "pfxh_trace_sample": (C.c_int, [_P, _U, _U, _S, _Z]),
    "pfxh_trace_sample_adaptive": (C.c_int, [_P, _P, _U, _P, _S, _Z]),
    "pfxh_trace_readback": (C.c_int, [_P, _P, _P, _P, _S, _Z]),
    "pfxh_trace_readback_deviation": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_trace_set_camera": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_trace_set_window": (C.c_int, [_P, _U, _U, _U, _U, _S, _Z]),
    "pfxh_trace_set_sun": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_trace_set_lens": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_float, _S, _Z]),
    "pfxh_trace_set_projection": (C.c_int, [_P, C.c_uint32, C.c_float, C.c_float, _S, _Z]),
    "pfxh_trace_set_sun_radius": (C.c_int, [_P, C.c_float, _S, _Z]),
    "pfxh_set_device_limits": (None, [_P]),
    "pfxh_learner_set_loss": (C.c_int, [_P, _U]),
    "pfxh_learner_set_passes": (C.c_int, [_P, _U]),
    "pfxh_stage_build": (C.c_int, [_P, C.POINTER(_P), _S, _Z]),
    "pfxh_layout_sobol": (C.c_int, [_U, _P, C.POINTER(_U)]),
    "pfxh_sobol_animated_seed": (_U, [_U, _U]),
    "pfxh_staged_scene": (_P, [_P]),
    "pfxh_staged_detail": (C.c_int, [_P, _P]),
    "pfxh_staged_trace": (C.c_int, [_P, _P, _U, _U, C.POINTER(_P), _S, _Z]),
    "pfxh_staged_free": (None, [_P]),
    "pfxh_get_device_limits": (None, [_P]),
    "pfxh_trace_set_environment": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_checkpoint_write": (C.c_int, [_P, _S, _P, _Z, C.POINTER(C.c_size_t), _S, _Z]),
    "pfxh_checkpoint_read": (C.c_int, [_P, _Z, _S, _P, _P, _P, _U, _S, _Z]),
    "pfxh_trace_save": (C.c_int, [_P, _S, _P, _Z, C.POINTER(C.c_size_t), _S, _Z]),
    "pfxh_atrous": (C.c_int, [_U, _U, _P, _P, _P, _U, _P, _S, _Z]),
    "pfxh_atrous_noalloc": (C.c_int, [_U, _U, _P, _P, _P, _U, _P, _P, _S, _Z]),
    "pfxh_trace_atrous": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_denoiser_new": (C.c_int, [_S, _Z, C.POINTER(_P), _S, _Z]),
    "pfxh_denoise": (C.c_int, [_P, _U, _U, _P, _P, _P, _P, _S, _Z]),
    "pfxh_trace_denoise": (C.c_int, [_P, _P, _P, _S, _Z]),
    "pfxh_denoiser_set_max_tile": (None, [_P, C.c_uint64]),
    "pfxh_denoiser_layers": (C.c_int, [_S, _Z, C.POINTER(_U), C.POINTER(C.c_uint64), _S, _Z]),
    "pfxh_denoise_plan": (C.c_int, [_U, _U, C.c_uint64, C.POINTER(_U), C.POINTER(_U), _P, _U, C.POINTER(_U)]),
    "pfxh_denoiser_free": (None, [_P]),
    "pfxh_batch_new": (C.c_int, [_P, _U, _P, C.POINTER(_P), _S, _Z]),
    "pfxh_batch_render": (C.c_int, [_P, _S, _Z]),
    "pfxh_batch_render_samples": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_batch_tune": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_batch_get_tuning": (None, [_P, _P]),
    "pfxh_batch_set_tuning": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_batch_count": (_U, [_P]),
    "pfxh_batch_trace": (_P, [_P, _U]),
    "pfxh_batch_stats": (C.c_int, [_P, _P]),
    "pfxh_batch_reset_stats": (None, [_P]),
    "pfxh_batch_free": (None, [_P]),
    "pfxh_learn_layer": (C.c_int, [_U, _P]),
    "pfxh_learn_param_count": (C.c_uint64, []),
    "pfxh_learner_new": (C.c_int, [_U, C.POINTER(_P), _S, _Z]),
    "pfxh_learner_set_params": (C.c_int, [_P, _P, C.c_uint64, _S, _Z]),
    "pfxh_learner_params": (C.c_int, [_P, _P, C.c_uint64, _S, _Z]),
    "pfxh_learner_grads": (C.c_int, [_P, _P, C.c_uint64, _S, _Z]),
    "pfxh_learner_reset_optimizer": (None, [_P]),
    "pfxh_learner_reserve": (C.c_int, [_P, _U, _U, _S, _Z]),
    "pfxh_learner_add": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _S, _Z]),
    "pfxh_learner_crops": (_U, [_P]),
    "pfxh_learner_gradient": (C.c_int, [_P, _P, _U, C.POINTER(C.c_float), _S, _Z]),
    "pfxh_learner_train": (C.c_int, [_P, _U, _U, C.c_float, C.c_uint64, _P, _S, _Z]),
    "pfxh_learned_denoise": (C.c_int, [_P, _U, _U, _P, _P, _P, _P, _P, _P, _S, _Z]),
    "pfxh_trace_learned_denoise": (C.c_int, [_P, _P, _P, _S, _Z]),
    "pfxh_trace_learned_denoise_checked": (C.c_int, [_P, _P, _P, C.c_float, _U, _P, C.POINTER(_U), _S, _Z]),
    "pfxh_learner_free": (None, [_P]),
    "pfxh_trace_restore": (C.c_int, [_P, _S, _P, _Z, _S, _Z]),
    "pfxh_trace_samples": (_U, [_P]),
    "pfxh_trace_gpu_bytes": (C.c_uint64, [_P]),
    "pfxh_trace_node_bytes": (_U, [_P]),
    "pfxh_trace_probe": (C.c_int, [_P, _P, _S, _Z]),
    "pfxh_trace_stats": (C.c_int, [_P, _P]),
    "pfxh_trace_reset_stats": (None, [_P]),
    "pfxh_trace_set_counting": (None, [_P, C.c_int]),
    "pfxh_trace_free": (None, [_P]),
    "pfxh_layout_bvh": (C.c_int, [_P, _U, _P, C.POINTER(_U), _P, _S, _Z]),
    "pfxh_layout_levels": (C.c_int, [_P, _P, C.POINTER(_U), _P, C.POINTER(_U), _S, _Z]),
    "pfxh_layout_nodes": (C.c_int, [_P, _P, _P, C.POINTER(_U), C.POINTER(_U), _S, _Z]),
    "pfxh_layout_lights": (C.c_int, [_P, _U, _P, C.POINTER(_U), C.POINTER(_U), C.POINTER(_U), _S, _Z]),
    "pfxh_layout_select": (C.c_int, [_P, _P, _Z, _Z, C.c_int, C.POINTER(_Z)]),
    "pfxh_hbm_probe": (C.c_int, [C.c_uint64, _U, C.POINTER(C.c_double), _S, _Z]),
    "pfxh_trace_cast": (C.c_int, [_P, _P, _P, _U, C.c_float, _P, _P, _S, _Z]),
    "pfxh_layout_emitters": (C.c_int, [_P, _P, _P, C.POINTER(_U), _S, _Z]),
    "pfxh_version": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """Load libpfxhip.so (built in-tree by pfx_amd.build); fail loudly if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pfx_amd.build` (there is no CPU fallback)"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def err_buffer() -> C.Array:
    return C.create_string_buffer(1024)
