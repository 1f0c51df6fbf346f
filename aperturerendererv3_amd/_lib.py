"""ctypes declarations of include/ar.h and the loader of libar_hip.so.

The library is loaded from the package directory only.  When it is missing
the import of the GPU API raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libar_hip.so"
HEADER = Path(__file__).resolve().parent.parent / "include" / "ar.h"

AR_OK = 0
AR_E_INVALID = -1
AR_E_DEVICE = -2
AR_E_NOSCENE = -3
AR_E_OOM = -4
AR_E_UNSUPPORTED = -5
AR_E_CAPACITY = -6
AR_MISS = 0xFFFFFFFF

AR_MODEL_LAMBERT = 0
AR_MODEL_GGX = 1
AR_MODEL_GLASS = 2
AR_OBJ_BOX = 0
AR_OBJ_SDF_TORUS = 2
AR_OBJ_SDF_VOLUME = 3
AR_SDF_SHAPE_TORUS = 0
AR_SDF_SHAPE_DIAMOND = 1
AR_GBUFFER_FLOATS = 16
AR_STATS_FLOATS = 12
AR_TONE_LINEAR = 0
AR_TONE_GAMMA = 1
AR_TONE_ACES = 2
AR_VIEW_IMAGE, AR_VIEW_DENOISED, AR_VIEW_ALBEDO, AR_VIEW_NORMAL, AR_VIEW_DEPTH = 0, 1, 2, 3, 4
AR_VIEW_METALLIC, AR_VIEW_ROUGHNESS, AR_VIEW_STAT_COUNT, AR_VIEW_STAT_MEAN = 5, 6, 7, 8
AR_VIEW_STAT_M2, AR_VIEW_STAT_M3 = 9, 10
AR_OBJ_SPHERE = 1
AR_HIT_INSIDE = 16

FP = C.POINTER(C.c_float)
UP = C.POINTER(C.c_uint32)


class ar_device_opts(C.Structure):
    _fields_ = [("device", C.c_int), ("flags", C.c_uint32)]


class ar_material(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("emission", C.c_float * 3), ("specular", C.c_float),
                ("metallic", C.c_float), ("roughness", C.c_float), ("model", C.c_uint32),
                ("ior", C.c_float), ("sigma_a", C.c_float), ("abbe", C.c_float), ("tilt", C.c_float)]


class ar_object(C.Structure):
    _fields_ = [("type", C.c_uint32), ("center", C.c_float * 3), ("extent", C.c_float * 3),
                ("rotation", C.c_float * 3), ("material", ar_material), ("sdf", C.c_float * 4),
                ("sdf_volume", C.c_int32)]


class ar_mesh(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("nv", C.c_uint32), ("indices", C.c_void_p), ("ntri", C.c_uint32),
                ("tri_material", C.c_void_p), ("materials", C.c_void_p), ("nmaterials", C.c_uint32)]


class ar_rays(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax", "meps")]


class ar_hits(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("prim", "t", "u", "v", "flags")]


class ar_bvh_opts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("leaf_size", C.c_uint32), ("bin_count", C.c_uint32),
                ("max_depth", C.c_uint32)]


class ar_bvh_info(C.Structure):
    _fields_ = [("node_count", C.c_uint64), ("leaf_count", C.c_uint64), ("depth", C.c_uint32),
                ("width", C.c_uint32), ("device_bytes", C.c_uint64), ("build_seconds", C.c_double),
                ("sah_cost", C.c_double)]


class ar_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("vfov_deg", C.c_float)]


class ar_render_opts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("sample_offset", C.c_uint32),
                ("seed", C.c_uint64), ("max_depth", C.c_uint32), ("rr_min_depth", C.c_uint32), ("nee", C.c_uint32),
                ("sky", C.c_uint32), ("env", C.c_float * 3), ("max_paths_per_pass", C.c_uint32),
                ("spectral", C.c_uint32), ("output", C.c_uint32), ("guiding", C.c_uint32),
                ("frame_spp", C.c_uint32), ("frame_growth", C.c_uint32)]


class ar_mlp_opts(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("batch", C.c_uint32), ("iters_per_frame", C.c_uint32),
                ("seed", C.c_uint64)]


MLP_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 * 48 + 48


class ar_tune_result(C.Structure):
    _fields_ = [("width", C.c_uint32), ("tile", C.c_uint32), ("ms", (C.c_float * 4) * 3)]


class ar_smlp_train_opts(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("batch", C.c_uint32), ("first_step", C.c_uint32), ("total_steps", C.c_uint32),
                ("seed", C.c_uint32), ("noise", C.c_uint32), ("lr", C.c_float), ("weight_decay", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("clip_norm", C.c_float),
                ("warmup_ratio", C.c_float), ("min_lr_ratio", C.c_float), ("ema_decay", C.c_float)]


SMLP_PARAMS = 16096
AR_BASIS_SIGMOID = 0
AR_BASIS_NETWORK = 1


class ar_guiding_info(C.Structure):
    _fields_ = [("probes", C.c_uint64), ("capacity", C.c_uint64), ("samples", C.c_uint64), ("dropped", C.c_uint64),
                ("device_bytes", C.c_uint64)]


AR_SKY_NONE = 0
AR_SKY_GRADIENT = 1
AR_SKY_CONSTANT = 2
AR_SPECTRAL_RGB = 0
AR_SPECTRAL_SINGLE = 1
AR_SPECTRAL_HERO4 = 4
AR_OUTPUT_XYZ = 0
AR_OUTPUT_SRGB = 1
AR_GUIDING_OFF = 0
AR_GUIDING_VOXEL = 1
AR_GUIDING_MLP = 2

# name -> (restype, argtypes)
SIGNATURES = {
    "ar_abi_version": (C.c_int, []),
    "ar_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(ar_device_opts)]),
    "ar_destroy": (C.c_int, [C.c_void_p]),
    "ar_last_error": (C.c_char_p, [C.c_void_p]),
    "ar_scene_set": (C.c_int, [C.c_void_p, C.POINTER(ar_mesh), C.c_size_t]),
    "ar_scene_set_objects": (C.c_int, [C.c_void_p, C.POINTER(ar_object), C.c_size_t]),
    "ar_bvh_build": (C.c_int, [C.c_void_p, C.POINTER(ar_bvh_opts)]),
    "ar_bvh_info_get": (C.c_int, [C.c_void_p, C.POINTER(ar_bvh_info)]),
    "ar_trace_closest": (C.c_int, [C.c_void_p, C.POINTER(ar_rays), C.c_size_t, C.POINTER(ar_hits), C.c_void_p]),
    "ar_trace_closest_noalloc": (C.c_int, [C.c_void_p, C.POINTER(ar_rays), C.c_size_t, C.POINTER(ar_hits),
                                           C.c_void_p]),
    "ar_trace_any": (C.c_int, [C.c_void_p, C.POINTER(ar_rays), C.c_size_t, C.c_void_p, C.c_void_p]),
    "ar_trace_closest_counted": (C.c_int, [C.c_void_p, C.POINTER(ar_rays), C.c_size_t, C.POINTER(ar_hits),
                                           C.c_void_p, C.c_void_p, C.c_void_p]),
    "ar_make_shadow_rays": (C.c_int, [C.c_void_p, C.POINTER(ar_rays), C.POINTER(ar_hits), C.c_size_t,
                                      C.POINTER(C.c_float * 3), C.c_float, C.c_float, C.POINTER(ar_rays),
                                      C.c_void_p]),
    "ar_render": (C.c_int, [C.c_void_p, C.POINTER(ar_camera), C.POINTER(ar_render_opts), C.c_void_p, C.c_void_p]),
    "ar_present": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                             C.c_float, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "ar_render_counters": (C.c_int, [C.c_void_p, C.c_int]),
    "ar_render_counters_read": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ar_render_host": (C.c_int, [C.c_void_p, C.POINTER(ar_camera), C.POINTER(ar_render_opts), C.c_void_p, C.c_void_p,
                                 C.c_float, C.c_uint32]),
    "ar_render_progressive": (C.c_int, [C.c_void_p, C.POINTER(ar_camera), C.POINTER(ar_render_opts), C.c_uint32,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ar_denoise": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                             C.c_void_p]),
    "ar_tonemap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                             C.c_uint32, C.c_void_p]),
    "ar_luminance_mse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_double), C.c_void_p]),
    "ar_write_bmp": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "ar_sdf_volume_bake": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
    "ar_sdf_volume_add": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
    "ar_sdf_volume_read": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "ar_sdf_volume_clear": (C.c_int, [C.c_void_p]),
    "ar_spectrum_lut_build": (C.c_int, [C.c_void_p, C.c_size_t]),
    "ar_guiding_configure": (C.c_int, [C.c_void_p, C.c_uint64]),
    "ar_batch_create": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "ar_batch_size": (C.c_int, [C.c_void_p]),
    "ar_render_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ar_set_tile": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ar_graph_capture": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_void_p)]),
    "ar_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ar_graph_destroy": (C.c_int, [C.c_void_p]),
    "ar_autotune_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                    C.POINTER(ar_tune_result)]),
    "ar_mlp_configure": (C.c_int, [C.c_void_p, C.POINTER(ar_mlp_opts)]),
    "ar_mlp_params_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "ar_mlp_params_set": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "ar_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    "ar_mlp_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "ar_mlp_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint32,
                               C.c_void_p, C.c_void_p]),
    "ar_mlp_records_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "ar_guiding_reset": (C.c_int, [C.c_void_p]),
    "ar_guiding_info_get": (C.c_int, [C.c_void_p, C.POINTER(ar_guiding_info)]),
    "ar_guiding_snapshot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint64)]),
    "ar_guiding_restore": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "ar_spectrum_uplift": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p]),
    "ar_spectral_set_basis": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ar_spectral_mlp_set": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ar_spectral_mlp_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "ar_spectral_mlp_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "ar_spectral_mlp_build_tables": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "ar_spectral_mlp_read_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ar_spectral_mlp_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "ar_spectral_mlp_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                        C.POINTER(ar_smlp_train_opts), C.c_void_p, C.c_void_p]),
    "ar_spectral_mlp_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint32,
                                           C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ar_spectral_uplift_dataset": (C.c_int, [C.c_void_p, C.c_size_t, C.c_uint32, C.c_void_p, C.c_void_p,
                                             C.c_void_p]),
}

_lib = None


def load() -> C.CDLL:
    """Load libar_hip.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built: run python -m aperturerendererv3_amd.build")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols() -> list[str]:
    """Function names declared in include/ar.h."""
    import re
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ar_[a-z_0-9]+)\s*\(", text, re.M)))
