"""ctypes declarations of include/pyre_hip.h and include/pyre_gen.h.

Loading fails loudly when libpyre_hip.so is missing: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyre_hip.so")

PYRE_OK = 0
PYRE_EINVAL = -1
PYRE_EHIP = -2
PYRE_ENOMEM = -3
PYRE_EOVERFLOW = -4
PYRE_ESTATE = -5
PYRE_EIO = -6
PACK_BEAUTY_PNG, PACK_ALBEDO_PNG, PACK_NORMAL_PNG, PACK_DEPTH_PNG, PACK_EXR = 0, 1, 2, 3, 4
AOV_ALBEDO, AOV_NORMAL, AOV_DEPTH = 1, 2, 4
PYRE_MISS = 0xFFFFFFFF
MAT_LAMBERTIAN = 0
MAT_DISNEY = 1

_u32 = C.c_uint32
_f = C.c_float
_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_vp = C.c_void_p


class Ray(C.Structure):
    _fields_ = [("o", _f * 3), ("d", _f * 3), ("t_min", _f), ("t_max", _f), ("time", _f)]


class Ray32(C.Structure):
    _fields_ = [("o", _f * 3), ("t_min", _f), ("d", _f * 3), ("t_max", _f)]


class Hit(C.Structure):
    _fields_ = [("inst", _u32), ("tri", _u32), ("t", _f), ("u", _f), ("v", _f)]


class Hit8(C.Structure):
    _fields_ = [("prim", _u32), ("t", _f)]


class Material(C.Structure):
    _fields_ = [("kind", _u32), ("base_color", _f * 3), ("metallic", _f), ("roughness", _f), ("specular", _f)]


class QuadLight(C.Structure):
    _fields_ = [("p0", _f * 3), ("edge_u", _f * 3), ("edge_v", _f * 3), ("emission", _f * 3)]


class MeshDesc(C.Structure):
    _fields_ = [("positions", _fp), ("indices", _u32p), ("normals", _fp), ("n_vertices", _u32),
                ("n_triangles", _u32), ("material_id", _u32), ("motion", _fp)]


class EnvDesc(C.Structure):
    _fields_ = [("width", _u32), ("height", _u32), ("rgb", _fp), ("intensity", _f)]


class MeshLight(C.Structure):
    _fields_ = [("positions", _fp), ("indices", _u32p), ("n_vertices", _u32), ("n_triangles", _u32),
                ("emission", _f * 3)]


class SceneDesc(C.Structure):
    _fields_ = [("meshes", C.POINTER(MeshDesc)), ("n_meshes", _u32), ("materials", C.POINTER(Material)),
                ("n_materials", _u32), ("lights", C.POINTER(QuadLight)), ("n_lights", _u32),
                ("env", C.POINTER(EnvDesc)), ("mesh_lights", C.POINTER(MeshLight)), ("n_mesh_lights", _u32)]


class BvhParams(C.Structure):
    _fields_ = [("bins", _u32), ("max_leaf", _u32), ("traversal_cost", _f), ("intersect_cost", _f)]


class CameraDesc(C.Structure):
    _fields_ = [("origin", _f * 3), ("target", _f * 3), ("up", _f * 3), ("vfov_deg", _f), ("aspect", _f),
                ("aperture_radius", _f), ("focus_distance", _f)]


class View(C.Structure):
    _fields_ = [("has_camera", C.c_int), ("origin", _f * 3), ("target", _f * 3), ("up", _f * 3), ("vfov_deg", _f),
                ("has_film", C.c_int), ("width", _u32), ("height", _u32)]


class Tuning(C.Structure):
    _fields_ = [("pool_paths", _u32), ("trace_blocks_per_cu", _u32), ("stack_ring", _u32), ("shade_waves", _u32),
                ("traversal", _u32), ("eager", _u32), ("memory_fraction", _f)]


class IntegratorParams(C.Structure):
    _fields_ = [("max_depth", _u32), ("min_rr_depth", _u32)]


class PixelSample(C.Structure):
    _fields_ = [("radiance", _f * 3), ("albedo", _f * 3), ("normal", _f * 3), ("depth", _f)]


# name -> (restype, argtypes); every symbol the two headers declare.
SIGNATURES = {
    "pyre_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pyre_ctx_destroy": (None, [_vp]),
    "pyre_last_error": (C.c_char_p, [_vp]),
    "pyre_ctx_device_info": (C.c_int, [_vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]),
    "pyre_ctx_set_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "pyre_ctx_get_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "pyre_scene_upload": (C.c_int, [_vp, C.POINTER(SceneDesc), C.POINTER(_vp)]),
    "pyre_scene_destroy": (None, [_vp]),
    "pyre_scene_bounds": (C.c_int, [_vp, _vp, _fp, _fp]),
    "pyre_bvh_build": (C.c_int, [_vp, _vp, C.POINTER(BvhParams)]),
    "pyre_bvh_info": (C.c_int, [_vp, _vp, _u32p, _u32p, _u32p, C.POINTER(C.c_double)]),
    "pyre_bvh_export": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pyre_intersect_closest": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "pyre_intersect_any": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "pyre_trace_closest_dev": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp, _vp]),
    "pyre_trace_any_dev": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp, _vp]),
    "pyre_trace_closest_dev_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp, _vp]),
    "pyre_bsdf_eval_batch": (C.c_int, [_vp, C.POINTER(Material), _vp, _vp, C.c_size_t, _vp, _vp]),
    "pyre_bsdf_sample_batch": (C.c_int, [_vp, C.POINTER(Material), _vp, _vp, C.c_size_t, _vp, _vp, _vp, _vp]),
    "pyre_env_sample_batch": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp, _vp, _vp]),
    "pyre_env_pdf_batch": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "pyre_env_le_batch": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "pyre_light_sample_batch": (C.c_int, [_vp, _vp, _u32, _vp, _vp, C.c_size_t, _vp, _vp, _vp, _vp, _vp]),
    "pyre_light_pdf_batch": (C.c_int, [_vp, _vp, _u32, _vp, _vp, C.c_size_t, _vp]),
    "pyre_render": (C.c_int, [_vp, _vp, C.POINTER(CameraDesc), C.POINTER(IntegratorParams), _u32, _u32, _u32,
                              _u32, _vp, _vp]),
    "pyre_render_split": (C.c_int, [_vp, _vp, C.POINTER(CameraDesc), C.POINTER(IntegratorParams), _u32, _u32, _u32,
                                    _u32, _f, _u32, _vp]),
    # pyre_gen.h
    "pyre_gen_random_triangles": (C.c_int, [C.c_uint64, _u32, _f, _vp]),
    "pyre_gen_random_rays": (C.c_int, [C.c_uint64, _u32, _u32, _vp, _vp]),
    "pyre_preset_cornell": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pyre_preset_studio": (C.c_int, [C.c_int, _f, C.POINTER(_vp)]),
    "pyre_preset_production": (C.c_int, [C.c_uint64, _u32, _u32, _u32, _u32, C.POINTER(_vp)]),
    "pyre_preset_instances": (C.c_int, [C.c_uint64, _u32, _u32, _u32, _u32, _f, C.POINTER(_vp)]),
    "pyre_preset_desc": (C.c_int, [_vp, C.POINTER(SceneDesc)]),
    "pyre_preset_camera": (C.c_int, [_vp, _f, _f, C.POINTER(CameraDesc)]),
    "pyre_preset_destroy": (None, [_vp]),
    "pyre_auto_frame": (None, [_fp, _fp, _f, _f, _fp, _fp]),
    "pyre_preset_load_pbrt": (C.c_int, [C.c_char_p, C.POINTER(_vp), C.c_char_p, C.c_size_t]),
    "pyre_preset_load_pbrt_ex": (C.c_int, [C.c_char_p, _u32, C.POINTER(_vp), C.c_char_p, C.c_size_t]),
    "pyre_preset_load_gltf": (C.c_int, [C.c_char_p, C.POINTER(_vp), C.c_char_p, C.c_size_t]),
    "pyre_load_hdr": (C.c_int, [C.c_char_p, C.POINTER(_u32), C.POINTER(_u32), _vp, C.c_size_t, C.c_char_p,
                                C.c_size_t]),
    "pyre_preset_set_env_hdr": (C.c_int, [_vp, C.c_char_p, _f, C.c_char_p, C.c_size_t]),
    "pyre_preset_view": (C.c_int, [_vp, C.POINTER(View)]),
    "pyre_film_pack": (C.c_int, [_vp, _u32, _u32, _u32, _vp]),
    "pyre_write_png": (C.c_int, [C.c_char_p, _u32, _u32, _u32, _vp, C.c_char_p, C.c_size_t]),
    "pyre_write_exr": (C.c_int, [C.c_char_p, _u32, _u32, _u32, _vp, C.c_char_p, C.c_size_t]),
    "pyre_session_pack": (C.c_int, [_vp, _vp, _u32, _vp]),
}

_lib = None


def load():
    """Loads libpyre_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"pyre_amd: {LIB_PATH} is missing; build it with `python -m pyre_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class PyreError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pyre status {code}: {msg}")
        self.code = code


class NrcParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("learning_rate", _f), ("beta1", _f), ("beta2", _f), ("epsilon", _f),
                ("bounds_min", _f * 3), ("bounds_max", _f * 3)]


SIGNATURES.update({
    "pyre_render_stats": (C.c_int, [_vp, _vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "pyre_nrc_create": (C.c_int, [_vp, C.POINTER(NrcParams), C.POINTER(_vp)]),
    "pyre_nrc_destroy": (None, [_vp]),
    "pyre_nrc_param_count": (C.c_size_t, []),
    "pyre_nrc_get_weights": (C.c_int, [_vp, _vp, _vp]),
    "pyre_nrc_set_weights": (C.c_int, [_vp, _vp, _vp]),
    "pyre_nrc_infer": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "pyre_nrc_gradients": (C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, _vp, C.POINTER(_f)]),
    "pyre_nrc_train_step": (C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, C.POINTER(_f)]),
    "pyre_nrc_infer_dev": (C.c_int, [_vp, _vp, _vp, C.c_size_t, _vp, _vp]),
    "pyre_nrc_train_step_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, _vp, _vp]),
})


class NrcRenderParams(C.Structure):
    _fields_ = [("frames", _u32), ("spp_per_frame", _u32), ("train_paths", _u32), ("train_batch", _u32),
                ("train_steps", _u32), ("diffuse_roughness", _f), ("warmup_frames", _u32), ("plain_frames", _u32)]


SIGNATURES.update({
    "pyre_nrc_stats": (C.c_int, [_vp, _vp, C.POINTER(C.c_uint64), C.POINTER(_f)]),
    "pyre_nrc_reserve": (C.c_int, [_vp, _vp, C.c_size_t]),
    "pyre_session_create": (C.c_int, [_vp, _vp, C.POINTER(CameraDesc), C.POINTER(IntegratorParams), _u32, _u32,
                                      _u32, C.POINTER(_vp)]),
    "pyre_session_destroy": (None, [_vp]),
    "pyre_session_render": (C.c_int, [_vp, _vp, _u32]),
    "pyre_session_reserve": (C.c_int, [_vp, _vp, _u32]),
    "pyre_session_render_nrc": (C.c_int, [_vp, _vp, _vp, C.POINTER(NrcRenderParams), _vp]),
    "pyre_session_read": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(_u32)]),
    "pyre_session_reset": (C.c_int, [_vp, _vp]),
    "pyre_session_read_viewer": (C.c_int, [_vp, _vp, _vp]),
    "pyre_session_stats": (C.c_int, [_vp, _vp, C.POINTER(C.c_uint64)]),
})


class Job(C.Structure):
    _fields_ = [("scene", _vp), ("camera", CameraDesc), ("integrator", IntegratorParams), ("width", _u32),
                ("height", _u32), ("spp", _u32), ("spp_offset", _u32), ("out", _vp), ("accum64", _vp)]


class BatchParams(C.Structure):
    _fields_ = [("pool_paths", _u32), ("trace_blocks_per_cu", _u32), ("stack_ring", _u32), ("autotune", _u32)]


class BatchStats(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("passes", _u32), ("gpu_ms", _f),
                ("pool_paths", _u32), ("trace_blocks_per_cu", _u32), ("stack_ring", _u32)]


SIGNATURES.update({
    "pyre_render_batch": (C.c_int, [_vp, C.POINTER(Job), C.c_size_t, C.POINTER(BatchParams),
                                    C.POINTER(BatchStats)]),
})
