"""ctypes bindings of libxenon_hip.so (include/xenon_hip.h).

Loading fails loudly: there is no CPU fallback for any call. The library is
built in-tree by ``xenon_amd.build`` (``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libxenon_hip.so")

XNH_OK = 0
ERRORS = {-1: "XNH_E_INVALID", -2: "XNH_E_HIP", -3: "XNH_E_NO_SCENE", -4: "XNH_E_OVERFLOW",
          -5: "XNH_E_NO_DEVICE", -6: "XNH_E_ALLOC"}


class XnhError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class Config(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("max_bounces", C.c_int32),
                ("min_bounces", C.c_int32), ("spp_per_launch", C.c_int32), ("block_size", C.c_int32),
                ("lds_stack", C.c_int32), ("flags", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("lower_left", C.c_float * 3), ("horizontal", C.c_float * 3),
                ("vertical", C.c_float * 3)]


class BvhNode(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("hi", C.c_float * 3), ("child", C.c_uint32), ("count", C.c_uint32)]


_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


class MeshView(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("num_triangles", C.c_int32),
                ("vx", _fp), ("vy", _fp), ("vz", _fp), ("nx", _fp), ("ny", _fp), ("nz", _fp),
                ("indices", _ip), ("mat_ids", _ip), ("num_nodes", C.c_int32),
                ("nodes", C.POINTER(BvhNode)), ("prims", _u32p)]


class InstanceRec(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("instance_id", C.c_int32), ("local_to_world", C.c_float * 12),
                ("world_to_local", C.c_float * 12)]


class MaterialRec(C.Structure):
    _fields_ = [("base_color", C.c_float * 3), ("metallic", C.c_float), ("ior", C.c_float), ("alpha", C.c_float),
                ("alpha_x", C.c_float), ("alpha_y", C.c_float), ("f0", C.c_float * 3), ("subsurface", C.c_float),
                ("subsurface_color", C.c_float * 3), ("flags", C.c_uint32)]


class LightRec(C.Structure):
    _fields_ = [("mesh_id", C.c_uint32), ("tri_idx", C.c_uint32), ("emission", C.c_float * 3), ("area", C.c_float)]


class DirLightRec(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("emission", C.c_float * 3)]


class SkyRec(C.Structure):
    _fields_ = [("zenith", C.c_float * 3), ("horizon", C.c_float * 3), ("ground", C.c_float * 3),
                ("sharpness", C.c_float), ("intensity", C.c_float)]


class PrimRec(C.Structure):
    _fields_ = [("type", C.c_int32), ("mat_id", C.c_int32), ("p", C.c_float * 10)]


PRIM_SPHERE, PRIM_PLANE, PRIM_BOX, PRIM_DISK, PRIM_QUAD = 0, 1, 2, 3, 4
PRIM_GEOMETRY = -1


class FlatScene(C.Structure):
    _fields_ = [("num_meshes", C.c_int32), ("meshes", C.POINTER(MeshView)),
                ("num_instances", C.c_int32), ("instances", C.POINTER(InstanceRec)),
                ("num_tlas_nodes", C.c_int32), ("tlas_nodes", C.POINTER(BvhNode)), ("tlas_inst_indices", _u32p),
                ("num_materials", C.c_int32), ("materials", C.POINTER(MaterialRec)),
                ("num_lights", C.c_int32), ("lights", C.POINTER(LightRec)),
                ("num_dir_lights", C.c_int32), ("dir_lights", C.POINTER(DirLightRec)),
                ("sky", SkyRec),
                ("num_prims", C.c_int32), ("prims", C.POINTER(PrimRec)),
                ("num_prim_nodes", C.c_int32), ("prim_nodes", C.POINTER(BvhNode)), ("prim_order", _u32p)]


class Ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("dir", C.c_float * 3), ("tmin", C.c_float), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("instance_id", C.c_int32),
                ("prim_id", C.c_int32), ("mat_id", C.c_int32), ("pos", C.c_float * 3), ("normal", C.c_float * 3),
                ("geo_normal", C.c_float * 3), ("front_face", C.c_int32)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("width", C.c_int32), ("height", C.c_int32),
                ("spp_first", C.c_int32), ("spp_count", C.c_int32), ("max_bounces", C.c_int32),
                ("min_bounces", C.c_int32), ("rgb_out", _fp), ("sum_out", _fp)]


class Stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("node_tests", C.c_uint64), ("interior", C.c_uint64), ("tri_tests", C.c_uint64),
                ("inst_visits", C.c_uint64), ("shadow_node_tests", C.c_uint64), ("shadow_interior", C.c_uint64),
                ("shadow_tri_tests", C.c_uint64), ("shadow_inst_visits", C.c_uint64), ("nonfinite", C.c_uint64), ("render_ms", C.c_double),
                ("closest_ms", C.c_double), ("shadow_ms", C.c_double), ("shade_ms", C.c_double),
                ("closest_launches", C.c_int32), ("stack_overflows", C.c_int32),
                ("interior_iters", C.c_uint64), ("leaf_iters", C.c_uint64), ("leaf_steps", C.c_uint64),
                ("round_iters", C.c_uint64), ("round_lanes", C.c_uint64),
                ("shadow_interior_iters", C.c_uint64), ("shadow_leaf_iters", C.c_uint64),
                ("shadow_leaf_steps", C.c_uint64), ("shadow_round_iters", C.c_uint64),
                ("shadow_round_lanes", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class GuideConfig(C.Structure):
    _fields_ = [("t_log2", C.c_int32), ("learning_rate", C.c_float), ("guide_fraction", C.c_float),
                ("seed", C.c_uint32), ("bounds_lo", C.c_float * 3), ("bounds_hi", C.c_float * 3),
                ("train_batch", C.c_int32), ("record_paths", C.c_int32), ("train_steps", C.c_int32),
                ("warmup_steps", C.c_int32), ("guide_depth", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("lds_stack", C.c_int32), ("spp_per_launch", C.c_int32),
                ("ms_per_spp", C.c_float), ("candidates", C.c_int32)]


class FileInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("samples", C.c_int32),
                ("min_bounces", C.c_int32), ("max_bounces", C.c_int32), ("meshes", C.c_int32),
                ("primitives", C.c_int32), ("missing_files", C.c_int32), ("instances", C.c_int32),
                ("unknown_materials", C.c_int32),
                ("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fov", C.c_float)]


class NoiseStats(C.Structure):
    _fields_ = [("mean_squared_score", C.c_float), ("rms_score", C.c_float), ("mean_absolute", C.c_float),
                ("median_absolute", C.c_float), ("p95_absolute", C.c_float), ("max_absolute", C.c_float),
                ("avg_luminance", C.c_float), ("edge_threshold", C.c_float), ("epsilon", C.c_float),
                ("valid_pixel_count", C.c_int32), ("total_interior_pixel_count", C.c_int32),
                ("valid_pixel_ratio", C.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p
SIGNATURES = {
    "xnh_create": (C.c_int, [C.POINTER(Config), C.POINTER(_vp)]),
    "xnh_destroy": (None, [_vp]),
    "xnh_last_error": (C.c_char_p, [_vp]),
    "xnh_upload_scene": (C.c_int, [_vp, C.POINTER(FlatScene)]),
    "xnh_render": (C.c_int, [_vp, C.POINTER(Camera), C.c_int32, C.c_int32, _fp]),
    "xnh_read_accum": (C.c_int, [_vp, _fp, _ip]),
    "xnh_image_noise_stats": (C.c_int, [_vp, _fp, C.c_int32, C.c_int32, C.POINTER(NoiseStats)]),
    "xnh_image_noise_stats_noalloc": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.POINTER(NoiseStats)]),
    "xnh_image_rgb8": (C.c_int, [_vp, _fp, C.c_int32, C.c_int32, _u8p]),
    "xnh_reset": (C.c_int, [_vp]),
    "xnh_spp": (C.c_int, [_vp]),
    "xnh_get_stats": (C.c_int, [_vp, C.POINTER(Stats)]),
    "xnh_set_profiling": (C.c_int, [_vp, C.c_int32]),
    "xnh_trace_closest": (C.c_int, [_vp, C.POINTER(Ray), C.c_int32, C.POINTER(Hit)]),
    "xnh_trace_any": (C.c_int, [_vp, C.POINTER(Ray), C.c_int32, C.POINTER(C.c_uint8)]),
    "xnh_scene_upload": (C.c_int, [_vp, C.POINTER(FlatScene), C.POINTER(_vp)]),
    "xnh_scene_release": (None, [_vp, _vp]),
    "xnh_render_jobs": (C.c_int, [_vp, C.POINTER(Job), C.c_int32]),
    "xnh_scene_new": (_vp, []),
    "xnh_scene_free": (None, [_vp]),
    "xnh_scene_add_material": (C.c_int, [_vp, _fp]),
    "xnh_scene_add_mesh": (C.c_int, [_vp, C.c_int32, _fp, _fp, C.c_int32, _ip, _ip]),
    "xnh_scene_add_light": (C.c_int, [_vp, C.c_int32, C.c_int32, _fp, C.c_float]),
    "xnh_scene_add_dir_light": (C.c_int, [_vp, _fp, _fp, C.c_float]),
    "xnh_scene_set_sky": (C.c_int, [_vp, _fp, _fp, _fp, C.c_float, C.c_float]),
    "xnh_scene_add_instance": (C.c_int, [_vp, C.c_int32, _fp, C.c_float, _fp, C.c_int32]),
    "xnh_scene_add_primitive": (C.c_int, [_vp, C.c_int32, _fp, C.c_int32]),
    "xnh_scene_build": (C.c_int, [_vp, C.c_int32]),
    "xnh_scene_build_device": (C.c_int, [_vp, _vp]),
    "xnh_build_blas": (C.c_int, [_vp, C.POINTER(MeshView), C.POINTER(BvhNode), _u32p, _ip]),
    "xnh_refit_blas": (C.c_int, [_vp, C.POINTER(MeshView), C.POINTER(BvhNode)]),
    "xnh_scene_flat": (C.c_int, [_vp, C.POINTER(FlatScene)]),
    "xnh_camera_look_at": (C.c_int, [_fp, _fp, _fp, C.c_float, C.c_float, C.POINTER(Camera)]),
    "xnh_scene_load_file": (C.c_int, [C.c_char_p, C.POINTER(_vp), C.POINTER(Camera), C.POINTER(FileInfo)]),
    "xnh_material_load": (C.c_int, [C.c_char_p, _fp, C.c_char_p, C.c_int32]),
    "xnh_set_tuning": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32]),
    "xnh_get_tuning": (C.c_int, [_vp, _ip, _ip, _ip]),
    "xnh_autotune": (C.c_int, [_vp, C.POINTER(Camera), C.c_int32, C.POINTER(TuneResult)]),
    "xnh_guide_enable": (C.c_int, [_vp, C.POINTER(GuideConfig)]),
    "xnh_guide_disable": (C.c_int, [_vp]),
    "xnh_guide_set_training": (C.c_int, [_vp, C.c_int32]),
    "xnh_guide_set_params": (C.c_int, [_vp, _fp, _fp]),
    "xnh_guide_get_params": (C.c_int, [_vp, _fp, _fp]),
    "xnh_guide_eval": (C.c_int, [_vp, C.c_int32, _fp, _fp, _fp]),
    "xnh_guide_query": (C.c_int, [_vp, C.c_int32, _fp, _fp, _fp, _fp, _fp]),
    "xnh_guide_train": (C.c_int, [_vp, C.c_int32, _fp, _fp, _fp, _fp, C.c_int32, _fp, _fp, _fp]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded backend library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: the HIP backend has not been built "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        dll = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(dll, name)
            fn.restype = res
            fn.argtypes = args
        _lib = dll
    return _lib


def check(rc: int, ctx=None) -> int:
    if rc < 0:
        msg = lib().xnh_last_error(ctx).decode() if ctx else ""
        raise XnhError(rc, msg)
    return rc
