"""ctypes front end of the CPU checker libraries (TEST INFRASTRUCTURE ONLY).

Loads either the oracle restatement (lib/liborc_strict.so,
lib/liborc_fast.so) or the reference harness (_ref/libxenon_ref*.so); both
export the C API of oracle_api.h. Imported only by tests/ and by bench.py's
CPU-baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {
    "strict": os.path.join(HERE, "lib", "liborc_strict.so"),
    "fast": os.path.join(HERE, "lib", "liborc_fast.so"),
    "ref": os.path.join(HERE, "_ref", "libxenon_ref.so"),
    "ref_fast": os.path.join(HERE, "_ref", "libxenon_ref_fast.so"),
}

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p

_SIGS = {
    "orc_impl_name": (C.c_char_p, []),
    "orc_scene_new": (_vp, []),
    "orc_scene_free": (None, [_vp]),
    "orc_scene_add_material": (C.c_int, [_vp, _fp]),
    "orc_material_derived": (None, [_vp, C.c_int, _fp]),
    "orc_scene_add_mesh": (C.c_int, [_vp, C.c_int, _fp, _fp, C.c_int, _ip, _ip]),
    "orc_mesh_normals": (None, [_vp, C.c_int, _fp]),
    "orc_scene_add_light": (C.c_int, [_vp, C.c_int, C.c_int, _fp, C.c_float]),
    "orc_scene_add_dir_light": (C.c_int, [_vp, _fp, _fp, C.c_float]),
    "orc_scene_set_sky": (None, [_vp, _fp, _fp, _fp, C.c_float, C.c_float]),
    "orc_scene_add_instance": (C.c_int, [_vp, C.c_int, _fp, C.c_float, _fp, C.c_int]),
    "orc_scene_build": (C.c_int, [_vp]),
    "orc_scene_set_camera": (None, [_vp, _fp, _fp, _fp, C.c_float, C.c_float]),
    "orc_camera_get": (None, [_vp, _fp]),
    "orc_blas_node_count": (C.c_int, [_vp, C.c_int]),
    "orc_blas_export": (None, [_vp, C.c_int, _vp, _u32p]),
    "orc_tlas_node_count": (C.c_int, [_vp]),
    "orc_instance_count": (C.c_int, [_vp]),
    "orc_tlas_export": (None, [_vp, _vp, _u32p]),
    "orc_instance_export": (None, [_vp, _fp, _ip, _ip]),
    "orc_trace_closest": (None, [_vp, C.c_int, _fp, _fp, _ip, _u64p]),
    "orc_trace_any": (None, [_vp, C.c_int, _fp, _u8p, _u64p]),
    "orc_render": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _u64p]),
    "orc_path": (None, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp]),
    "orc_pcg_stream": (None, [C.c_uint64, C.c_uint64, C.c_int, _u32p]),
    "orc_bsdf_sample": (None, [_fp, _fp, C.c_uint64, C.c_int, _fp]),
    "orc_bsdf_eval": (None, [_fp, C.c_int, _fp, _fp, _fp]),
    "orc_noise_stats": (None, [_fp, C.c_int, C.c_int, _fp, _ip]),
    "orc_rgb8": (None, [_fp, C.c_int, C.c_int, _u8p]),
}

NOISE_FIELDS = ("mean_squared_score", "rms_score", "mean_absolute", "median_absolute", "p95_absolute",
                "max_absolute", "avg_luminance", "edge_threshold", "epsilon")

NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("child", "<u4"), ("count", "<u4")])


def default_threads() -> int:
    """Worker threads for CPU renders: the CPUs this process may run on,
    capped at 16 (the CPU share of one GPU on a shared host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(_fp)


def _arr(a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    return a


class Lib:
    """One loaded checker library."""

    def __init__(self, kind: str = "strict"):
        path = PATHS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"checker library {path} is not built (run `make -C oracle all ref`)")
        self.kind = kind
        self.dll = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(self.dll, name)
            fn.restype = res
            fn.argtypes = args
        self.is_reference = self.dll.orc_impl_name().decode() == "reference"

    # -- scene --------------------------------------------------------------
    def load(self, desc) -> "OScene":
        return OScene(self, desc)

    def pcg(self, seed: int, seq: int, n: int) -> np.ndarray:
        out = np.zeros(n, np.uint32)
        self.dll.orc_pcg_stream(seed, seq, n, out.ctypes.data_as(_u32p))
        return out

    def bsdf_sample(self, params16, wo, seed: int, n: int) -> np.ndarray:
        out = np.zeros((n, 9), np.float32)
        p = _arr(params16, np.float32)
        w = _arr(wo, np.float32)
        self.dll.orc_bsdf_sample(p.ctypes.data_as(_fp), w.ctypes.data_as(_fp), seed, n, out.ctypes.data_as(_fp))
        return out

    def noise_stats(self, rgb, width: int, height: int) -> dict:
        """NoiseStats of a framebuffer (W*H*3 floats, the layout save_image
        receives): the nine float fields by name plus valid_pixel_count,
        total_interior_pixel_count and valid_pixel_ratio."""
        img = _arr(rgb, np.float32).reshape(-1)
        o = np.zeros(9, np.float32)
        oi = np.zeros(3, np.int32)
        self.dll.orc_noise_stats(img.ctypes.data_as(_fp), width, height, o.ctypes.data_as(_fp), oi.ctypes.data_as(_ip))
        out = {k: np.float32(v) for k, v in zip(NOISE_FIELDS, o)}
        out.update(valid_pixel_count=int(oi[0]), total_interior_pixel_count=int(oi[1]),
                   valid_pixel_ratio=oi[2:3].view(np.float32)[0])
        return out

    def rgb8(self, rgb, width: int, height: int) -> np.ndarray:
        """The 8-bit pixels save_image writes, (H, W, 3) in file row order."""
        img = _arr(rgb, np.float32).reshape(-1)
        out = np.zeros((height, width, 3), np.uint8)
        self.dll.orc_rgb8(img.ctypes.data_as(_fp), width, height, out.ctypes.data_as(_u8p))
        return out

    def bsdf_eval(self, params16, wo, wi) -> np.ndarray:
        wo = _arr(wo, np.float32).reshape(-1, 3)
        wi = _arr(wi, np.float32).reshape(-1, 3)
        out = np.zeros((wo.shape[0], 4), np.float32)
        p = _arr(params16, np.float32)
        self.dll.orc_bsdf_eval(p.ctypes.data_as(_fp), wo.shape[0], wo.ctypes.data_as(_fp), wi.ctypes.data_as(_fp),
                               out.ctypes.data_as(_fp))
        return out


def _prims_export(lib, h, n):
    """(type, mat_id) pairs and 10 parameters per analytic primitive, as the
    reference's Scene::all_prims holds them (reference build only)."""
    tm = np.zeros((n, 2), np.int32)
    p = np.zeros((n, 10), np.float32)
    if n:
        fn = lib.dll.orc_prims_export
        fn.restype, fn.argtypes = None, [_vp, _ip, _fp]
        fn(h, tm.ctypes.data_as(_ip), p.ctypes.data_as(_fp))
    return tm, p


class OScene:
    """A scene built inside one checker library from a SceneDesc."""

    def __init__(self, lib: Lib, desc):
        self.lib = lib
        self.desc = desc
        d = lib.dll
        self.h = d.orc_scene_new()
        for m in desc.materials:
            p = m.params16()
            d.orc_scene_add_material(self.h, p.ctypes.data_as(_fp))
        self._keep = []
        for m in desc.meshes:
            pos = _arr(m.positions, np.float32)
            idx = _arr(m.indices, np.int32)
            mat = _arr(m.mat_ids, np.int32)
            nrm = _arr(m.normals, np.float32) if m.normals is not None else None
            d.orc_scene_add_mesh(self.h, pos.shape[0], pos.ctypes.data_as(_fp),
                                 nrm.ctypes.data_as(_fp) if nrm is not None else None,
                                 idx.shape[0], idx.ctypes.data_as(_ip), mat.ctypes.data_as(_ip))
        for lt in desc.lights:
            e = np.array(lt.emission, np.float32)
            d.orc_scene_add_light(self.h, lt.mesh, lt.tri, e.ctypes.data_as(_fp), float(np.float32(lt.area)))
        for s in desc.suns:
            dd = np.array(s.direction, np.float32)
            cc = np.array(s.color, np.float32)
            d.orc_scene_add_dir_light(self.h, dd.ctypes.data_as(_fp), cc.ctypes.data_as(_fp), s.intensity)
        sk = desc.sky
        z, h_, g = (np.array(v, np.float32) for v in (sk.zenith, sk.horizon, sk.ground))
        d.orc_scene_set_sky(self.h, z.ctypes.data_as(_fp), h_.ctypes.data_as(_fp), g.ctypes.data_as(_fp),
                            sk.intensity, sk.sharpness)
        prims = getattr(desc, "prims", [])
        if prims:
            if not lib.is_reference:
                raise NotImplementedError("analytic primitives are checked against the reference build only")
            fn = d.orc_scene_add_prim
            fn.restype, fn.argtypes = C.c_int, [_vp, C.c_int, _fp, C.c_int]
            for pr in prims:
                q = np.zeros(10, np.float32)
                q[:len(pr.params)] = pr.params
                fn(self.h, pr.type_id, q.ctypes.data_as(_fp), pr.mat)
        for ins in desc.instances:
            t = np.array(ins.translation, np.float32)
            r = np.array(ins.rot_deg, np.float32)
            d.orc_scene_add_instance(self.h, ins.mesh, t.ctypes.data_as(_fp), ins.scale, r.ctypes.data_as(_fp),
                                     ins.instance_id)
        cam = desc.camera
        e, tg, up = (np.array(v, np.float32) for v in (cam.eye, cam.target, cam.up))
        d.orc_scene_set_camera(self.h, e.ctypes.data_as(_fp), tg.ctypes.data_as(_fp), up.ctypes.data_as(_fp),
                               cam.fov, desc.aspect)
        rc = d.orc_scene_build(self.h)
        if rc != 0:
            raise RuntimeError(f"orc_scene_build failed ({rc})")

    def __del__(self):
        try:
            self.lib.dll.orc_scene_free(self.h)
        except Exception:
            pass

    def camera(self) -> np.ndarray:
        out = np.zeros(12, np.float32)
        self.lib.dll.orc_camera_get(self.h, out.ctypes.data_as(_fp))
        return out

    def blas(self, mesh: int):
        n = self.lib.dll.orc_blas_node_count(self.h, mesh)
        nodes = np.zeros(n, NODE_DTYPE)
        slots = np.zeros(self.desc.meshes[mesh].num_tris, np.uint32)
        self.lib.dll.orc_blas_export(self.h, mesh, nodes.ctypes.data_as(_vp), slots.ctypes.data_as(_u32p))
        return nodes, slots

    def tlas(self):
        n = self.lib.dll.orc_tlas_node_count(self.h)
        ni = self.lib.dll.orc_instance_count(self.h)
        nodes = np.zeros(n, NODE_DTYPE)
        order = np.zeros(ni, np.uint32)
        self.lib.dll.orc_tlas_export(self.h, nodes.ctypes.data_as(_vp), order.ctypes.data_as(_u32p))
        xf = np.zeros((ni, 24), np.float32)
        ids = np.zeros(ni, np.int32)
        meshes = np.zeros(ni, np.int32)
        self.lib.dll.orc_instance_export(self.h, xf.ctypes.data_as(_fp), ids.ctypes.data_as(_ip),
                                         meshes.ctypes.data_as(_ip))
        return nodes, order, xf, ids, meshes

    def prims(self):
        return _prims_export(self.lib, self.h, len(getattr(self.desc, "prims", [])))

    def normals(self, mesh: int) -> np.ndarray:
        out = np.zeros((self.desc.meshes[mesh].positions.shape[0], 3), np.float32)
        self.lib.dll.orc_mesh_normals(self.h, mesh, out.ctypes.data_as(_fp))
        return out

    def trace_closest(self, rays: np.ndarray, counters: bool = False):
        rays = _arr(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hf = np.zeros((n, 12), np.float32)
        hi = np.zeros((n, 4), np.int32)
        cnt = np.zeros(4, np.uint64)
        self.lib.dll.orc_trace_closest(self.h, n, rays.ctypes.data_as(_fp), hf.ctypes.data_as(_fp),
                                       hi.ctypes.data_as(_ip), cnt.ctypes.data_as(_u64p) if counters else None)
        return (hf, hi, cnt) if counters else (hf, hi)

    def trace_any(self, rays: np.ndarray, counters: bool = False):
        rays = _arr(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint8)
        cnt = np.zeros(4, np.uint64)
        self.lib.dll.orc_trace_any(self.h, n, rays.ctypes.data_as(_fp), occ.ctypes.data_as(_u8p),
                                   cnt.ctypes.data_as(_u64p) if counters else None)
        return (occ, cnt) if counters else occ

    def render_sum(self, spp_first: int, spp_count: int, nthreads: int = 0, width=None, height=None,
                   accum=None):
        """Restatement only: adds samples spp_first..spp_first+spp_count-1
        into a (H*W, 3) float32 sum (pixel order); returns (sum, stats)."""
        if self.lib.is_reference:
            raise RuntimeError("the reference harness only renders averages from sample 1")
        W = width or self.desc.width
        H = height or self.desc.height
        if accum is None:
            accum = np.zeros((H * W, 3), np.float32)
        stats = np.zeros(3, np.uint64)
        nthreads = nthreads or default_threads()
        rc = self.lib.dll.orc_render(self.h, W, H, spp_first, spp_count, self.desc.min_bounces,
                                     self.desc.max_bounces, nthreads, accum.ctypes.data_as(_fp),
                                     stats.ctypes.data_as(_u64p))
        if rc != 0:
            raise RuntimeError(f"orc_render failed ({rc})")
        return accum, stats

    def render_average(self, spp_count: int, nthreads: int = 0, width=None, height=None) -> np.ndarray:
        """Average of samples 1..spp_count in pixel order, (H*W, 3), computed
        as the reference finalises a frame: sum * (1/spp) in float32."""
        W = width or self.desc.width
        H = height or self.desc.height
        if self.lib.is_reference:
            out = np.zeros((H * W, 3), np.float32)
            rc = self.lib.dll.orc_render(self.h, W, H, 1, spp_count, self.desc.min_bounces,
                                         self.desc.max_bounces, nthreads, out.ctypes.data_as(_fp), None)
            if rc != 0:
                raise RuntimeError(f"reference render failed ({rc})")
            return out
        acc, _ = self.render_sum(1, spp_count, nthreads, W, H)
        inv = np.float32(1.0) / np.float32(spp_count)
        return acc * inv

    def path(self, x, y, spp, width=None, height=None):
        out = np.zeros(3, np.float32)
        self.lib.dll.orc_path(self.h, width or self.desc.width, height or self.desc.height, x, y, spp,
                              self.desc.min_bounces, self.desc.max_bounces, out.ctypes.data_as(_fp))
        return out

    def path_rays(self, x, y, spp, width=None, height=None, max_rays=64):
        """The closest-hit rays of one path (restatement only): (n, 8) rays
        (o, d, tmin, tmax) and (n, 3) hits (t, instance, triangle)."""
        fn = self.lib.dll.orc_path_rays
        fn.restype = C.c_int
        fn.argtypes = [_vp] + [C.c_int] * 8 + [_fp]
        out = np.zeros((max_rays, 11), np.float32)
        n = fn(self.h, width or self.desc.width, height or self.desc.height, x, y, spp, self.desc.min_bounces,
               self.desc.max_bounces, max_rays, out.ctypes.data_as(_fp))
        n = min(n, max_rays)
        return out[:n, :8].copy(), out[:n, 8:].copy()


# ---- scene files: the reference's own load_scene (oracle/_ref only) --------
_FILE_SIGS = {
    "orc_ref_load_scene_file": (_vp, [C.c_char_p, _ip]),
    "orc_ref_scene_sizes": (None, [_vp, _ip]),
    "orc_ref_mesh_size": (None, [_vp, C.c_int, _ip]),
    "orc_ref_mesh_export": (None, [_vp, C.c_int, _fp, _fp, _ip, _ip]),
    "orc_ref_material_params": (None, [_vp, C.c_int, _fp]),
    "orc_ref_lights_export": (None, [_vp, _fp]),
    "orc_ref_suns_export": (None, [_vp, _fp]),
    "orc_ref_sky_export": (None, [_vp, _fp]),
}


class RefFileScene:
    """A scene loaded from a .xenon file by the reference's load_scene
    (scene/scene_file.cpp:49-240), inside oracle/_ref/libxenon_ref.so."""

    def __init__(self, lib: Lib, path: str):
        if not lib.is_reference:
            raise ValueError("scene files load through the reference library only")
        d = lib.dll
        for name, (res, args) in _FILE_SIGS.items():
            fn = getattr(d, name)
            fn.restype, fn.argtypes = res, args
        self.lib = lib
        cfg = np.zeros(5, np.int32)
        self.h = d.orc_ref_load_scene_file(str(path).encode(), cfg.ctypes.data_as(_ip))
        if not self.h:
            raise FileNotFoundError(path)
        self.config = cfg
        sz = np.zeros(5, np.int32)
        d.orc_ref_scene_sizes(self.h, sz.ctypes.data_as(_ip))
        self.num_meshes, self.num_materials, self.num_lights, self.num_suns, self.num_prims = (int(v) for v in sz)

    def __del__(self):
        try:
            self.lib.dll.orc_scene_free(self.h)
        except Exception:
            pass

    def mesh(self, i: int):
        d = self.lib.dll
        nvt = np.zeros(2, np.int32)
        d.orc_ref_mesh_size(self.h, i, nvt.ctypes.data_as(_ip))
        nv, nt = int(nvt[0]), int(nvt[1])
        pos = np.zeros((nv, 3), np.float32)
        nrm = np.zeros((nv, 3), np.float32)
        idx = np.zeros((nt, 3), np.int32)
        mat = np.zeros(nt, np.int32)
        d.orc_ref_mesh_export(self.h, i, pos.ctypes.data_as(_fp), nrm.ctypes.data_as(_fp), idx.ctypes.data_as(_ip),
                              mat.ctypes.data_as(_ip))
        return pos, nrm, idx, mat

    def material_params(self, i: int) -> np.ndarray:
        out = np.zeros(16, np.float32)
        self.lib.dll.orc_ref_material_params(self.h, i, out.ctypes.data_as(_fp))
        return out

    def material_derived(self, i: int) -> np.ndarray:
        out = np.zeros(11, np.float32)
        self.lib.dll.orc_material_derived(self.h, i, out.ctypes.data_as(_fp))
        return out

    def lights(self) -> np.ndarray:
        out = np.zeros((self.num_lights, 6), np.float32)
        self.lib.dll.orc_ref_lights_export(self.h, out.ctypes.data_as(_fp))
        return out

    def suns(self) -> np.ndarray:
        out = np.zeros((self.num_suns, 6), np.float32)
        self.lib.dll.orc_ref_suns_export(self.h, out.ctypes.data_as(_fp))
        return out

    def sky(self) -> np.ndarray:
        out = np.zeros(11, np.float32)
        self.lib.dll.orc_ref_sky_export(self.h, out.ctypes.data_as(_fp))
        return out

    def prims(self):
        return _prims_export(self.lib, self.h, self.num_prims)

    def camera(self) -> np.ndarray:
        out = np.zeros(12, np.float32)
        self.lib.dll.orc_camera_get(self.h, out.ctypes.data_as(_fp))
        return out

    def blas(self, mesh: int):
        d = self.lib.dll
        n = d.orc_blas_node_count(self.h, mesh)
        nvt = np.zeros(2, np.int32)
        d.orc_ref_mesh_size(self.h, mesh, nvt.ctypes.data_as(_ip))
        nodes = np.zeros(n, NODE_DTYPE)
        slots = np.zeros(int(nvt[1]), np.uint32)
        d.orc_blas_export(self.h, mesh, nodes.ctypes.data_as(_vp), slots.ctypes.data_as(_u32p))
        return nodes, slots

    def tlas(self):
        d = self.lib.dll
        n = d.orc_tlas_node_count(self.h)
        ni = d.orc_instance_count(self.h)
        nodes = np.zeros(n, NODE_DTYPE)
        order = np.zeros(ni, np.uint32)
        d.orc_tlas_export(self.h, nodes.ctypes.data_as(_vp), order.ctypes.data_as(_u32p))
        xf = np.zeros((ni, 24), np.float32)
        ids = np.zeros(ni, np.int32)
        meshes = np.zeros(ni, np.int32)
        d.orc_instance_export(self.h, xf.ctypes.data_as(_fp), ids.ctypes.data_as(_ip), meshes.ctypes.data_as(_ip))
        return nodes, order, xf, ids, meshes

    def place_instances(self, instances):
        """Replaces the TLAS by one over (mesh, t3, scale, rot3, id) placements,
        as orc_scene_build does for explicit instances."""
        d = self.lib.dll
        for mesh, t, s, rot, iid in instances:
            tt = np.array(t, np.float32)
            rr = np.array(rot, np.float32)
            d.orc_scene_add_instance(self.h, int(mesh), tt.ctypes.data_as(_fp), float(np.float32(s)),
                                     rr.ctypes.data_as(_fp), int(iid))
        if d.orc_scene_build(self.h) != 0:
            raise RuntimeError("orc_scene_build failed")
