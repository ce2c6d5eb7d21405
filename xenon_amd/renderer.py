"""Python host mirror of the reference's GPU-backend contract.

``HipRenderer`` mirrors ``CudaRenderer`` (src/cuda/cuda_renderer.cuh:14-55):
constructed with (width, height, max_bounces, min_bounces), it takes a built
scene once through ``upload_scene`` and adds exactly one sample per
``render_frame``; ``spp()`` reports the count. ``HipScene`` plays the role of
the reference's ``Scene`` after ``build_acceleration`` (scene/scene.h:79-121).
Every call goes through libxenon_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N
from .scenes import SceneDesc

NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("child", "<u4"), ("count", "<u4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("instance_id", "<i4"), ("prim_id", "<i4"),
                      ("mat_id", "<i4"), ("pos", "<f4", 3), ("normal", "<f4", 3), ("geo_normal", "<f4", 3),
                      ("front_face", "<i4")])
assert HIT_DTYPE.itemsize == C.sizeof(N.Hit)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class HipScene:
    """Scene built by the backend's host builder from a ``SceneDesc``."""

    def __init__(self, desc: SceneDesc, device_blas: "HipRenderer | None" = None):
        """desc's scene, built on the host as Scene::build_acceleration builds
        it (the reference's trees); with device_blas, every mesh's BLAS is
        built on that renderer's device instead (xnh_build_blas: a linear
        BVH, hits equal up to equal-t ties)."""
        L = N.lib()
        self.desc = desc
        self.h = L.xnh_scene_new()
        if not self.h:
            raise MemoryError("xnh_scene_new")
        for m in desc.materials:
            N.check(L.xnh_scene_add_material(self.h, _fp(m.params16())))
        for m in desc.meshes:
            pos = np.ascontiguousarray(m.positions, np.float32)
            idx = np.ascontiguousarray(m.indices, np.int32)
            mat = np.ascontiguousarray(m.mat_ids, np.int32)
            nrm = None if m.normals is None else np.ascontiguousarray(m.normals, np.float32)
            N.check(L.xnh_scene_add_mesh(self.h, pos.shape[0], _fp(pos), _fp(nrm) if nrm is not None else None,
                                         idx.shape[0], _ip(idx), _ip(mat)))
        for lt in desc.lights:
            e = np.array(lt.emission, np.float32)
            N.check(L.xnh_scene_add_light(self.h, lt.mesh, lt.tri, _fp(e), float(np.float32(lt.area))))
        for s in desc.suns:
            d = np.array(s.direction, np.float32)
            c = np.array(s.color, np.float32)
            N.check(L.xnh_scene_add_dir_light(self.h, _fp(d), _fp(c), s.intensity))
        sk = desc.sky
        z, hz, g = (np.array(v, np.float32) for v in (sk.zenith, sk.horizon, sk.ground))
        N.check(L.xnh_scene_set_sky(self.h, _fp(z), _fp(hz), _fp(g), sk.intensity, sk.sharpness))
        for pr in desc.prims:
            q = np.zeros(10, np.float32)
            q[:len(pr.params)] = pr.params
            N.check(L.xnh_scene_add_primitive(self.h, pr.type_id, _fp(q), pr.mat))
        for ins in desc.instances:
            t = np.array(ins.translation, np.float32)
            r = np.array(ins.rot_deg, np.float32)
            N.check(L.xnh_scene_add_instance(self.h, ins.mesh, _fp(t), ins.scale, _fp(r), ins.instance_id))
        if device_blas is None:
            N.check(L.xnh_scene_build(self.h, 0))
        else:
            N.check(L.xnh_scene_build_device(self.h, device_blas.h), device_blas.h)
        self._flat = N.FlatScene()
        N.check(L.xnh_scene_flat(self.h, C.byref(self._flat)))

    @classmethod
    def from_file(cls, path: str) -> "HipScene":
        """A .xenon scene file (load_scene, scene/scene_file.cpp:49-240) with
        its .mat materials and OBJ meshes. ``info`` holds the file's render
        settings and counts (meshes, analytic primitives, instances, files
        that did not open, unknown material names); ``camera()`` uses the
        file's camera."""
        self = cls.__new__(cls)
        self.desc = None
        h = C.c_void_p()
        cam = N.Camera()
        info = N.FileInfo()
        N.check(N.lib().xnh_scene_load_file(str(path).encode(), C.byref(h), C.byref(cam), C.byref(info)))
        self.h = h.value
        self.info = info
        self.file_camera = cam
        self._flat = N.FlatScene()
        N.check(N.lib().xnh_scene_flat(self.h, C.byref(self._flat)))
        return self

    @property
    def width(self) -> int:
        return self.desc.width if self.desc is not None else self.info.width

    @property
    def height(self) -> int:
        return self.desc.height if self.desc is not None else self.info.height

    def __del__(self):
        try:
            if self.h:
                N.lib().xnh_scene_free(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def flat(self) -> N.FlatScene:
        return self._flat

    def camera(self, width=None, height=None) -> N.Camera:
        W = width or self.width
        H = height or self.height
        aspect = float(np.float32(W) / np.float32(H))
        if self.desc is not None:
            cam = self.desc.camera
            e, t, u = (np.array(v, np.float32) for v in (cam.eye, cam.target, cam.up))
            fov = cam.fov
        else:
            e, t, u = (np.array(v[:], np.float32) for v in (self.info.eye, self.info.target, self.info.up))
            fov = self.info.fov
        out = N.Camera()
        N.check(N.lib().xnh_camera_look_at(_fp(e), _fp(t), _fp(u), fov, aspect, C.byref(out)))
        return out

    def blas(self, mesh: int):
        mv = self._flat.meshes[mesh]
        if mv.num_nodes == 0:
            return np.zeros(0, NODE_DTYPE), np.zeros(0, np.uint32)
        nodes = np.ctypeslib.as_array(C.cast(mv.nodes, C.POINTER(C.c_uint8)), shape=(mv.num_nodes * 32,))
        nodes = nodes.view(NODE_DTYPE).copy()
        prims = np.ctypeslib.as_array(mv.prims, shape=(mv.num_triangles,)).copy()
        return nodes, prims

    def normals(self, mesh: int) -> np.ndarray:
        mv = self._flat.meshes[mesh]
        n = mv.num_vertices
        return np.stack([np.ctypeslib.as_array(p, shape=(n,)) for p in (mv.nx, mv.ny, mv.nz)], 1).copy()

    def tlas(self):
        f = self._flat
        nodes = np.ctypeslib.as_array(C.cast(f.tlas_nodes, C.POINTER(C.c_uint8)), shape=(f.num_tlas_nodes * 32,))
        nodes = nodes.view(NODE_DTYPE).copy()
        order = np.ctypeslib.as_array(f.tlas_inst_indices, shape=(f.num_instances,)).copy()
        ni = f.num_instances
        xf = np.zeros((ni, 24), np.float32)
        ids = np.zeros(ni, np.int32)
        meshes = np.zeros(ni, np.int32)
        for i in range(ni):
            r = f.instances[i]
            xf[i, :12] = np.array(r.local_to_world[:], np.float32)
            xf[i, 12:] = np.array(r.world_to_local[:], np.float32)
            ids[i] = r.instance_id
            meshes[i] = r.mesh
        return nodes, order, xf, ids, meshes


    def prims(self):
        """The analytic primitives as stored: ((type, mat_id) per primitive,
        10 parameters per primitive in xnh_prim order)."""
        f = self._flat
        tm = np.zeros((f.num_prims, 2), np.int32)
        p = np.zeros((f.num_prims, 10), np.float32)
        for i in range(f.num_prims):
            r = f.prims[i]
            tm[i] = (r.type, r.mat_id)
            p[i] = r.p[:]
        return tm, p

    def prim_blas(self):
        """The primitives' BVH: (nodes, leaf slot -> primitive index)."""
        f = self._flat
        if f.num_prim_nodes == 0:
            return np.zeros(0, NODE_DTYPE), np.zeros(0, np.uint32)
        nodes = np.ctypeslib.as_array(C.cast(f.prim_nodes, C.POINTER(C.c_uint8)), shape=(f.num_prim_nodes * 32,))
        nodes = nodes.view(NODE_DTYPE).copy()
        order = np.ctypeslib.as_array(f.prim_order, shape=(f.num_prims,)).copy()
        return nodes, order

    def bounds(self):
        """World bounding box of the scene (the TLAS root box): (lo, hi)."""
        f = self._flat
        if f.num_tlas_nodes == 0:
            return np.zeros(3, np.float32), np.ones(3, np.float32)
        root = np.ctypeslib.as_array(C.cast(f.tlas_nodes, C.POINTER(C.c_uint8)), shape=(32,)).view(NODE_DTYPE)[0]
        return np.array(root["lo"], np.float32), np.array(root["hi"], np.float32)


class GpuScene:
    """A scene kept resident on the device for batched jobs (xnh_scene_upload)."""

    def __init__(self, renderer: "HipRenderer", scene: HipScene):
        self.renderer = renderer
        self.scene = scene
        h = C.c_void_p()
        renderer._chk(N.lib().xnh_scene_upload(renderer.h, C.byref(scene.flat), C.byref(h)))
        self.h = h

    def release(self):
        if getattr(self, "h", None) and self.renderer.h:
            N.lib().xnh_scene_release(self.renderer.h, self.h)
        self.h = None


class HipRenderer:
    """The backend renderer: ``CudaRenderer``'s three-method contract plus
    batch rendering and the ray-batch parity API."""

    def __init__(self, width: int, height: int, max_bounces: int = 8, min_bounces: int = 3,
                 spp_per_launch: int = 0, block_size: int = 0, lds_stack: int = 0, use_graph: bool = True):
        self.width, self.height = int(width), int(height)
        cfg = N.Config(self.width, self.height, int(max_bounces), int(min_bounces), int(spp_per_launch),
                       int(block_size), int(lds_stack), 0 if use_graph else 1)
        h = C.c_void_p()
        N.check(N.lib().xnh_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.scene = None

    def close(self):
        if getattr(self, "h", None):
            N.lib().xnh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        return N.check(rc, self.h)

    # -- CudaRenderer contract ---------------------------------------------
    def upload_scene(self, scene: HipScene) -> None:
        self._chk(N.lib().xnh_upload_scene(self.h, C.byref(scene.flat)))
        self.scene = scene

    def render_frame(self, camera: N.Camera, chain=None) -> np.ndarray:
        """Adds exactly one sample (index spp()+1) and returns the running
        average as the reference's swapchain holds it: (H, W, 3), row 0 at the
        top (rows flipped, render/cpu_tiled.cpp:228-229). With a
        TripleSwapchain the frame goes into its write buffer, which is then
        published (cuda_renderer.cu:501-590)."""
        img = self.render(camera, self.spp() + 1, 1)
        frame = img.reshape(self.height, self.width, 3)[::-1]
        if chain is None:
            return frame.copy()
        out = chain.get_write_buffer()
        np.copyto(out, frame)
        chain.swap_writer()
        return out

    def spp(self) -> int:
        return N.lib().xnh_spp(self.h)

    # -- batch API ----------------------------------------------------------
    def render(self, camera: N.Camera, spp_first: int, spp_count: int, want_image: bool = True):
        """Adds samples spp_first..spp_first+spp_count-1 (1-based); returns the
        running average in pixel order, (H*W, 3), or None."""
        out = np.empty((self.height * self.width, 3), np.float32) if want_image else None
        self._chk(N.lib().xnh_render(self.h, C.byref(camera), int(spp_first), int(spp_count),
                                     _fp(out) if out is not None else None))
        return out

    # -- BLAS on the device ------------------------------------------------
    @staticmethod
    def _mesh_view(positions, indices, nodes=None, prims=None):
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
        idx = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
        v = N.MeshView()
        v.num_vertices, v.num_triangles = pos.shape[0], idx.shape[0]
        v.vx, v.vy, v.vz, v.indices = _fp(x), _fp(y), _fp(z), _ip(idx)
        keep = [x, y, z, idx]
        if nodes is not None:
            nd = np.ascontiguousarray(nodes)
            pr = np.ascontiguousarray(prims, np.uint32)
            v.num_nodes = nd.shape[0]
            v.nodes = nd.ctypes.data_as(C.POINTER(N.BvhNode))
            v.prims = pr.ctypes.data_as(N._u32p)
            keep += [nd, pr]
        return v, keep

    def build_blas(self, positions, indices):
        """A BLAS built on the device (xnh_build_blas): (nodes, leaf slot ->
        triangle), nodes in NODE_DTYPE."""
        v, keep = self._mesh_view(positions, indices)
        n = v.num_triangles
        nodes = np.zeros(max(2 * n - 1, 0), NODE_DTYPE)
        prims = np.zeros(n, np.uint32)
        nn = C.c_int32()
        self._chk(N.lib().xnh_build_blas(self.h, C.byref(v), nodes.ctypes.data_as(C.POINTER(N.BvhNode)),
                                         prims.ctypes.data_as(N._u32p), C.byref(nn)))
        del keep
        return nodes[:nn.value], prims

    def refit_blas(self, positions, indices, nodes, prims):
        """The boxes of (nodes, prims) recomputed for new vertex positions
        (xnh_refit_blas); the topology is kept."""
        v, keep = self._mesh_view(positions, indices, np.asarray(nodes, NODE_DTYPE), prims)
        out = np.zeros(v.num_nodes, NODE_DTYPE)
        self._chk(N.lib().xnh_refit_blas(self.h, C.byref(v), out.ctypes.data_as(C.POINTER(N.BvhNode))))
        del keep
        return out

    # -- output stage (main.cpp:25-52, math/image.h) ---------------------
    def _framebuffer(self, rgb, width, height):
        if rgb is None:
            return None, self.width, self.height
        a = np.ascontiguousarray(rgb, np.float32)
        if width is None:
            if a.ndim != 3:
                raise ValueError("give width and height for a flat framebuffer")
            height, width = a.shape[:2]
        if a.size != width * height * 3:
            raise ValueError("framebuffer size does not match width x height x 3")
        return a, width, height

    def noise_stats(self, rgb=None, width=None, height=None) -> dict:
        """NoiseStats (math/image.h:35-118) computed on the device, of a
        framebuffer in the swapchain layout (row 0 at the top, as
        render_frame returns it), or of the running average when rgb is
        None."""
        a, w, h = self._framebuffer(rgb, width, height)
        st = N.NoiseStats()
        self._chk(N.lib().xnh_image_noise_stats(self.h, _fp(a) if a is not None else None, w, h, C.byref(st)))
        return st.as_dict()

    def image_rgb8(self, rgb=None, width=None, height=None) -> np.ndarray:
        """The 8-bit pixels the reference's save_image writes (main.cpp:26-31),
        (H, W, 3) in the PNG's row order."""
        a, w, h = self._framebuffer(rgb, width, height)
        out = np.empty((h, w, 3), np.uint8)
        self._chk(N.lib().xnh_image_rgb8(self.h, _fp(a) if a is not None else None, w, h,
                                         out.ctypes.data_as(N._u8p)))
        return out

    def save_image(self, path: str, rgb=None, width=None, height=None, report: bool = True) -> dict:
        """save_image (main.cpp:25-52): the PNG of the running average (or of
        the given swapchain-layout framebuffer) with the reference's pixel
        bytes and row order, and its NoiseStats, printed as the reference
        prints them when `report`."""
        from .imageio import write_png_rows
        write_png_rows(path, self.image_rgb8(rgb, width, height))
        st = self.noise_stats(rgb, width, height)
        if report:
            for k in ("mean_squared_score", "rms_score", "mean_absolute", "median_absolute", "p95_absolute",
                      "max_absolute", "avg_luminance", "edge_threshold", "epsilon", "valid_pixel_count",
                      "valid_pixel_ratio"):
                name = "".join(w.capitalize() if i else w for i, w in enumerate(k.split("_")))
                v = st[k]
                print(f"{name:<16} = {v if isinstance(v, int) else float(f'{v:.10g}')}")
            print(f"Saved image to {path}")
        return st

    def read_accum(self):
        out = np.empty((self.height * self.width, 3), np.float32)
        spp = C.c_int32()
        self._chk(N.lib().xnh_read_accum(self.h, _fp(out), C.byref(spp)))
        return out, spp.value

    def reset(self):
        self._chk(N.lib().xnh_reset(self.h))

    PROFILE_TIMING, PROFILE_WORK = 1, 2

    def set_profiling(self, timing: bool = False, work: bool = False):
        """Per-kernel event timing and/or traversal work counters for the
        following renders (both off by default)."""
        self._chk(N.lib().xnh_set_profiling(self.h, (self.PROFILE_TIMING if timing else 0) |
                                            (self.PROFILE_WORK if work else 0)))

    def stats(self) -> dict:
        s = N.Stats()
        self._chk(N.lib().xnh_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def render_jobs(self, jobs, want_sums: bool = False, want_images: bool = True):
        """Renders a list of jobs together. Each job is a dict with keys
        scene (GpuScene or None for the uploaded scene), camera, width,
        height, spp_first, spp_count, max_bounces, min_bounces. Returns one
        (H*W, 3) average per job (and the sums when want_sums); with
        want_images=False and no sums nothing is copied back and the
        averages are None."""
        arr = (N.Job * len(jobs))()
        outs, sums = [], []
        for k, j in enumerate(jobs):
            W, H = int(j["width"]), int(j["height"])
            o = np.empty((W * H, 3), np.float32) if want_images else None
            sm = np.empty((W * H, 3), np.float32) if want_sums else None
            outs.append(o)
            sums.append(sm)
            sc = j.get("scene")
            arr[k] = N.Job(sc.h if sc is not None else None, j["camera"], W, H, int(j.get("spp_first", 1)),
                           int(j["spp_count"]), int(j["max_bounces"]), int(j["min_bounces"]),
                           _fp(o) if o is not None else None,
                           _fp(sm) if sm is not None else None)
        self._chk(N.lib().xnh_render_jobs(self.h, arr, len(jobs)))
        return (outs, sums) if want_sums else outs

    # -- tuning -----------------------------------------------------------------
    def set_tuning(self, block_size: int = 0, lds_stack: int = 0, spp_per_launch: int = 0):
        """Traversal workgroup size, LDS stack entries per lane and samples per
        batch (queue capacity); 0 = default. Clears the accumulation."""
        self._chk(N.lib().xnh_set_tuning(self.h, int(block_size), int(lds_stack), int(spp_per_launch)))

    def tuning(self) -> dict:
        b, l, n = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(N.lib().xnh_get_tuning(self.h, C.byref(b), C.byref(l), C.byref(n)))
        return {"block_size": b.value, "lds_stack": l.value, "spp_per_launch": n.value}

    def autotune(self, camera: N.Camera, probe_spp: int = 8, cache: str | None = "default") -> dict:
        """Picks the fastest (workgroup size, LDS stack depth, queue capacity)
        for the uploaded scene by timing every candidate on `probe_spp`
        samples, or reuses the choice cached for the same scene shape,
        resolution and depth (cache=None: always measure). Clears the
        accumulation."""
        import json
        import os
        key = None
        path = None
        if cache is not None and self.scene is not None:
            path = (os.path.join(os.path.expanduser("~"), ".cache", "xenon_amd", "autotune.json")
                    if cache == "default" else cache)
            f = self.scene._flat
            key = (f"{self.width}x{self.height}:tris{sum(f.meshes[m].num_triangles for m in range(f.num_meshes))}"
                   f":inst{f.num_instances}:probe{probe_spp}")
            try:
                with open(path) as fh:
                    hit = json.load(fh).get(key)
            except (OSError, ValueError):
                hit = None
            if hit:
                self.set_tuning(hit["block_size"], hit["lds_stack"], hit["spp_per_launch"])
                return dict(hit, cached=True)
        res = N.TuneResult()
        self._chk(N.lib().xnh_autotune(self.h, C.byref(camera), int(probe_spp), C.byref(res)))
        out = {"block_size": res.block_size, "lds_stack": res.lds_stack, "spp_per_launch": res.spp_per_launch,
               "ms_per_spp": res.ms_per_spp, "candidates": res.candidates}
        if key is not None:
            try:
                os.makedirs(os.path.dirname(path), exist_ok=True)
                try:
                    with open(path) as fh:
                        db = json.load(fh)
                except (OSError, ValueError):
                    db = {}
                db[key] = out
                with open(path, "w") as fh:
                    json.dump(db, fh, indent=1)
            except OSError:
                pass
        return dict(out, cached=False)

    # -- path guiding -----------------------------------------------------------
    GUIDE_LEVELS, GUIDE_FEATURES, GUIDE_WEIGHTS = 16, 2, 48 * 64 + 64 * 64 + 64 * 64 + 64 * 48

    def guide_enable(self, bounds_lo=None, bounds_hi=None, t_log2: int = 19, learning_rate: float = 5e-3,
                     guide_fraction: float = 0.5, seed: int = 0, train_batch: int = 0, record_paths: int = 0,
                     train_steps: int = 0, warmup_steps: int = 0, guide_depth: int = 0):
        """Turns learned path guiding on with fresh parameters; positions
        inside [bounds_lo, bounds_hi] (default: the uploaded scene's box)
        are mapped to the hash grid. Zero arguments take the library's
        defaults (xnh_guide_config)."""
        if bounds_lo is None or bounds_hi is None:
            if self.scene is None:
                raise ValueError("guide_enable needs bounds or an uploaded scene")
            bounds_lo, bounds_hi = self.scene.bounds()
        cfg = N.GuideConfig(int(t_log2), float(learning_rate), float(guide_fraction), int(seed) & 0xFFFFFFFF,
                            (C.c_float * 3)(*bounds_lo), (C.c_float * 3)(*bounds_hi), int(train_batch),
                            int(record_paths), int(train_steps), int(warmup_steps), int(guide_depth))
        self._chk(N.lib().xnh_guide_enable(self.h, C.byref(cfg)))
        self._guide_t = 1 << int(t_log2 or 19)

    def guide_disable(self):
        self._chk(N.lib().xnh_guide_disable(self.h))

    def guide_set_training(self, on: bool):
        """Train after every rendered batch (default) or keep the network frozen."""
        self._chk(N.lib().xnh_guide_set_training(self.h, 1 if on else 0))

    def guide_params(self):
        """FP32 master parameters: (grid (16, T, 2), weights (12288,))."""
        grid = np.empty((self.GUIDE_LEVELS, self._guide_t, self.GUIDE_FEATURES), np.float32)
        w = np.empty(self.GUIDE_WEIGHTS, np.float32)
        self._chk(N.lib().xnh_guide_get_params(self.h, _fp(grid), _fp(w)))
        return grid, w

    def guide_set_params(self, grid: np.ndarray, weights: np.ndarray):
        grid = np.ascontiguousarray(grid, np.float32)
        weights = np.ascontiguousarray(weights, np.float32)
        if grid.size != self.GUIDE_LEVELS * self._guide_t * self.GUIDE_FEATURES or weights.size != self.GUIDE_WEIGHTS:
            raise ValueError("guide parameter sizes do not match the enabled configuration")
        self._chk(N.lib().xnh_guide_set_params(self.h, _fp(grid), _fp(weights)))

    def guide_eval(self, pos: np.ndarray, nrm: np.ndarray) -> np.ndarray:
        """Raw network outputs, (n, 48)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        out = np.empty((pos.shape[0], 48), np.float32)
        self._chk(N.lib().xnh_guide_eval(self.h, pos.shape[0], _fp(pos), _fp(nrm), _fp(out)))
        return out

    def guide_sample(self, pos, nrm, u):
        """One mixture sample per point from uniforms u (n, 3): (dir, pdf)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 3)
        d = np.empty_like(pos)
        pdf = np.empty(pos.shape[0], np.float32)
        self._chk(N.lib().xnh_guide_query(self.h, pos.shape[0], _fp(pos), _fp(nrm), _fp(u), _fp(d), _fp(pdf)))
        return d, pdf

    def guide_pdf(self, pos, nrm, dirs):
        """Mixture pdf (per steradian) of each point at the given directions."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3).copy()
        pdf = np.empty(pos.shape[0], np.float32)
        self._chk(N.lib().xnh_guide_query(self.h, pos.shape[0], _fp(pos), _fp(nrm), None, _fp(d), _fp(pdf)))
        return pdf

    def guide_train(self, pos, nrm, dirs, lam, apply: bool = True, want_grads: bool = False):
        """One training step on a host batch. Returns the mean scaled loss,
        and with want_grads (apply False) the summed gradients
        (grid (16, T, 2), weights (12288,))."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        lam = np.ascontiguousarray(lam, np.float32).reshape(-1)
        loss = C.c_float()
        gg = gw = None
        if want_grads:
            gg = np.empty((self.GUIDE_LEVELS, self._guide_t, self.GUIDE_FEATURES), np.float32)
            gw = np.empty(self.GUIDE_WEIGHTS, np.float32)
        self._chk(N.lib().xnh_guide_train(self.h, pos.shape[0], _fp(pos), _fp(nrm), _fp(dirs), _fp(lam),
                                          1 if apply else 0, C.byref(loss), _fp(gg) if gg is not None else None,
                                          _fp(gw) if gw is not None else None))
        return (loss.value, gg, gw) if want_grads else loss.value

    # -- ray batches ----------------------------------------------------------
    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros(n, HIT_DTYPE)
        self._chk(N.lib().xnh_trace_closest(self.h, rays.ctypes.data_as(C.POINTER(N.Ray)), n,
                                            hits.ctypes.data_as(C.POINTER(N.Hit))))
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint8)
        self._chk(N.lib().xnh_trace_any(self.h, rays.ctypes.data_as(C.POINTER(N.Ray)), n,
                                        occ.ctypes.data_as(C.POINTER(C.c_uint8))))
        return occ
