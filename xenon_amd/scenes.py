"""Seeded synthetic scene generators for the benchmark configurations.

The reference ships no meshes for its sample scenes, so every workload is
built here from a seed. A ``SceneDesc`` is plain numpy data in the
reference's scene model (scene/scene.h:52-161): triangle meshes with
per-triangle material ids, principled materials (material/material.h:24-37),
area lights named by (mesh index, triangle index) with a caller-given area
(scene/scene.h:19-24), directional lights given toward the sun
(scene/scene.h:43-49), a gradient sky (scene/environment.h:26-51), a
look-at camera (camera/camera.h:15-32) and optional TRS instances
(geometry/transform.h:105-148). With no instances every mesh gets an
identity instance whose id is its mesh index (scene/scene.h:87-99).

All consumers (the HIP backend's scene builder and the CPU checkers) receive
the same arrays, so vertex data never has to be regenerated bit-identically
by two implementations.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


@dataclass
class Material:
    base: tuple = (0.5, 0.5, 0.5)
    roughness: float = 0.5
    metallic: float = 0.0
    ior: float = 1.5
    transmission: float = 0.0
    subsurface: float = 0.0
    ss_color: tuple = (1.0, 1.0, 1.0)
    anisotropy: float = 0.0
    emission: tuple = (0.0, 0.0, 0.0)

    def params16(self) -> np.ndarray:
        """Authoring parameters in the 16-float order of the C ABIs."""
        return np.array([*self.base, self.roughness, self.metallic, self.ior, self.transmission,
                         self.subsurface, *self.ss_color, self.anisotropy, *self.emission, 0.0], dtype=F32)


@dataclass
class Mesh:
    positions: np.ndarray            # (V, 3) float32
    indices: np.ndarray              # (T, 3) int32
    mat_ids: np.ndarray              # (T,) int32
    normals: np.ndarray | None = None  # (V, 3) float32; None -> area-weighted smooth normals

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=F32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.int32).reshape(-1, 3)
        self.mat_ids = np.ascontiguousarray(self.mat_ids, dtype=np.int32).reshape(-1)
        if self.normals is not None:
            self.normals = np.ascontiguousarray(self.normals, dtype=F32).reshape(-1, 3)

    @property
    def num_tris(self) -> int:
        return int(self.indices.shape[0])


@dataclass
class Light:
    mesh: int
    tri: int
    emission: tuple
    area: float


@dataclass
class Sun:
    direction: tuple   # toward the sun; normalised by the consumer
    color: tuple
    intensity: float


@dataclass
class Sky:
    zenith: tuple = (0.10, 0.30, 0.80)
    horizon: tuple = (0.55, 0.70, 0.85)
    ground: tuple = (0.10, 0.08, 0.05)
    intensity: float = 1.0
    sharpness: float = 3.0


@dataclass
class Instance:
    mesh: int
    translation: tuple
    scale: float
    rot_deg: tuple
    instance_id: int


@dataclass
class Prim:
    """Analytic primitive (geometry/primitives.h:54-318); params by kind:
    sphere (cx, cy, cz, r), plane (nx, ny, nz, offset), box (cx, cy, cz, hx,
    hy, hz), disk (cx, cy, cz, nx, ny, nz, r), quad (ox, oy, oz, ux, uy, uz,
    vx, vy, vz). Plane and disk normals are used as given."""
    kind: str
    params: tuple
    mat: int

    KINDS = ("sphere", "plane", "box", "disk", "quad")

    @property
    def type_id(self) -> int:
        return self.KINDS.index(self.kind)


@dataclass
class Camera:
    eye: tuple
    target: tuple
    up: tuple = (0.0, 1.0, 0.0)
    fov: float = 40.0


@dataclass
class SceneDesc:
    name: str
    meshes: list
    materials: list
    lights: list
    camera: Camera
    width: int
    height: int
    spp: int
    min_bounces: int = 3
    max_bounces: int = 8
    suns: list = field(default_factory=list)
    sky: Sky = field(default_factory=Sky)
    instances: list = field(default_factory=list)
    prims: list = field(default_factory=list)

    @property
    def aspect(self) -> float:
        return float(F32(self.width) / F32(self.height))

    @property
    def num_tris(self) -> int:
        return sum(m.num_tris for m in self.meshes)


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------

def _quad(verts, tris, p0, p1, p2, p3, facing):
    """Appends the planar quad p0-p1-p2-p3 (corners in cyclic order) as two
    triangles with their own four vertices (flat shading), wound so that the
    geometric normal cross(p1 - p0, p2 - p0) points along `facing`. The
    reference never face-forwards normals and takes NEE only on the normal's
    side (render/cpu_tiled.cpp:117-121), so orientation matters."""
    P = [np.asarray(p, np.float64) for p in (p0, p1, p2, p3)]
    n = np.cross(P[1] - P[0], P[2] - P[0])
    if float(np.dot(n, facing)) < 0.0:
        P = [P[0], P[3], P[2], P[1]]
    base = len(verts)
    verts.extend([tuple(p) for p in P])
    tris.append((base, base + 1, base + 2))
    tris.append((base, base + 2, base + 3))


def _box(verts, tris, lo, hi, rot_y_deg=0.0, center=None):
    """Closed box with outward-facing quads, rotated about the vertical axis
    through `center`."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
         (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    if rot_y_deg:
        cx, cz = center
        a = math.radians(rot_y_deg)
        ca, sa = math.cos(a), math.sin(a)
        c = [(cx + (x - cx) * ca + (z - cz) * sa, y, cz - (x - cx) * sa + (z - cz) * ca) for x, y, z in c]
    C = np.asarray(c, np.float64)
    mid = C.mean(0)
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2)]
    for f in faces:
        fc = C[list(f)].mean(0)
        _quad(verts, tris, *[c[i] for i in f], facing=fc - mid)


def _tri_area(p0, p1, p2) -> float:
    p0, p1, p2 = (np.asarray(p, dtype=np.float64) for p in (p0, p1, p2))
    return float(0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0)))


# ---------------------------------------------------------------------------
# config 0: Cornell box
# ---------------------------------------------------------------------------

def cornell(width=512, height=512, spp=64, max_bounces=8, min_bounces=3) -> SceneDesc:
    """Cornell box: five wall quads (white 0.73, red 0.65/0.05/0.05, green
    0.12/0.45/0.15), a tall and a short box, an emissive ceiling quad with
    radiance (17, 12, 4). One mesh, walls facing inward; the two light
    triangles are area lights with their true areas. "Diffuse" is the
    principled model with metallic 0, roughness 1 and ior 1 (F0 = 0, so the
    specular lobe has zero selection weight)."""
    S = 5.55
    verts, tris, mats = [], [], []

    def add(mat, fn, *a, **k):
        n0 = len(tris)
        fn(verts, tris, *a, **k)
        mats.extend([mat] * (len(tris) - n0))

    WHITE, RED, GREEN, LIGHT = 0, 1, 2, 3
    add(WHITE, _quad, (0, 0, 0), (0, 0, -S), (S, 0, -S), (S, 0, 0), facing=(0, 1, 0))     # floor
    add(WHITE, _quad, (0, S, 0), (S, S, 0), (S, S, -S), (0, S, -S), facing=(0, -1, 0))    # ceiling
    add(WHITE, _quad, (0, 0, -S), (0, S, -S), (S, S, -S), (S, 0, -S), facing=(0, 0, 1))   # back
    add(RED, _quad, (0, 0, 0), (0, S, 0), (0, S, -S), (0, 0, -S), facing=(1, 0, 0))       # left
    add(GREEN, _quad, (S, 0, 0), (S, 0, -S), (S, S, -S), (S, S, 0), facing=(-1, 0, 0))    # right
    add(WHITE, _box, (1.30, 0.0, -4.72), (2.95, 3.30, -3.07), rot_y_deg=15.0, center=(2.125, -3.895))
    add(WHITE, _box, (3.10, 0.0, -2.50), (4.75, 1.65, -0.85), rot_y_deg=-18.0, center=(3.925, -1.675))
    ly = S - 0.01
    lq = [(2.13, ly, -2.27), (3.43, ly, -2.27), (3.43, ly, -3.32), (2.13, ly, -3.32)]
    add(LIGHT, _quad, *lq, facing=(0, -1, 0))
    nt = len(tris)
    mesh = Mesh(np.array(verts, dtype=F32), np.array(tris, dtype=np.int32), np.array(mats, dtype=np.int32))
    materials = [
        Material(base=(0.73, 0.73, 0.73), roughness=1.0, ior=1.0),
        Material(base=(0.65, 0.05, 0.05), roughness=1.0, ior=1.0),
        Material(base=(0.12, 0.45, 0.15), roughness=1.0, ior=1.0),
        Material(base=(0.78, 0.78, 0.78), roughness=1.0, ior=1.0, emission=(17.0, 12.0, 4.0)),
    ]
    P = mesh.positions
    lights = []
    for t in (nt - 2, nt - 1):
        i0, i1, i2 = mesh.indices[t]
        lights.append(Light(0, t, (17.0, 12.0, 4.0), _tri_area(P[i0], P[i1], P[i2])))
    cam = Camera(eye=(S / 2, S / 2, 8.0), target=(S / 2, S / 2, 0.0), fov=37.5)
    return SceneDesc("cornell", [mesh], materials, lights, cam, width, height, spp,
                     min_bounces=min_bounces, max_bounces=max_bounces)


# ---------------------------------------------------------------------------
# displaced spheres
# ---------------------------------------------------------------------------

def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    return v, f


def icosphere(level: int):
    """Unit icosphere by recursive 4:1 midpoint subdivision:
    20*4^level triangles, 10*4^level + 2 vertices (float64)."""
    v, f = _icosahedron()
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        m = inv.reshape(3, -1) + nv
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def geodesic_sphere(freq: int):
    """Unit class-I geodesic sphere: each icosahedron face split into freq^2
    triangles (20*freq^2 in total), shared vertices merged."""
    v0, f0 = _icosahedron()
    pts, tris = [], []
    n_face = (freq + 1) * (freq + 2) // 2
    ii, jj = [], []
    for i in range(freq + 1):
        for j in range(freq + 1 - i):
            ii.append(i)
            jj.append(j)
    ii, jj = np.array(ii), np.array(jj)
    kk = freq - ii - jj
    lid = -np.ones((freq + 1, freq + 1), dtype=np.int64)
    lid[ii, jj] = np.arange(n_face)
    loc = []
    for i in range(freq):
        for j in range(freq - i):
            loc.append((lid[i, j], lid[i + 1, j], lid[i, j + 1]))
            if j < freq - i - 1:
                loc.append((lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]))
    loc = np.array(loc, dtype=np.int64)
    for fi, (a, b, c) in enumerate(f0):
        p = (ii[:, None] * v0[b] + jj[:, None] * v0[c] + kk[:, None] * v0[a]) / freq
        pts.append(p / np.linalg.norm(p, axis=1, keepdims=True))
        tris.append(loc + fi * n_face)
    P = np.concatenate(pts)
    T = np.concatenate(tris)
    q = np.round(P * (1 << 22)).astype(np.int64)
    _, first, inv = np.unique(q, axis=0, return_index=True, return_inverse=True)
    V = P[first]
    T = inv.reshape(-1)[T]
    # keep orientation outward
    a, b, c = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return V, T


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15))
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / float(1 << 53)) * 2.0 - 1.0


def value_noise(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0) -> np.ndarray:
    """Fractal 3D value noise in [-1, 1]: lattice values from an integer hash,
    smoothstep-weighted trilinear interpolation, frequency x2 and amplitude
    x0.5 per octave, normalised by the amplitude sum."""
    out = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    with np.errstate(over="ignore"):
        for o in range(octaves):
            q = p * freq + 17.0 * (o + 1)
            i0 = np.floor(q).astype(np.int64)
            f = q - i0
            w = f * f * (3.0 - 2.0 * f)
            acc = np.zeros(p.shape[0])
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        val = _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed * 131 + o)
                        wt = ((w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1])
                              * (w[:, 2] if dz else 1 - w[:, 2]))
                        acc += wt * val
            out += amp * acc
            norm += amp
            amp *= 0.5
            freq *= 2.0
    return out / norm


def displaced(V: np.ndarray, seed: int, amplitude: float = 0.05) -> np.ndarray:
    """Radial displacement r = 1 + amplitude * noise(p) of unit-sphere points."""
    n = value_noise(V, seed)
    return V * (1.0 + amplitude * n)[:, None]


# ---------------------------------------------------------------------------
# config 1: one large displaced icosphere
# ---------------------------------------------------------------------------

def big_mesh(width=1920, height=1080, spp=256, level=8, max_bounces=8, min_bounces=3, seed=1) -> SceneDesc:
    """Icosphere L`level` (20*4^level triangles; L8 = 1,310,720) radially
    displaced by 4-octave value noise (amplitude 5%), GGX conductor
    roughness 0.3, on a diffuse ground quad; sun (3, 3, 2.8) and the
    gradient sky. The configuration's sun direction (1, -1, 0.5) is the
    direction sunlight travels; the reference's sunlight takes the direction
    toward the sun, so the scene passes (-1, 1, -0.5)."""
    V, F = icosphere(level)
    P = displaced(V, seed)
    sphere = Mesh(P.astype(F32), F.astype(np.int32), np.zeros(F.shape[0], np.int32))
    g = 12.0
    gy = -1.06
    ground = Mesh(np.array([(-g, gy, g), (g, gy, g), (g, gy, -g), (-g, gy, -g)], F32),
                  np.array([(0, 1, 2), (0, 2, 3)], np.int32), np.array([1, 1], np.int32))
    materials = [Material(base=(0.95, 0.64, 0.54), roughness=0.3, metallic=1.0),
                 Material(base=(0.6, 0.6, 0.6), roughness=1.0, ior=1.0)]
    cam = Camera(eye=(0.0, 0.55, 3.6), target=(0.0, -0.1, 0.0), fov=40.0)
    return SceneDesc("big_mesh", [sphere, ground], materials, [], cam, width, height, spp,
                     min_bounces=min_bounces, max_bounces=max_bounces,
                     suns=[Sun((-1.0, 1.0, -0.5), (3.0, 3.0, 2.8), 1.0)], sky=Sky())


# ---------------------------------------------------------------------------
# config 2: instancing
# ---------------------------------------------------------------------------

def _random_rotation_deg(rng):
    """Uniform random rotation (Shoemake quaternion) expressed as the
    reference's Euler angles: M = Rz * Rx * Ry (geometry/transform.h:105-122)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    M = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    # M = Rz(c) Rx(a) Ry(b): M[2][1] = sin(a); M[2][0] = -cos(a) sin(b); M[2][2] = cos(a) cos(b);
    # M[0][1] = -sin(c) cos(a); M[1][1] = cos(c) cos(a)
    a = math.asin(max(-1.0, min(1.0, M[2][1])))
    b = math.atan2(-M[2][0], M[2][2])
    c = math.atan2(-M[0][1], M[1][1])
    return (math.degrees(a), math.degrees(b), math.degrees(c))


def instancing(width=1920, height=1080, spp=64, max_bounces=10, min_bounces=3, n_inst=4096, n_blas=10,
               tri_range=(20000, 200000), n_emit=64, box=100.0) -> SceneDesc:
    """TLAS over `n_inst` instances of `n_blas` displaced geodesic spheres.

    BLAS triangle counts are log-uniform in `tri_range` (seed 2; geodesic
    frequency round(sqrt(n/20))). Instance transforms (seed 3): uniform
    random rotation, scale U[0.5, 2], position U[0, box]^3. Materials
    (seed 4), one per BLAS: 40% diffuse, 30% GGX conductor (roughness
    U[0.05, 0.6]), 30% rough dielectric (ior 1.5). `n_emit` emissive
    triangles (seed 5) form one extra mesh with an identity instance whose id
    equals its mesh index, as the reference's light lookup requires
    (render/cpu_tiled.cpp:63-65, 93-106); the instanced spheres take ids
    above it."""
    rng2 = np.random.default_rng(2)
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    counts = np.exp(rng2.uniform(lo, hi, n_blas))
    meshes = []
    kinds = ["diffuse"] * int(round(0.4 * n_blas)) + ["ggx"] * int(round(0.3 * n_blas))
    kinds += ["glass"] * (n_blas - len(kinds))
    rng4 = np.random.default_rng(4)
    materials = []
    for b in range(n_blas):
        freq = max(2, int(round(math.sqrt(counts[b] / 20.0))))
        V, F = geodesic_sphere(freq)
        P = displaced(V, seed=100 + b)
        kind = kinds[b]
        col = tuple(float(x) for x in rng4.uniform(0.2, 0.9, 3))
        if kind == "diffuse":
            materials.append(Material(base=col, roughness=1.0, ior=1.0))
        elif kind == "ggx":
            materials.append(Material(base=col, roughness=float(rng4.uniform(0.05, 0.6)), metallic=1.0))
        else:
            materials.append(Material(base=(1.0, 1.0, 1.0), roughness=0.1, ior=1.5, transmission=1.0))
        meshes.append(Mesh(P.astype(F32), F.astype(np.int32), np.full(F.shape[0], b, np.int32)))
    # emitters
    rng5 = np.random.default_rng(5)
    ev, et = [], []
    for k in range(n_emit):
        c = rng5.uniform(0.0, box, 3)
        nrm = rng5.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        t1 = np.cross(nrm, [0.0, 1.0, 0.0] if abs(nrm[1]) < 0.9 else [1.0, 0.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(nrm, t1)
        s = 3.0
        p0, p1, p2 = c + s * t1, c + s * (-0.5 * t1 + 0.866 * t2), c + s * (-0.5 * t1 - 0.866 * t2)
        ev.extend([p0, p1, p2])
        et.append((3 * k, 3 * k + 1, 3 * k + 2))
    emit_mat = len(materials)
    materials.append(Material(base=(0.8, 0.8, 0.8), roughness=1.0, ior=1.0, emission=(1.0, 1.0, 1.0)))
    emesh = Mesh(np.array(ev, F32), np.array(et, np.int32), np.full(n_emit, emit_mat, np.int32))
    emesh_idx = len(meshes)
    meshes.append(emesh)
    lights = []
    P = emesh.positions
    for k in range(n_emit):
        e = tuple(float(x) for x in rng5.uniform(10.0, 40.0, 3))
        lights.append(Light(emesh_idx, k, e, _tri_area(P[3 * k], P[3 * k + 1], P[3 * k + 2])))
    rng3 = np.random.default_rng(3)
    instances = [Instance(emesh_idx, (0.0, 0.0, 0.0), 1.0, (0.0, 0.0, 0.0), emesh_idx)]
    which = rng3.integers(0, n_blas, n_inst)
    for i in range(n_inst):
        rot = _random_rotation_deg(rng3)
        s = float(rng3.uniform(0.5, 2.0))
        pos = tuple(float(x) for x in rng3.uniform(0.0, box, 3))
        instances.append(Instance(int(which[i]), pos, s, rot, emesh_idx + 1 + i))
    cam = Camera(eye=(box * 0.5, box * 0.62, -box * 0.55), target=(box * 0.5, box * 0.45, box * 0.5), fov=55.0)
    return SceneDesc("instancing", meshes, materials, lights, cam, width, height, spp,
                     min_bounces=min_bounces, max_bounces=max_bounces, instances=instances)


CONFIGS = {0: cornell, 1: big_mesh, 2: instancing}


# ---------------------------------------------------------------------------
# config 4: caustics through a wall opening
# ---------------------------------------------------------------------------

def caustic_room(width=1280, height=720, spp=64, freq=158, max_bounces=8, min_bounces=3, seed=7,
                 emission=(400.0, 370.0, 320.0)) -> SceneDesc:
    """Closed diffuse room (10 x 6 x 10) lit only by a 0.5 x 0.5 emissive quad
    at the back of a small alcove behind a 1 x 1 opening in the +x wall; a
    displaced geodesic sphere of 20*freq^2 triangles (freq 158: 499,280) of
    delta glass (ior 1.5, roughness 0.02, alpha 4e-4 < 1e-3) stands in the
    room. The sky is off, so every photon enters through the opening."""
    verts, tris, mats = [], [], []

    def add(mat, fn, *a, **k):
        n0 = len(tris)
        fn(verts, tris, *a, **k)
        mats.extend([mat] * (len(tris) - n0))

    X, Y, Z = 5.0, 6.0, 5.0
    WALL, FLOOR, LIGHTM = 0, 1, 2
    inward = {"+y": (0, 1, 0), "-y": (0, -1, 0), "+z": (0, 0, 1), "-z": (0, 0, -1), "+x": (1, 0, 0),
              "-x": (-1, 0, 0)}
    add(FLOOR, _quad, (-X, 0, Z), (X, 0, Z), (X, 0, -Z), (-X, 0, -Z), facing=inward["+y"])    # floor
    add(WALL, _quad, (-X, Y, Z), (-X, Y, -Z), (X, Y, -Z), (X, Y, Z), facing=inward["-y"])     # ceiling
    add(WALL, _quad, (-X, 0, -Z), (X, 0, -Z), (X, Y, -Z), (-X, Y, -Z), facing=inward["+z"])   # back
    add(WALL, _quad, (-X, 0, Z), (-X, Y, Z), (X, Y, Z), (X, 0, Z), facing=inward["-z"])       # front
    add(WALL, _quad, (-X, 0, Z), (-X, 0, -Z), (-X, Y, -Z), (-X, Y, Z), facing=inward["+x"])   # left
    # right wall (x = +X, facing -x) with a 1 x 1 opening at y in [3.5, 4.5], z in [-0.5, 0.5]
    y0, y1, z0, z1 = 3.5, 4.5, -0.5, 0.5
    add(WALL, _quad, (X, 0, Z), (X, y0, Z), (X, y0, -Z), (X, 0, -Z), facing=inward["-x"])
    add(WALL, _quad, (X, y1, Z), (X, Y, Z), (X, Y, -Z), (X, y1, -Z), facing=inward["-x"])
    add(WALL, _quad, (X, y0, Z), (X, y1, Z), (X, y1, z1), (X, y0, z1), facing=inward["-x"])
    add(WALL, _quad, (X, y0, z0), (X, y1, z0), (X, y1, -Z), (X, y0, -Z), facing=inward["-x"])
    # alcove behind the opening (depth 0.8), facing inward
    D = X + 0.8
    add(WALL, _quad, (X, y0, z1), (D, y0, z1), (D, y0, z0), (X, y0, z0), facing=inward["+y"])
    add(WALL, _quad, (X, y1, z1), (X, y1, z0), (D, y1, z0), (D, y1, z1), facing=inward["-y"])
    add(WALL, _quad, (X, y0, z0), (D, y0, z0), (D, y1, z0), (X, y1, z0), facing=inward["+z"])
    add(WALL, _quad, (X, y0, z1), (X, y1, z1), (D, y1, z1), (D, y0, z1), facing=inward["-z"])
    add(WALL, _quad, (D, y0, z1), (D, y1, z1), (D, y1, z0), (D, y0, z0), facing=inward["-x"])
    # emitter: 0.5 x 0.5, just in front of the alcove back, facing -x
    e = D - 0.01
    lq = [(e, 3.75, 0.25), (e, 4.25, 0.25), (e, 4.25, -0.25), (e, 3.75, -0.25)]
    add(LIGHTM, _quad, *lq, facing=inward["-x"])
    nt = len(tris)
    room = Mesh(np.array(verts, F32), np.array(tris, np.int32), np.array(mats, np.int32))
    V, F = geodesic_sphere(freq)
    P = displaced(V, seed=seed, amplitude=0.05) * 1.3 + np.array([0.3, 1.45, -0.2])
    glass = Mesh(P.astype(F32), F.astype(np.int32), np.full(F.shape[0], 3, np.int32))
    materials = [Material(base=(0.75, 0.75, 0.75), roughness=1.0, ior=1.0),
                 Material(base=(0.60, 0.55, 0.50), roughness=1.0, ior=1.0),
                 Material(base=(0.8, 0.8, 0.8), roughness=1.0, ior=1.0, emission=emission),
                 Material(base=(1.0, 1.0, 1.0), roughness=0.02, ior=1.5, transmission=1.0)]
    Pr = room.positions
    lights = []
    for t in (nt - 2, nt - 1):
        i0, i1, i2 = room.indices[t]
        lights.append(Light(0, t, tuple(emission), _tri_area(Pr[i0], Pr[i1], Pr[i2])))
    cam = Camera(eye=(-4.2, 3.2, 4.2), target=(0.4, 1.3, -0.3), fov=58.0)
    sky = Sky(intensity=0.0)
    return SceneDesc("caustic_room", [room, glass], materials, lights, cam, width, height, spp,
                     min_bounces=min_bounces, max_bounces=max_bounces, sky=sky)


# ---------------------------------------------------------------------------
# analytic primitives beside meshes
# ---------------------------------------------------------------------------

def prim_shapes(width=800, height=450, spp=64, n_spheres=24, max_bounces=8, min_bounces=3, seed=5) -> SceneDesc:
    """The five primitive kinds of the reference's prim_types scene (ground
    plane, sphere, box, disk, tilted quad, thin metal box) plus a glass
    sphere and `n_spheres` small random spheres resting on the ground, next
    to two meshes: an emissive quad overhead (two area-light triangles,
    facing down) and a glossy icosphere. Sun and sky are on, so rays end on
    primitives, meshes, the light and the sky."""
    rng = np.random.default_rng(seed)
    PLANE, RED, BLUE, YELLOW, GREEN, METAL, GLASS, LIGHTM, GLOSSY = range(9)
    materials = [Material(base=(0.72, 0.70, 0.68), roughness=0.8),
                 Material(base=(0.85, 0.12, 0.08), roughness=0.3),
                 Material(base=(0.10, 0.45, 0.80), roughness=0.5),
                 Material(base=(0.90, 0.75, 0.10), roughness=0.4),
                 Material(base=(0.15, 0.80, 0.25), roughness=0.5),
                 Material(base=(0.90, 0.90, 0.90), roughness=0.05, metallic=1.0),
                 Material(base=(1.0, 1.0, 1.0), roughness=0.0, ior=1.5, transmission=1.0),
                 Material(base=(0.8, 0.8, 0.8), roughness=1.0, ior=1.0, emission=(14.0, 12.0, 9.0)),
                 Material(base=(0.95, 0.64, 0.54), roughness=0.25, metallic=1.0)]
    dn = np.array([0.0, 0.5, 1.0], F32)
    dn = dn / np.sqrt(np.float32(dn @ dn))
    prims = [Prim("plane", (0.0, 1.0, 0.0, 0.0), PLANE),
             Prim("sphere", (-7.0, 1.4, 0.0, 1.4), RED),
             Prim("box", (-3.5, 1.2, 0.0, 1.2, 1.2, 1.2), BLUE),
             Prim("disk", (0.0, 1.8, 0.5, *map(float, dn), 1.3), YELLOW),
             Prim("quad", (2.8, 0.0, -0.8, 2.4, 0.0, 0.0, 0.0, 2.8, -0.8), GREEN),
             Prim("box", (7.0, 0.5, 0.0, 1.6, 0.5, 1.0), METAL),
             Prim("sphere", (-1.6, 0.9, 3.2, 0.9), GLASS)]
    for _ in range(n_spheres):
        r = float(rng.uniform(0.15, 0.45))
        x, z = float(rng.uniform(-9.0, 9.0)), float(rng.uniform(-5.0, 6.0))
        prims.append(Prim("sphere", (x, r, z, r), int(rng.choice([RED, BLUE, YELLOW, GREEN, METAL, GLASS]))))
    verts, tris = [], []
    _quad(verts, tris, (-1.5, 7.0, -1.0), (1.5, 7.0, -1.0), (1.5, 7.0, 1.0), (-1.5, 7.0, 1.0), facing=(0, -1, 0))
    lamp = Mesh(np.array(verts, F32), np.array(tris, np.int32), np.full(len(tris), LIGHTM, np.int32))
    V, F = icosphere(2)
    ball = Mesh((V * 0.9 + np.array([4.6, 0.9, 2.6])).astype(F32), F.astype(np.int32),
                np.full(F.shape[0], GLOSSY, np.int32))
    P = lamp.positions
    lights = [Light(0, t, (14.0, 12.0, 9.0), _tri_area(*P[lamp.indices[t]])) for t in range(2)]
    cam = Camera(eye=(0.0, 5.0, 20.0), target=(0.0, 2.0, 0.0), fov=50.0)
    return SceneDesc("prim_shapes", [lamp, ball], materials, lights, cam, width, height, spp,
                     min_bounces=min_bounces, max_bounces=max_bounces,
                     suns=[Sun((-0.4, 1.0, 0.3), (2.0, 1.9, 1.7), 1.0)],
                     sky=Sky((0.15, 0.25, 0.55), (0.65, 0.72, 0.80), (0.15, 0.12, 0.10), 0.8, 3.0), prims=prims)


# ---------------------------------------------------------------------------
# config 3: batched jobs
# ---------------------------------------------------------------------------

@dataclass
class Job:
    kind: int            # generator: 0 cornell, 1 big mesh, 2 instancing
    width: int
    height: int
    spp: int
    max_bounces: int
    seed: int


def job_list(n=32, seed=6, res_range=(256, 2048), spp_choices=(16, 32, 64, 128, 256, 512, 1024),
             depth_range=(4, 12)) -> list:
    """`n` render jobs drawn by seed: generator 0-2, width and height uniform
    in `res_range` (multiples of 16), spp from `spp_choices`, max depth
    uniform in `depth_range`."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        w, h = (int(rng.integers(res_range[0] // 16, res_range[1] // 16 + 1)) * 16 for _ in range(2))
        jobs.append(Job(int(rng.integers(0, 3)), w, h, int(rng.choice(spp_choices)),
                        int(rng.integers(depth_range[0], depth_range[1] + 1)), int(rng.integers(0, 1 << 30))))
    return jobs


def job_scene(job: Job, small: bool = False) -> SceneDesc:
    """The scene of a job; `small` shrinks the geometry (for tests)."""
    if job.kind == 0:
        d = cornell(job.width, job.height, job.spp, max_bounces=job.max_bounces)
    elif job.kind == 1:
        d = big_mesh(job.width, job.height, job.spp, level=4 if small else 6, max_bounces=job.max_bounces)
    else:
        d = instancing(job.width, job.height, job.spp, max_bounces=job.max_bounces,
                       n_inst=64 if small else 512, tri_range=(500, 2000) if small else (2000, 20000),
                       n_emit=16)
    return d


CONFIGS[4] = caustic_room
