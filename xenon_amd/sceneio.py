"""Scene files out: scene descriptions as .xenon scene files with .mat
materials and OBJ meshes, in the formats load_scene reads
(scene/scene_file.cpp:49-240, material/material.cpp:31-104,
geometry/mesh.cpp:42-151), and stand-ins for the OBJ meshes that the
reference's sample scenes name but do not ship. Reading is
``HipScene.from_file`` (the library's own parser). Floats are written with 9
significant digits, so every float32 reads back to the same bits."""
from __future__ import annotations

import os
import sys

import numpy as np

from . import scenes


def g(x) -> str:
    return f"{float(np.float32(x)):.9g}"


def vec(v) -> str:
    return " ".join(g(x) for x in v)


def write_mat(path: str, name: str, params16, extra_lines=()) -> None:
    p = np.asarray(params16, np.float32)
    lines = [f'name "{name}"', "type principled", f"baseColor {vec(p[0:3])}", f"roughness {g(p[3])}",
             f"metallic {g(p[4])}", f"ior {g(p[5])}", f"transmission {g(p[6])}", f"subsurface {g(p[7])}",
             f"subsurfaceColor {vec(p[8:11])}", f"anisotropy {g(p[11])}", f"emission {vec(p[12:15])}",
             *extra_lines]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_obj(path: str, positions, indices, mat_ids=None, mat_names=None, normals=None, quads=False) -> None:
    """One `v` per position (and `vn` per normal when given, referenced by
    the same index), faces as `f p//n` or `f p`; `usemtl` whenever the
    material changes. quads=True joins consecutive triangle pairs (a b c)
    (a c d) into one quad face, which load_obj fans back into the same two
    triangles."""
    P = np.asarray(positions, np.float32)
    F = np.asarray(indices, np.int64).reshape(-1, 3)
    M = np.zeros(len(F), np.int64) if mat_ids is None else np.asarray(mat_ids, np.int64)
    out = ["# written by xenon_amd.sceneio"]
    out += [f"v {vec(p)}" for p in P]
    if normals is not None:
        out += [f"vn {vec(n)}" for n in np.asarray(normals, np.float32)]

    def corner(i):
        return f"{i + 1}//{i + 1}" if normals is not None else f"{i + 1}"

    cur = None
    t = 0
    while t < len(F):
        if mat_names is not None and M[t] != cur:
            cur = M[t]
            out.append(f"usemtl {mat_names[cur]}")
        a, b, c = F[t]
        if quads and t + 1 < len(F) and M[t + 1] == M[t] and F[t + 1][0] == a and F[t + 1][1] == c:
            out.append(f"f {corner(a)} {corner(b)} {corner(c)} {corner(F[t + 1][2])}")
            t += 2
            continue
        out.append(f"f {corner(a)} {corner(b)} {corner(c)}")
        t += 1
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_scene(desc, directory: str, name: str = "scene", with_normals: bool = False, mesh_root=None) -> str:
    """The description as <directory>/<name>.xenon plus materials/*.mat and
    meshes/*.obj. Mesh paths are opened as written, relative to the working
    directory of the process that loads the file (load_scene does not
    resolve them against the scene file): they are written absolute, or
    relative to ``mesh_root`` when given. Explicit instances are written as
    `instance` lines (this backend's extension of the format)."""
    os.makedirs(os.path.join(directory, "materials"), exist_ok=True)
    os.makedirs(os.path.join(directory, "meshes"), exist_ok=True)
    lines = ["# scene written by xenon_amd.sceneio", f"config {desc.width} {desc.height} {desc.spp}",
             f"camera {vec(desc.camera.eye)}  {vec(desc.camera.target)}  {g(desc.camera.fov)}"]
    names = []
    for i, m in enumerate(desc.materials):
        nm = f"m{i}"
        names.append(nm)
        write_mat(os.path.join(directory, "materials", f"{nm}.mat"), nm, m.params16())
        lines.append(f"matfile {nm}.mat")
    for i, m in enumerate(desc.meshes):
        obj = os.path.abspath(os.path.join(directory, "meshes", f"mesh{i}.obj"))
        nrm = m.normals if with_normals else None
        write_obj(obj, m.positions, m.indices, m.mat_ids, names, normals=nrm)
        lines.append(f"mesh {obj if mesh_root is None else os.path.relpath(obj, mesh_root)}")
    for ins in desc.instances:
        lines.append(f"instance {ins.mesh} {vec(ins.translation)} {g(ins.scale)} {vec(ins.rot_deg)} "
                     f"{ins.instance_id}")
    for lt in desc.lights:
        lines.append(f"light {lt.mesh} {lt.tri} {vec(lt.emission)} {g(lt.area)}")
    for s in desc.suns:
        lines.append(f"sunlight {vec(s.direction)} {vec(s.color)} {g(s.intensity)}")
    sk = desc.sky
    lines.append(f"sky {vec(sk.zenith)} {vec(sk.horizon)} {vec(sk.ground)} {g(sk.intensity)} {g(sk.sharpness)}")
    path = os.path.join(directory, f"{name}.xenon")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


# ---------------------------------------------------------------------------
# stand-ins for the sample scenes' missing meshes
# ---------------------------------------------------------------------------

def reference_box(blocks: bool):
    """The box the sample scenes' cameras frame (eye (0, 5, 15) looking at
    (0, 5, 0), fov 40): room x, z in [-5, 5], y in [0, 10], open towards +z,
    walls facing inward, usemtl groups floor, ceiling, back, left, right
    [white] light. The light quad under the ceiling comes last, so that its
    triangles are 10 and 11 (empty box) or 34 and 35 (with the tall and short
    blocks), the indices the scene files' `light` lines name. Returns
    (positions, indices, mat_ids, names)."""
    c = scenes.cornell()
    m = c.meshes[0]
    S = 5.55
    P = m.positions.astype(np.float64)
    # the Cornell description's room [0, S] x [0, S] x [-S, 0] onto the frame above
    Q = np.stack([P[:, 0] * (10.0 / S) - 5.0, P[:, 1] * (10.0 / S), (P[:, 2] + S) * (10.0 / S) - 5.0], 1)
    T = m.indices
    keep = np.arange(len(T)) if blocks else np.r_[0:10, len(T) - 2:len(T)]
    # Cornell materials: 0 white, 1 red, 2 green, 3 light; walls in the order
    # floor, ceiling, back, left, right
    names = ["floor", "ceiling", "back", "left", "right", "white", "light"]
    wall = [0, 0, 1, 1, 2, 2, 3, 3, 4, 4]
    mats = []
    for t in keep:
        if t < 10:
            mats.append(wall[t])
        elif m.mat_ids[t] == 3:
            mats.append(6)
        else:
            mats.append(5)
    used = sorted(set(np.asarray(T[keep]).reshape(-1).tolist()))
    remap = {v: i for i, v in enumerate(used)}
    idx = np.array([[remap[v] for v in T[t]] for t in keep], np.int32)
    return Q[used].astype(np.float32), idx, np.array(mats, np.int32), names


def blob(radius: float, freq: int = 24, seed: int = 3, amplitude: float = 0.12):
    """A displaced geodesic sphere of the given radius (stand-in for a
    scanned model)."""
    V, T = scenes.geodesic_sphere(freq)
    V = scenes.displaced(V, seed, amplitude)
    return (V * radius).astype(np.float32), T.astype(np.int32)


def write_reference_meshes(root: str) -> list:
    """Writes, under ``root`` (the directory the sample scenes are run from),
    the OBJ files their `mesh` lines name: scenes/cornell_box/meshes/
    cornell_box.obj and empty_box.obj (the box above), and stand-ins for the
    scanned models, scenes/cow/meshes/cow.obj and scenes/lucy/meshes/lucy.obj
    (displaced spheres sized so that the scene files' transforms put them
    inside the box). Returns the paths written."""
    out = []
    for name, blocks in (("cornell_box", True), ("empty_box", False)):
        P, T, M, names = reference_box(blocks)
        path = os.path.join(root, "scenes", "cornell_box", "meshes", f"{name}.obj")
        os.makedirs(os.path.dirname(path), exist_ok=True)
        write_obj(path, P, T, M, names, quads=True)
        out.append(path)
    # cow.xenon: scale 1.1 at (0, 3.8, 0); lucy.xenon: scale 0.0065 at (3, 3, 0)
    for scene, mat, radius in (("cow", "glossy_metal", 2.0), ("lucy", "glass", 300.0)):
        P, T = blob(radius)
        path = os.path.join(root, "scenes", scene, "meshes", f"{scene}.obj")
        os.makedirs(os.path.dirname(path), exist_ok=True)
        write_obj(path, P, T, np.zeros(len(T), np.int32), [mat])
        out.append(path)
    return out


if __name__ == "__main__":
    for p in write_reference_meshes(sys.argv[1] if len(sys.argv) > 1 else "."):
        print(p)
