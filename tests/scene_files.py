"""Writes scene descriptions as .xenon scene files with .mat materials and
OBJ meshes, in the formats load_scene reads (scene/scene_file.cpp:49-240,
material/material.cpp:31-104, geometry/mesh.cpp:42-151). Floats are written
with 9 significant digits, so every float32 reads back to the same bits."""
from __future__ import annotations

import os

import numpy as np


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
    out = ["# written by tests/scene_files.py"]
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


def write_scene(desc, directory: str, name: str = "scene", with_normals: bool = False) -> str:
    """The description as <directory>/<name>.xenon plus materials/*.mat and
    meshes/*.obj; mesh paths are written relative to the scene file's
    directory. Instances other than one identity instance per mesh are not
    expressible in the format (load_scene bakes transforms into vertices)."""
    os.makedirs(os.path.join(directory, "materials"), exist_ok=True)
    os.makedirs(os.path.join(directory, "meshes"), exist_ok=True)
    lines = ["# scene written by tests/scene_files.py", f"config {desc.width} {desc.height} {desc.spp}",
             f"camera {vec(desc.camera.eye)}  {vec(desc.camera.target)}  {g(desc.camera.fov)}"]
    names = []
    for i, m in enumerate(desc.materials):
        nm = f"m{i}"
        names.append(nm)
        write_mat(os.path.join(directory, "materials", f"{nm}.mat"), nm, m.params16())
        lines.append(f"matfile {nm}.mat")
    for i, m in enumerate(desc.meshes):
        rel = os.path.join("meshes", f"mesh{i}.obj")
        nrm = m.normals if with_normals else None
        write_obj(os.path.join(directory, rel), m.positions, m.indices, m.mat_ids, names, normals=nrm)
        lines.append(f"mesh {rel}")
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
