"""Scene files, CPU only: xnh_scene_load_file reads .xenon scene files, .mat
materials and OBJ meshes into the scene the reference's load_scene
(scene/scene_file.cpp:49-240) builds from the same files, bit for bit:
render settings, camera, prepared materials, mesh vertices, normals,
triangles and material ids, BLAS and TLAS, area and directional lights and
the sky. The reference side is oracle/_ref (the reference's own parser,
compiled from its sources); those tests skip where it is not built. Its
shipped scene files are read where they lie; the OBJ meshes they name are
not shipped, so the mesh-bearing scenes run with the stand-ins that
xenon_amd.sceneio writes at the paths the files name."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

from xenon_amd import sceneio as SF
from xenon_amd import HipScene, scenes
from xenon_amd import native as N

REF_SCENES = "/root/reference/scenes"


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32)


def _flat_mesh(hs, i):
    mv = hs.flat.meshes[i]
    nv, nt = mv.num_vertices, mv.num_triangles
    pos = np.stack([np.ctypeslib.as_array(p, shape=(nv,)) for p in (mv.vx, mv.vy, mv.vz)], 1) if nv else \
        np.zeros((0, 3), np.float32)
    nrm = hs.normals(i) if nv else np.zeros((0, 3), np.float32)
    idx = np.ctypeslib.as_array(mv.indices, shape=(3 * nt,)).reshape(nt, 3) if nt else np.zeros((0, 3), np.int32)
    mat = np.ctypeslib.as_array(mv.mat_ids, shape=(nt,)) if nt else np.zeros(0, np.int32)
    return pos, nrm, idx, mat


def _cam_array(c):
    return np.array([*c.eye, *c.lower_left, *c.horizontal, *c.vertical], np.float32)


def assert_same_scene(hs, rs):
    """hs: HipScene.from_file; rs: pyoracle.RefFileScene of the same file."""
    info = hs.info
    assert [info.width, info.height, info.samples, info.min_bounces, info.max_bounces] == list(rs.config)
    assert np.array_equal(_bits(_cam_array(hs.file_camera)), _bits(rs.camera()))
    assert np.array_equal(_bits(_cam_array(hs.camera())), _bits(rs.camera()))
    f = hs.flat
    assert f.num_materials == rs.num_materials
    for i in range(f.num_materials):
        m = f.materials[i]
        p, d = rs.material_params(i), rs.material_derived(i)
        ours = [*m.base_color, m.metallic, m.ior, m.subsurface, *m.subsurface_color]
        theirs = [*p[0:3], p[4], p[5], p[7], *p[8:11]]
        assert np.array_equal(_bits(ours), _bits(theirs)), i
        assert np.array_equal(_bits([m.alpha, m.alpha_x, m.alpha_y, *m.f0]), _bits(d[0:6])), i
        flags = [bool(m.flags & b) for b in (N_CONDUCTOR, N_TRANSMISSIVE, N_SUBSURFACE, N_DELTA, N_EMISSIVE)]
        assert flags == [bool(x) for x in d[6:11]], i
    assert f.num_meshes == rs.num_meshes == info.meshes
    for i in range(f.num_meshes):
        for a, b in zip(_flat_mesh(hs, i), rs.mesh(i)):
            assert a.shape == b.shape
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
        (n1, s1), (n2, s2) = hs.blas(i), rs.blas(i)
        assert n1.tobytes() == n2.tobytes() and np.array_equal(s1, s2)
    assert info.primitives == f.num_prims == rs.num_prims
    for a, b in zip(hs.prims(), rs.prims()):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
    # the TLAS, the primitives' instance (mesh -1) included
    for a, b in zip(hs.tlas(), rs.tlas()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    L = rs.lights()
    assert f.num_lights == len(L)
    for i in range(f.num_lights):
        lt = f.lights[i]
        assert (lt.mesh_id, lt.tri_idx) == (int(L[i, 0]), int(L[i, 1]))
        assert np.array_equal(_bits([*lt.emission, lt.area]), _bits(L[i, 2:6]))
    S = rs.suns()
    assert f.num_dir_lights == len(S)
    for i in range(f.num_dir_lights):
        s = f.dir_lights[i]
        assert np.array_equal(_bits([*s.direction, *s.emission]), _bits(S[i]))
    sk = f.sky
    assert np.array_equal(_bits([*sk.zenith, *sk.horizon, *sk.ground, sk.sharpness, sk.intensity]), _bits(rs.sky()))


N_CONDUCTOR, N_TRANSMISSIVE, N_SUBSURFACE, N_DELTA, N_EMISSIVE = 1, 2, 4, 8, 16


def load_both(reference, path):
    import pyoracle

    return HipScene.from_file(path), pyoracle.RefFileScene(reference, path)


# ---------------------------------------------------------------------------
# the reference's own scene files
# ---------------------------------------------------------------------------

def _ref_scene_files():
    if not os.path.isdir(REF_SCENES):
        return []
    out = []
    for d in sorted(os.listdir(REF_SCENES)):
        for f in sorted(os.listdir(os.path.join(REF_SCENES, d))):
            if f.endswith(".xenon"):
                out.append(f"{d}/{f}")
    return out


def _mirror(tmp_path, rel):
    """The reference scene directory copied under tmp_path/scenes/ (scene
    file and materials) and the OBJ files the sample scenes name written
    under tmp_path by sceneio.write_reference_meshes, so that mesh paths
    relative to tmp_path (the working directory, as the reference's run
    script sets it) resolve."""
    d = rel.split("/")[0]
    dst = tmp_path / "scenes" / d
    shutil.copytree(os.path.join(REF_SCENES, d), dst, ignore=shutil.ignore_patterns("*.py"))
    SF.write_reference_meshes(str(tmp_path))
    return str(dst / rel.split("/")[1])


@pytest.mark.parametrize("rel", _ref_scene_files())
def test_reference_scene_files(reference, tmp_path, monkeypatch, rel):
    path = _mirror(tmp_path, rel)
    monkeypatch.chdir(tmp_path)
    hs, rs = load_both(reference, path)
    assert hs.info.missing_files == 0
    assert_same_scene(hs, rs)
    # area lights sit on the stand-in box's light quad
    f = hs.flat
    for i in range(f.num_lights):
        lt = f.lights[i]
        mv = f.meshes[lt.mesh_id]
        assert lt.tri_idx < mv.num_triangles
        assert f.materials[mv.mat_ids[lt.tri_idx]].flags & N_EMISSIVE or "cow" in rel or "lucy" in rel


# ---------------------------------------------------------------------------
# written scenes: every directive, optional field, path rule and OBJ form
# ---------------------------------------------------------------------------

def _edge_scene(tmp_path):
    rng = np.random.default_rng(5)
    (tmp_path / "sub" / "materials").mkdir(parents=True)
    (tmp_path / "materials").mkdir()
    (tmp_path / "meshes").mkdir()
    # .mat: comments, indentation, quoted and unquoted names, unknown keys,
    # keys this backend ignores, a key with too few values, absent keys
    SF.write_mat(str(tmp_path / "sub" / "materials" / "glass.mat"), "glass",
                 [0.9, 0.95, 1.0, 0.05, 0.0, 1.45, 1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0],
                 extra_lines=["   # indented comment", "clearcoat 0.5", "clearcoatRoughness 0.1",
                              "subsurfaceMFP 1 2 3", "emissionTemperature 3000", "bogusKey 1 2 3"])
    with open(tmp_path / "sub" / "materials" / "partial.mat", "w") as f:
        f.write("# only some keys\nname partial\nbaseColor 0.2 0.3\nmetallic 0.75\nanisotropy -0.6\n"
                "subsurface 0.4\nsubsurfaceColor 0.9 0.5 0.25\n")
    # found relative to the working directory first (materials/ under cwd)
    SF.write_mat(str(tmp_path / "materials" / "cwd.mat"), "cwdmat",
                 [0.1, 0.2, 0.3, 0.7, 1.0, 1.5, 0.0, 0.0, 1.0, 1.0, 1.0, 0.3, 0.0, 0.0, 0.0, 0.0])
    SF.write_mat(str(tmp_path / "sub" / "materials" / "lamp.mat"), "lamp",
                 [1, 1, 1, 0.0, 0.0, 1.5, 0, 0, 1, 1, 1, 0, 4.0, 3.0, 2.0, 0])
    # OBJ with normals, every face form, quads, a pentagon, usemtl switches
    # (including an unknown name, which keeps the current material)
    V = rng.uniform(-2, 2, (40, 3)).astype(np.float32)
    Nn = rng.normal(size=(25, 3)).astype(np.float32)
    lines = ["# mixed forms", "o thing", "g group"] + [f"v {SF.vec(v)}" for v in V]
    lines += [f"vn {SF.vec(n)}" for n in Nn] + ["vt 0.5 0.5", "s 1"]
    forms = ["{p}", "{p}/{t}", "{p}//{n}", "{p}/{t}/{n}"]
    mats = ["glass", "partial", "nosuch", "cwdmat", "lamp"]
    for k in range(60):
        if k % 7 == 0:
            lines.append(f"usemtl {mats[(k // 7) % len(mats)]}")
        nvert = 3 + (k % 3)
        ps = rng.choice(40, nvert, replace=False)
        ns = rng.integers(0, 25, nvert)
        form = forms[k % 4]
        lines.append("f " + " ".join(form.format(p=p + 1, t=1, n=n + 1) for p, n in zip(ps, ns)))
    lines.append("f 1 2")   # degenerate face: no triangle
    (tmp_path / "meshes" / "mixed.obj").write_text("\n".join(lines) + "\n")
    # OBJ without normals (smooth normals, before the transform), CRLF
    # line ends and tab separators
    desc = scenes.big_mesh(width=8, height=8, level=3)
    m = desc.meshes[0]
    objtxt = "\r\n".join([f"v\t{SF.vec(v)}" for v in m.positions] +
                         ["f " + " ".join(str(i + 1) for i in t) for t in m.indices]) + "\r\n"
    (tmp_path / "sub" / "blob.obj").write_text(objtxt)
    scene = f"""# edge-case scene
config 320 200 48

camera 1.5 2.5 9   0.25 1 -0.5   55
matfile "glass.mat"
matfile partial.mat
matfile cwd.mat
matfile lamp.mat
matfile missing.mat
material legacy_a 0.8 0.2 0.1 0.0 0.3
material legacy_b 0.9 0.9 0.9 0.0 0.02 0.5 1.33 1.0 0.2
material legacy_c 0.5 0.5 0.5 1.0 0.4 0.5 1.5 0.005 0.1
material glass 0.1 0.1 0.1 0 0.9
mesh meshes/mixed.obj
mesh sub/blob.obj   0.5 -1.25 2   1.75   30 -45 12.5
mesh blob.obj
mesh meshes/mixed.obj 0 0 0 1 0 0 0
mesh meshes/absent.obj
mesh meshes/mixed.obj 1 2 3
sphere 0 1 0 1 glass
box 0 0 0 1 1 1 legacy_a
light 1 3 4 5 6 0.25
light 0 0 1 1 1 2
sunlight 0 0 0 1 1 1 3
sunlight 0.3 0.9 -0.2 1 0.9 0.8 6.5
sky 0.2 0.3 0.9  0.6 0.7 0.8
sky 0.1 0.2 0.8  0.5 0.6 0.7  0.2 0.1 0.05
sky 0.15 0.25 0.85  0.55 0.65 0.75  0.25 0.15 0.1  0.7 2.25
unknowndirective 1 2 3
"""
    path = tmp_path / "sub" / "edge.xenon"
    path.write_text(scene)
    return str(path)


def test_written_scene_every_directive(reference, tmp_path, monkeypatch):
    path = _edge_scene(tmp_path)
    monkeypatch.chdir(tmp_path)
    hs, rs = load_both(reference, path)
    assert hs.info.missing_files == 3          # missing.mat, blob.obj (mesh paths are not
    # resolved beside the scene file), meshes/absent.obj
    assert hs.info.meshes == 4
    assert hs.info.primitives == 2
    assert_same_scene(hs, rs)


def test_scene_written_from_description(reference, tmp_path, monkeypatch):
    """A description written out and loaded back is the description's own
    scene (builder path) and the reference's (file path)."""
    desc = scenes.cornell(64, 48)
    desc.suns.append(scenes.Sun((0.2, 1.0, 0.1), (1.0, 0.9, 0.8), 2.0))
    desc.sky = scenes.Sky((0.1, 0.2, 0.7), (0.5, 0.6, 0.7), (0.2, 0.1, 0.1), 0.5, 2.0)
    path = SF.write_scene(desc, str(tmp_path), with_normals=False)
    monkeypatch.chdir(tmp_path)
    hs, rs = load_both(reference, path)
    assert_same_scene(hs, rs)
    built = HipScene(desc)
    for a, b in zip(built.tlas(), hs.tlas()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    for a, b in zip(built.blas(0), hs.blas(0)):
        assert a.tobytes() == b.tobytes()


# ---------------------------------------------------------------------------
# without the reference
# ---------------------------------------------------------------------------

def test_load_errors(tmp_path):
    with pytest.raises(N.XnhError):
        HipScene.from_file(str(tmp_path / "nope.xenon"))
    # a face naming a vertex that does not exist refuses the mesh file
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    (tmp_path / "bad.xenon").write_text(f"mesh {tmp_path / 'bad.obj'}\n")
    hs = HipScene.from_file(str(tmp_path / "bad.xenon"))
    assert hs.info.missing_files == 1 and hs.info.meshes == 0


def test_defaults_without_directives(tmp_path):
    (tmp_path / "empty.xenon").write_text("# nothing\n\n")
    hs = HipScene.from_file(str(tmp_path / "empty.xenon"))
    i = hs.info
    assert (i.width, i.height, i.samples, i.min_bounces, i.max_bounces) == (800, 600, 64, 3, 8)
    assert list(i.eye) == [0.0, 5.0, 20.0] and list(i.target) == [0.0, 5.0, 0.0] and i.fov == 40.0
    f = hs.flat
    assert f.num_meshes == 0 and f.num_materials == 0 and f.num_lights == 0
    sk = f.sky
    assert _bits([*sk.zenith, *sk.horizon, *sk.ground, sk.sharpness, sk.intensity]).tolist() == \
        _bits([0.10, 0.30, 0.80, 0.55, 0.70, 0.85, 0.10, 0.08, 0.05, 3.0, 1.0]).tolist()


def test_material_load(tmp_path):
    SF.write_mat(str(tmp_path / "a.mat"), "metal", [0.9, 0.8, 0.7, 0.2, 1.0, 1.5, 0, 0, 1, 1, 1, 0.5, 0, 0, 0, 0])
    p = np.zeros(16, np.float32)
    name = C.create_string_buffer(4)
    rc = N.lib().xnh_material_load(str(tmp_path / "a.mat").encode(), p.ctypes.data_as(C.POINTER(C.c_float)),
                                   name, 4)
    assert rc == 0 and name.value == b"met"
    assert _bits(p).tolist() == _bits([0.9, 0.8, 0.7, 0.2, 1.0, 1.5, 0, 0, 1, 1, 1, 0.5, 0, 0, 0, 0]).tolist()
    rc = N.lib().xnh_material_load(str(tmp_path / "none.mat").encode(), p.ctypes.data_as(C.POINTER(C.c_float)),
                                   None, 0)
    assert rc == -1
    assert _bits(p).tolist() == _bits([0.5, 0.5, 0.5, 0.5, 0, 1.5, 0, 0, 1, 1, 1, 0, 0, 0, 0, 0]).tolist()


def test_instance_directive(reference, tmp_path, monkeypatch):
    """`instance` lines (this backend's extension; the reference's parser
    skips them) build the TLAS the reference builds over the same TRS
    placements."""
    desc = scenes.big_mesh(width=8, height=8, level=3)
    m = desc.meshes[0]
    SF.write_obj(str(tmp_path / "blob.obj"), m.positions, m.indices)
    place = [(0, (0, 0, 0), 1.0, (0, 0, 0), 0), (1, (3.5, -1, 2), 0.5, (10, 20, 30), 7),
             (0, (-4, 2, 1.25), 2.0, (0, 90, 0), 2), (1, (1, 1, 1), 1.5, (-45, 0, 12), 3)]
    lines = ["mesh blob.obj", "mesh blob.obj 0 1 0 0.75 0 30 0"]
    for mi, t, s, r, iid in place:
        tail = f" {iid}" if iid != 2 else ""        # the default id is the line's index
        lines.append(f"instance {mi} {SF.vec(t)} {SF.g(s)} {SF.vec(r)}{tail}")
    (tmp_path / "inst.xenon").write_text("\n".join(lines) + "\nlight 1 4 2 2 2 0.5\n")
    monkeypatch.chdir(tmp_path)
    hs, rs = load_both(reference, str(tmp_path / "inst.xenon"))
    assert hs.info.instances == 4 and hs.info.meshes == 2
    rs.place_instances(place)
    assert_same_scene(hs, rs)
    for a, b in zip(hs.tlas(), rs.tlas()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.parametrize("line", ["instance 2 0 0 0 1 0 0 0", "instance 0 0 0 0 1 0 0", "instance -1 0 0 0 1 0 0 0"])
def test_instance_directive_errors(tmp_path, line):
    desc = scenes.big_mesh(width=8, height=8, level=2)
    m = desc.meshes[0]
    SF.write_obj(str(tmp_path / "blob.obj"), m.positions, m.indices)
    (tmp_path / "bad.xenon").write_text(f"mesh {tmp_path / 'blob.obj'}\nmesh {tmp_path / 'blob.obj'}\n{line}\n")
    with pytest.raises(N.XnhError):
        HipScene.from_file(str(tmp_path / "bad.xenon"))


def test_written_instancing_scene_round_trip(tmp_path):
    """A description with explicit instances written out (instance lines)
    loads into the scene the builder makes from the description."""
    desc = scenes.instancing(width=32, height=18, n_inst=60, tri_range=(300, 1500), n_blas=4, n_emit=4)
    path = SF.write_scene(desc, str(tmp_path), mesh_root=str(tmp_path))
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        hs = HipScene.from_file(path)
    finally:
        os.chdir(cwd)
    built = HipScene(desc)
    assert hs.info.instances == len(desc.instances) and hs.info.missing_files == 0
    for a, b in zip(built.tlas(), hs.tlas()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    for m in range(len(desc.meshes)):
        for a, b in zip(built.blas(m), hs.blas(m)):
            assert a.tobytes() == b.tobytes()


# ---------------------------------------------------------------------------
# the reference-side binding (xenon_amd/binding/hip_renderer.cpp)
# ---------------------------------------------------------------------------

def test_reference_binding_flattens_like_the_library(reference, tmp_path):
    """The binding a maintainer adds to the reference (src/hip/hip_renderer.cpp,
    kept as xenon_amd/binding/hip_renderer.cpp) compiled against the
    reference's own headers (oracle/Makefile target `binding`): for every
    sample scene file, the edge-case scene and a written description, the
    xnh_flat_scene it makes from the reference's Scene (load_scene +
    build_acceleration) equals, array by array and byte for byte, the one
    the library makes from the same file (xnh_scene_load_file +
    xnh_scene_flat). Host code only; no GPU call."""
    import subprocess
    oracle_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    subprocess.run(["make", "-C", oracle_dir, "-s", "binding"], check=True, capture_output=True)
    exe = os.path.join(oracle_dir, "_ref", "check_binding")
    paths = []
    for rel in _ref_scene_files():
        d, f = rel.split("/")
        shutil.copytree(os.path.join(REF_SCENES, d), tmp_path / "scenes" / d,
                        ignore=shutil.ignore_patterns("*.py"), dirs_exist_ok=True)
        paths.append(str(tmp_path / "scenes" / d / f))
    SF.write_reference_meshes(str(tmp_path))
    paths.append(_edge_scene(tmp_path))
    desc = scenes.cornell(64, 48)
    desc.suns.append(scenes.Sun((0.2, 1.0, 0.1), (1.0, 0.9, 0.8), 2.0))
    (tmp_path / "written").mkdir()
    paths.append(SF.write_scene(desc, str(tmp_path / "written"), with_normals=True))
    r = subprocess.run([exe, *paths], cwd=str(tmp_path), capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert sum(line.startswith("OK ") for line in r.stdout.splitlines()) == len(paths)
    assert len(paths) >= 10
