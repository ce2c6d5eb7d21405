"""Host half of the HIP backend, CPU only: the C-ABI library loads and exports
every declared symbol; the host scene builder reproduces the reference's
trees, normals, instance transforms and camera bit for bit (checked against
the oracle, itself pinned to the reference); argument validation errors are
returned, not raised as crashes. No call here touches a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from xenon_amd import HipScene, scenes
from xenon_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "xenon_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xnh_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
        assert s in N.SIGNATURES, f"{s} has no ctypes signature"


def test_library_resolves_its_own_symbols():
    """Every function the library calls inside its own namespace is linked
    in (a shared library links with undefined symbols, which would only
    fail at load or first call on the GPU box)."""
    import shutil
    import subprocess
    if shutil.which("nm") is None:
        pytest.skip("nm is not installed")
    out = subprocess.run(["nm", "-D", "--undefined-only", "-C", N.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert not [l for l in out.splitlines() if "xnh" in l], out


def test_struct_layouts():
    assert C.sizeof(N.NoiseStats) == 48
    assert C.sizeof(N.BvhNode) == 32          # == BLASNode / TLASNode
    assert C.sizeof(N.Hit) == 64
    assert C.sizeof(N.Ray) == 32
    assert C.sizeof(N.InstanceRec) == 104
    assert C.sizeof(N.MaterialRec) == 64


SCENES = {
    "cornell": lambda: scenes.cornell(32, 32),
    "big_mesh": lambda: scenes.big_mesh(width=32, height=18, level=5),
    "instancing": lambda: scenes.instancing(width=32, height=18, n_inst=300, tri_range=(500, 4000), n_blas=8,
                                            n_emit=16),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_builder_matches_oracle(oracle, name):
    desc = SCENES[name]()
    hs, osc = HipScene(desc), oracle.load(desc)
    for m in range(len(desc.meshes)):
        (n1, s1), (n2, s2) = hs.blas(m), osc.blas(m)
        assert n1.tobytes() == n2.tobytes() and np.array_equal(s1, s2)
        assert np.array_equal(hs.normals(m).view(np.uint32), osc.normals(m).view(np.uint32))
    for a, b in zip(hs.tlas(), osc.tlas()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    cam = hs.camera()
    flat = np.array([*cam.eye, *cam.lower_left, *cam.horizontal, *cam.vertical], np.float32)
    assert np.array_equal(flat.view(np.uint32), osc.camera().view(np.uint32))


def test_builder_degenerate_inputs(oracle):
    """Coincident centroids (forced leaves), zero-area and duplicate
    triangles, a mesh with no triangles."""
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [2, 2, 2]], np.float32)
    F = np.array([[0, 1, 2]] * 5 + [[0, 3, 3], [4, 4, 4], [0, 1, 2]], np.int32)
    mesh = scenes.Mesh(V, F, np.zeros(len(F), np.int32))
    empty = scenes.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), np.zeros(0, np.int32))
    desc = scenes.SceneDesc("degenerate", [mesh, empty], [scenes.Material()], [],
                            scenes.Camera((0.3, 0.3, 3), (0.3, 0.3, 0)), 8, 8, 1)
    hs, osc = HipScene(desc), oracle.load(desc)
    (n1, s1), (n2, s2) = hs.blas(0), osc.blas(0)
    assert n1.tobytes() == n2.tobytes() and np.array_equal(s1, s2)
    assert hs.blas(1)[0].shape == (0,)
    for a, b in zip(hs.tlas(), osc.tlas()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def test_scene_builder_rejects_bad_input():
    lib = N.lib()
    s = lib.xnh_scene_new()
    try:
        pos = np.zeros((3, 3), np.float32)
        idx = np.array([[0, 1, 5]], np.int32)      # vertex index out of range
        mat = np.zeros(1, np.int32)
        fp = pos.ctypes.data_as(C.POINTER(C.c_float))
        rc = lib.xnh_scene_add_mesh(s, 3, fp, None, 1, idx.ctypes.data_as(C.POINTER(C.c_int32)),
                                    mat.ctypes.data_as(C.POINTER(C.c_int32)))
        assert rc == -1
        rc = lib.xnh_scene_add_mesh(s, -1, fp, None, 0, None, None)
        assert rc == -1
        t = np.zeros(3, np.float32)
        rc = lib.xnh_scene_add_instance(s, 7, t.ctypes.data_as(C.POINTER(C.c_float)), 1.0,
                                        t.ctypes.data_as(C.POINTER(C.c_float)), 0)
        assert rc == 0
        assert lib.xnh_scene_build(s, 0) == -1      # instance of a mesh that does not exist
        flat = N.FlatScene()
        assert lib.xnh_scene_flat(s, C.byref(flat)) == -1
    finally:
        lib.xnh_scene_free(s)
    assert lib.xnh_create(None, None) == -1


def test_create_reports_missing_device():
    """Without a GPU the backend refuses to create a renderer (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    cfg = N.Config(8, 8, 8, 3, 0, 0, 0, 0)
    h = C.c_void_p()
    assert N.lib().xnh_create(C.byref(cfg), C.byref(h)) == -5


FULL_SIZE = {
    "big_mesh_L8": lambda: scenes.big_mesh(width=32, height=18),
    "instancing_4096": lambda: scenes.instancing(width=32, height=18),
    "caustic_room_f158": lambda: scenes.caustic_room(width=32, height=18),
}


@pytest.mark.parametrize("name", list(FULL_SIZE))
def test_builder_matches_reference_at_full_size(reference, name):
    """The product builder against the reference's own BLAS::build /
    TLAS::build (blas.cpp:97-203, tlas.cpp:40-145, compiled into oracle/_ref)
    on the benchmark geometries at their full sizes: the 1,310,722-triangle
    displaced icosphere, the 4096-instance TLAS over ten 20k-200k-triangle
    BLAS, and the 499,280-triangle caustic-room sphere. Every node, leaf
    slot, vertex normal, TLAS node, instance order and transform equal."""
    desc = FULL_SIZE[name]()
    hs, ref = HipScene(desc), reference.load(desc)
    ntri = 0
    for m in range(len(desc.meshes)):
        (n1, s1), (n2, s2) = hs.blas(m), ref.blas(m)
        assert n1.tobytes() == n2.tobytes(), f"mesh {m}: BLAS nodes differ"
        assert np.array_equal(s1, s2), f"mesh {m}: leaf slots differ"
        assert np.array_equal(hs.normals(m).view(np.uint32), ref.normals(m).view(np.uint32))
        ntri += len(s1)
    for a, b in zip(hs.tlas(), ref.tlas()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    if name == "big_mesh_L8":
        assert ntri == 1_310_722
    elif name == "instancing_4096":
        assert len(hs.tlas()[1]) == 4097 and len(desc.meshes) == 11   # + the emitters' identity instance
        assert all(20_000 * 0.9 <= len(desc.meshes[m].indices) <= 200_000 * 1.1 for m in range(10))
    else:
        assert ntri > 499_000


def test_fast_div_matches_integer_division(tmp_path):
    """The multiply-shift division the camera-ray mapping uses (FastDiv in
    common/dev_types.h, its multiplier computed on the host) equals integer
    division for dividends below 2^31: tests/native/fast_div_check.cpp
    compiled against the header itself."""
    import shutil
    import subprocess
    from xenon_amd import build as xb
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    here = os.path.dirname(os.path.abspath(__file__))
    exe = str(tmp_path / "fast_div_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{xb.ROCM}/include", f"-I{xb.CSRC}",
                    os.path.join(here, "native", "fast_div_check.cpp"), "-o", exe], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "agree" in r.stdout
