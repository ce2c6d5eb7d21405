"""Analytic primitives (geometry/primitives.h, geometry/prim_blas.h), CPU
only: the builder stores the primitives as given and places them in the
TLAS as one identity instance (mesh -1) after the meshes, as
Scene::build_acceleration (scene/scene.h:79-125) does; the TLAS, BLAS and
primitive records equal the reference's, byte for byte, and the golden
fixture the reference wrote for the scene is this scene. Hit parity of the
device walk against that fixture is in test_gpu_prims.py."""
import importlib.util
import os

import numpy as np
import pytest

from xenon_amd import HipScene, scenes
from xenon_amd import native as N

HERE = os.path.dirname(os.path.abspath(__file__))


def _mg():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


def _bytes_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def test_primitives_stored_as_given():
    desc = scenes.prim_shapes(32, 18, n_spheres=6)
    hs = HipScene(desc)
    tm, p = hs.prims()
    assert len(tm) == len(desc.prims) == hs.flat.num_prims
    for i, pr in enumerate(desc.prims):
        assert (tm[i, 0], tm[i, 1]) == (pr.type_id, pr.mat)
        q = np.zeros(10, np.float32)
        q[:len(pr.params)] = pr.params
        assert _bytes_equal(p[i], q)
    nodes, order = hs.prim_blas()
    assert sorted(order.tolist()) == list(range(len(desc.prims)))
    # the identity instance over the primitives comes after the meshes
    _, _, xf, ids, meshes = hs.tlas()
    assert meshes.tolist() == [0, 1, N.PRIM_GEOMETRY] and ids.tolist() == [0, 1, 2]
    assert np.array_equal(xf[2], np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0] * 2, np.float32))


def test_prim_blas_bounds_every_primitive():
    """Every primitive lies inside the box of each node on its leaf's path
    (plane boxes are the reference's large finite slab)."""
    desc = scenes.prim_shapes(32, 18, n_spheres=40, seed=9)
    hs = HipScene(desc)
    nodes, order = hs.prim_blas()
    seen = np.zeros(len(order), bool)

    def walk(i, lo, hi):
        n = nodes[i]
        blo, bhi = np.asarray(n["lo"]), np.asarray(n["hi"])
        assert np.all(blo >= lo) and np.all(bhi <= hi)
        if n["count"] > 0:
            for s in range(n["child"], n["child"] + n["count"]):
                pr = desc.prims[order[s]]
                seen[s] = True
                if pr.kind == "sphere":
                    c, r = np.asarray(pr.params[:3], np.float32), np.float32(pr.params[3])
                    assert np.all(c - r >= blo) and np.all(c + r <= bhi)
        else:
            walk(n["child"], blo, bhi)
            walk(n["child"] + 1, blo, bhi)

    walk(0, np.full(3, -np.inf, np.float32), np.full(3, np.inf, np.float32))
    assert seen.all()


def test_bad_primitives_refused():
    import ctypes as C
    L = N.lib()
    h = L.xnh_scene_new()
    E = -1   # XNH_E_INVALID
    try:
        ones = (C.c_float * 10)(*([1.0] * 10))
        assert L.xnh_scene_add_primitive(h, 5, ones, 0) == E
        assert L.xnh_scene_add_primitive(h, -1, ones, 0) == E
        assert L.xnh_scene_add_primitive(h, 0, ones, -1) == E
        bad = (C.c_float * 10)(0.0, 0.0, float("nan"), 1.0, 0, 0, 0, 0, 0, 0)
        assert L.xnh_scene_add_primitive(h, 0, bad, 0) == E
        # a sphere reads its first 4 parameters only
        tail = (C.c_float * 10)(0.0, 0.0, 0.0, 1.0, float("nan"), 0, 0, 0, 0, 0)
        assert L.xnh_scene_add_primitive(h, 0, tail, 0) == 0
    finally:
        L.xnh_scene_free(h)


def test_fixture_is_this_scene():
    """The golden fixture (written by the reference) was made from this
    generator's scene, and its TLAS and mesh BLAS digests are those of the
    tree built here."""
    mg = _mg()
    g = np.load(os.path.join(HERE, "golden", "prim_shapes.npz"), allow_pickle=False)
    desc = mg.make_scene("prim_shapes")
    assert mg.scene_digest(desc) == str(g["scene_digest"])
    assert mg.tree_digest(HipScene(desc), desc) == str(g["tree_digest"])


@pytest.mark.parametrize("n_spheres,seed", [(24, 5), (0, 1), (200, 3)])
def test_scene_matches_reference_build(reference, n_spheres, seed):
    desc = scenes.prim_shapes(40, 30, n_spheres=n_spheres, seed=seed)
    hs, rs = HipScene(desc), reference.load(desc)
    for a, b in zip(hs.prims(), rs.prims()):
        assert _bytes_equal(a, b)
    for a, b in zip(hs.tlas(), rs.tlas()):
        assert _bytes_equal(a, b)
    for m in range(len(desc.meshes)):
        (n1, s1), (n2, s2) = hs.blas(m), rs.blas(m)
        assert n1.tobytes() == n2.tobytes() and np.array_equal(s1, s2)


def test_reference_reproduces_fixture(reference):
    mg = _mg()
    g = np.load(os.path.join(HERE, "golden", "prim_shapes.npz"), allow_pickle=False)
    desc = mg.make_scene("prim_shapes")
    rs = reference.load(desc)
    hf, hi = rs.trace_closest(g["rays"])
    assert np.array_equal(hi, g["hit_i"]) and np.array_equal(hf.view(np.uint32), g["hit_f"].view(np.uint32))
    assert np.array_equal(rs.trace_any(g["shadow_rays"]), g["occluded"])
