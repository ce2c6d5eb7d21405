"""The traversal stack's limit and its error path.

The walker keeps one stack of 128 entries per ray (the reference's TLAS and
BLAS stacks of 64 each, tlas.cpp:187 and blas.cpp:219). The first 16 by
default sit in LDS and the rest spill to device memory. A push beyond 128
ends the ray and the call returns XNH_E_OVERFLOW with the number of such
rays in the stats.

The scenes here are flat scenes (the upload's C-ABI form) with a
hand-built BLAS: a chain of `depth` interior nodes, each with a leaf
child holding one triangle at x = 10.5 and the next chain node as its
other child, the last chain node a leaf holding a triangle at x = 5. A
ray along x through every box pushes one entry per level. A closest-hit
ray visits the nearer child first and pushes the farther one. An any-hit
ray visits the farther child first and pushes the nearer one. The boxes
therefore put the chain nearer for a ray coming from -x (closest-hit
test) and the leaves nearer for a ray coming from +x (any-hit test).
Below the limit, the expected hits follow from the geometry alone: the
bottom triangle at t = 6 from x = -1, and an occluder on any ray from
x = 12."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OVERFLOW = -4


def _chain_scene(depth: int):
    """A HipScene of depth + 1 triangles (flat, facing +x, covering
    y, z in [-1, 3] with y + z <= 2) and a flat scene replacing its BLAS
    by the chain. Returns the HipScene, the flat scene and the arrays
    the flat scene points into."""
    from xenon_amd import HipScene, native as N, scenes
    ntri = depth + 1
    pos, idx = [], []
    for k in range(ntri):
        x = 5.0 if k == depth else 10.5
        pos += [(x, -1.0, -1.0), (x, 3.0, -1.0), (x, -1.0, 3.0)]
        idx.append((3 * k, 3 * k + 1, 3 * k + 2))
    mesh = scenes.Mesh(np.array(pos, np.float32), np.array(idx, np.int32), np.zeros(ntri, np.int32))
    desc = scenes.SceneDesc("stack_chain", [mesh], [scenes.Material()], [],
                            scenes.Camera((-1.0, 0.1, 0.2), (10.0, 0.1, 0.2), fov=5.0), 8, 8, 1)
    hs = HipScene(desc)
    # node 2k: chain interior node k, children 2k+1 (leaf k) and 2k+2
    # (chain node k+1); node 2*depth: the bottom leaf
    nn = 2 * depth + 1
    nodes = (N.BvhNode * nn)()
    for k in range(depth):
        c, leaf = nodes[2 * k], nodes[2 * k + 1]
        c.lo[:] = (5.0, -1.0, -1.0)
        c.hi[:] = (10.4, 3.0, 3.0)
        c.child, c.count = 2 * k + 1, 0
        leaf.lo[:] = (10.5, -1.0, -1.0)
        leaf.hi[:] = (10.5, 3.0, 3.0)
        leaf.child, leaf.count = k, 1
    b = nodes[2 * depth]
    b.lo[:] = (5.0, -1.0, -1.0)
    b.hi[:] = (5.0, 3.0, 3.0)
    b.child, b.count = depth, 1
    prims = np.arange(ntri, dtype=np.uint32)
    # the chain is a valid BLAS: children after their parent and inside
    # the node array, leaves inside the slot range
    for j in range(nn):
        n = nodes[j]
        if n.count == 0:
            assert j < n.child and n.child + 1 < nn
        else:
            assert n.child + n.count <= ntri
    src = hs.flat
    assert src.num_meshes == 1 and src.num_instances == 1
    mv = N.MeshView()
    C.memmove(C.byref(mv), src.meshes, C.sizeof(N.MeshView))
    mv.num_nodes = nn
    mv.nodes = C.cast(nodes, C.POINTER(N.BvhNode))
    mv.prims = prims.ctypes.data_as(C.POINTER(C.c_uint32))
    flat = N.FlatScene()
    C.memmove(C.byref(flat), C.byref(src), C.sizeof(N.FlatScene))
    flat.meshes = C.pointer(mv)
    return hs, flat, (nodes, prims, mv)


def _upload(r, flat):
    from xenon_amd import native as N
    r._chk(N.lib().xnh_upload_scene(r.h, C.byref(flat)))


def _rays(n, org, d, tmax):
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3:6] = d
    rays[:, 6] = 0.0
    rays[:, 7] = tmax
    return rays


def test_chain_below_the_limit_spills_and_hits():
    """Depth 100: 100 entries, most of them in the spill; the bottom
    triangle at t = 6 is found and an any-hit ray is occluded."""
    from xenon_amd import HipRenderer
    depth = 100
    hs, flat, keep = _chain_scene(depth)
    with HipRenderer(8, 8, 8, 3, spp_per_launch=4) as r:
        _upload(r, flat)
        print(f"[stack] chain depth {depth}: closest", flush=True)
        hits = r.trace_closest(_rays(256, (-1.0, 0.1, 0.2), (1.0, 0.0, 0.0), np.inf))
        assert np.all(hits["prim_id"] == depth)
        assert np.all(np.abs(hits["t"] - 6.0) <= 4e-6)
        assert r.stats()["stack_overflows"] == 0
        print(f"[stack] chain depth {depth}: any", flush=True)
        occ = r.trace_any(_rays(256, (12.0, 0.1, 0.2), (-1.0, 0.0, 0.0), 100.0))
        assert np.all(occ == 1)
        print(f"[stack] chain depth {depth}: render", flush=True)
        img = r.render(hs.camera(), 1, 4)
        assert np.all(np.isfinite(img))
        assert r.stats()["stack_overflows"] == 0


def test_chain_beyond_the_limit_reports_overflow_and_recovers():
    """Depth 200: every ray pushes past 128 entries. Closest-hit and
    any-hit batches and a render each return XNH_E_OVERFLOW and count
    the rays; the context then renders a valid scene normally."""
    from xenon_amd import HipRenderer, native as N
    hs, flat, keep = _chain_scene(200)
    hs_ok, flat_ok, keep_ok = _chain_scene(100)
    n = 256
    with HipRenderer(8, 8, 8, 3, spp_per_launch=4) as r:
        _upload(r, flat)
        print("[stack] chain depth 200: closest", flush=True)
        with pytest.raises(N.XnhError) as e:
            r.trace_closest(_rays(n, (-1.0, 0.1, 0.2), (1.0, 0.0, 0.0), np.inf))
        assert e.value.code == OVERFLOW
        assert r.stats()["stack_overflows"] == n
        print("[stack] chain depth 200: any", flush=True)
        with pytest.raises(N.XnhError) as e:
            r.trace_any(_rays(n, (12.0, 0.1, 0.2), (-1.0, 0.0, 0.0), 100.0))
        assert e.value.code == OVERFLOW
        assert r.stats()["stack_overflows"] == n
        print("[stack] chain depth 200: render", flush=True)
        with pytest.raises(N.XnhError) as e:
            r.render(hs.camera(), 1, 4)
        assert e.value.code == OVERFLOW
        assert r.stats()["stack_overflows"] >= 8 * 8 * 4
        print("[stack] chain depth 100 after the overflow", flush=True)
        r.reset()
        _upload(r, flat_ok)
        hits = r.trace_closest(_rays(n, (-1.0, 0.1, 0.2), (1.0, 0.0, 0.0), np.inf))
        assert np.all(hits["prim_id"] == 100)
        img = r.render(hs_ok.camera(), 1, 4)
        assert np.all(np.isfinite(img))
