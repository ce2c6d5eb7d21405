"""Seeded ray batches for traversal parity tests (rows: o.xyz, d.xyz, tmin, tmax)."""
from __future__ import annotations

import numpy as np

F32 = np.float32


def _unit(v):
    v = v.astype(np.float64)
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(F32)


def scene_bounds(desc):
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    if desc.instances:
        for ins in desc.instances:
            P = desc.meshes[ins.mesh].positions.astype(np.float64)
            r = np.abs(P).max() * ins.scale * 1.8 + 1e-3
            c = np.array(ins.translation)
            lo = np.minimum(lo, c - r)
            hi = np.maximum(hi, c + r)
    else:
        for m in desc.meshes:
            lo = np.minimum(lo, m.positions.min(0))
            hi = np.maximum(hi, m.positions.max(0))
    for pr in getattr(desc, "prims", []):   # planes are unbounded: left out
        p = np.asarray(pr.params, np.float64)
        if pr.kind in ("sphere", "disk"):
            pts = np.stack([p[:3] - p[-1], p[:3] + p[-1]])
        elif pr.kind == "box":
            pts = np.stack([p[:3] - p[3:6], p[:3] + p[3:6]])
        elif pr.kind == "quad":
            pts = np.stack([p[:3], p[:3] + p[3:6], p[:3] + p[6:9], p[:3] + p[3:6] + p[6:9]])
        else:
            continue
        lo = np.minimum(lo, pts.min(0))
        hi = np.maximum(hi, pts.max(0))
    return lo, hi


def camera_rays(cam12, n, rng):
    eye, llc, hz, vt = cam12[0:3], cam12[3:6], cam12[6:9], cam12[9:12]
    s = rng.random(n).astype(F32)
    t = rng.random(n).astype(F32)
    d = llc[None] + s[:, None] * hz[None] + t[:, None] * vt[None] - eye[None]
    d = d / np.sqrt((d * d).sum(1, keepdims=True)).astype(F32)
    o = np.repeat(eye[None], n, 0)
    return np.concatenate([o, d.astype(F32), np.full((n, 1), 1e-4, F32), np.full((n, 1), np.inf, F32)], 1)


def random_rays(desc, n, rng, tmax=np.inf):
    lo, hi = scene_bounds(desc)
    ext = hi - lo
    o = (lo + (rng.random((n, 3)) * 1.2 - 0.1) * ext).astype(F32)
    d = _unit(rng.normal(size=(n, 3)))
    return np.concatenate([o, d, np.full((n, 1), 1e-4, F32), np.full((n, 1), tmax, F32)], 1)


def vertex_rays(desc, n, rng):
    """Rays aimed exactly at mesh vertices and edge midpoints: shared edges and
    corners, where equal-t ties between neighbouring triangles occur."""
    lo, hi = scene_bounds(desc)
    mi = rng.integers(0, len(desc.meshes), n)
    out = np.zeros((n, 8), F32)
    for k in range(n):
        m = desc.meshes[mi[k]]
        t = rng.integers(0, m.num_tris)
        a, b = m.indices[t][rng.permutation(3)[:2]]
        p = m.positions[a].astype(np.float64)
        if rng.random() < 0.5:
            p = 0.5 * (p + m.positions[b])
        if desc.instances:   # aim at the identity instance when there is one
            pass
        o = lo + rng.random(3) * (hi - lo)
        d = p - o
        out[k, :3] = o
        out[k, 3:6] = d / np.linalg.norm(d)
        out[k, 6] = 1e-4
        out[k, 7] = np.inf
    return out


def bounce_rays(hf, hi, n, rng):
    """Rays leaving closest-hit points along random directions, offset as the
    integrator offsets bounce rays (1e-4 along the geometric normal)."""
    ok = np.where(hi[:, 0] >= 0)[0]
    if len(ok) == 0:
        return np.zeros((0, 8), F32)
    pick = ok[rng.integers(0, len(ok), n)]
    pos = hf[pick, 3:6]
    gn = hf[pick, 9:12]
    d = _unit(rng.normal(size=(n, 3)))
    flip = (gn * d).sum(1) < 0
    g = gn.copy()
    g[flip] = -g[flip]
    o = (pos + g * F32(1e-4)).astype(F32)
    return np.concatenate([o, d, np.full((n, 1), 1e-4, F32), np.full((n, 1), np.inf, F32)], 1)


def shadow_rays(rays, hf, hi, rng):
    """Segments between two surface points of a ray batch (each hit point,
    nudged off its surface, toward the hit point of another ray) with finite
    tmax, the shape of the integrator's area-light shadow rays; misses fall
    back to a point two units along their ray."""
    n = rays.shape[0]
    has = hi[:, 0] >= 0
    pts = rays[:, :3] + rays[:, 3:6] * F32(2.0)
    pts[has] = hf[has, 3:6]
    partner = rng.permutation(n)
    o = pts.copy()
    gn = hf[:, 9:12]
    tgt = pts[partner]
    d = tgt - o
    flip = (gn * d).sum(1) < 0
    g = gn.copy()
    g[flip] = -g[flip]
    o[has] = (o[has] + g[has] * F32(5e-4)).astype(F32)
    d = tgt - o
    dist = np.sqrt((d * d).sum(1)).astype(F32)
    dist[dist == 0] = 1
    d = (d / dist[:, None]).astype(F32)
    return np.concatenate([o, d, np.full((n, 1), 5e-4, F32), (dist - F32(5e-4))[:, None]], 1).astype(F32)


def prim_rays(desc, n, rng):
    """Rays from random points of the scene box aimed at points of the
    finite analytic primitives; half aim at a sphere's surface, a box's
    edges and corners, a disk's rim or a quad's edges, where the slab,
    radius and edge tests decide on equality."""
    lo, hi = scene_bounds(desc)
    fin = [p for p in desc.prims if p.kind != "plane"]
    out = np.zeros((n, 8), F32)
    for k in range(n):
        pr = fin[rng.integers(0, len(fin))]
        p = np.asarray(pr.params, np.float64)
        edge = rng.random() < 0.5
        if pr.kind == "sphere":
            v = rng.normal(size=3)
            tgt = p[:3] + p[3] * v / np.linalg.norm(v)
        elif pr.kind == "box":
            s = rng.uniform(-1.0, 1.0, 3)
            ax = rng.permutation(3)
            s[ax[0]] = rng.choice([-1.0, 1.0])
            if edge:
                s[ax[1]] = rng.choice([-1.0, 1.0])
                if rng.random() < 0.3:
                    s[ax[2]] = rng.choice([-1.0, 1.0])
            tgt = p[:3] + s * p[3:6]
        elif pr.kind == "disk":
            nn = p[3:6] / np.linalg.norm(p[3:6])
            t1 = np.cross(nn, [1.0, 0.0, 0.0] if abs(nn[0]) < 0.9 else [0.0, 1.0, 0.0])
            t1 /= np.linalg.norm(t1)
            t2 = np.cross(nn, t1)
            a = rng.uniform(0.0, 2.0 * np.pi)
            rad = p[6] if edge else p[6] * np.sqrt(rng.random())
            tgt = p[:3] + rad * (np.cos(a) * t1 + np.sin(a) * t2)
        else:   # quad
            a, b = rng.random(2)
            if edge:
                a = float(rng.choice([0.0, 1.0]))
                if rng.random() < 0.5:
                    a, b = b, a
            tgt = p[:3] + a * p[3:6] + b * p[6:9]
        org = lo + (rng.random(3) * 1.2 - 0.1) * (hi - lo)
        d = tgt - org
        out[k, :3] = org
        out[k, 3:6] = d / np.linalg.norm(d)
        out[k, 6:8] = (1e-4, np.inf)
    return out


def standard_batch(desc, cam12, n_each=4096, seed=0):
    rng = np.random.default_rng(seed)
    return np.concatenate([camera_rays(cam12, n_each, rng), random_rays(desc, n_each, rng),
                           vertex_rays(desc, n_each // 4, rng)]).astype(F32)
