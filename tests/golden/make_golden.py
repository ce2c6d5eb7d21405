"""Writes the golden fixtures under tests/golden/ from the reference's own
render core (oracle/_ref/libxenon_ref.so, compiled from the reference
sources by oracle/Makefile). Run in a container that has the reference:

    make -C oracle ref && python tests/golden/make_golden.py [name ...]

(no names: every fixture). Each fixture holds the inputs (a scene named by generator + arguments and
its array digest, a ray batch) and the reference's outputs (closest hits,
occlusion, BLAS/TLAS digests, a small averaged image).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from raysets import bounce_rays, prim_rays, shadow_rays, standard_batch  # noqa: E402

import pyoracle  # noqa: E402
from xenon_amd import scenes  # noqa: E402

SCENES = {
    "cornell": ("cornell", dict(width=16, height=16, spp=4)),
    "big_mesh": ("big_mesh", dict(width=16, height=9, spp=4, level=4)),
    "instancing": ("instancing", dict(width=16, height=9, spp=4, n_inst=96, tri_range=(400, 2000), n_blas=6,
                                      n_emit=12)),
    "prim_shapes": ("prim_shapes", dict(width=48, height=27, spp=16, n_spheres=24)),
}
# camera rays per fixture (as many bounce rays follow)
RAYS = {"prim_shapes": 2048}


def make_scene(name):
    fn, kw = SCENES[name]
    return getattr(scenes, fn)(**kw)


def scene_digest(desc) -> str:
    h = hashlib.sha256()
    for m in desc.meshes:
        for a in (m.positions, m.indices, m.mat_ids):
            h.update(np.ascontiguousarray(a).tobytes())
    for mat in desc.materials:
        h.update(mat.params16().tobytes())
    for lt in desc.lights:
        h.update(np.array([lt.mesh, lt.tri, *lt.emission, lt.area], np.float64).tobytes())
    for ins in desc.instances:
        h.update(np.array([ins.mesh, *ins.translation, ins.scale, *ins.rot_deg, ins.instance_id], np.float64).tobytes())
    if desc.prims:   # scenes with primitives also digest them, the suns and the sky
        for pr in desc.prims:
            h.update(np.array([pr.type_id, pr.mat, *pr.params], np.float64).tobytes())
        for su in desc.suns:
            h.update(np.array([*su.direction, *su.color, su.intensity], np.float64).tobytes())
        sk = desc.sky
        h.update(np.array([*sk.zenith, *sk.horizon, *sk.ground, sk.intensity, sk.sharpness], np.float64).tobytes())
    return h.hexdigest()


def tree_digest(osc, desc) -> str:
    h = hashlib.sha256()
    for m in range(len(desc.meshes)):
        nodes, slots = osc.blas(m)
        h.update(nodes.tobytes())
        h.update(slots.tobytes())
    for a in osc.tlas():
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main(names):
    ref = pyoracle.Lib("ref")
    for name in names or SCENES:
        desc = make_scene(name)
        osc = ref.load(desc)
        n = RAYS.get(name, 768)
        rays = standard_batch(desc, osc.camera(), n, seed=11)
        if desc.prims:
            rays = np.concatenate([rays, prim_rays(desc, n, np.random.default_rng(14))])
        hf, hi = osc.trace_closest(rays)
        rays2 = bounce_rays(hf, hi, n, np.random.default_rng(12))
        rays = np.concatenate([rays, rays2])
        hf, hi = osc.trace_closest(rays)
        srays = shadow_rays(rays, hf, hi, np.random.default_rng(13))
        occ = osc.trace_any(srays)
        img = osc.render_average(desc.spp)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), rays=rays, hit_f=hf, hit_i=hi, shadow_rays=srays,
                            occluded=occ, image=img, camera=osc.camera(),
                            scene_digest=np.array(scene_digest(desc)), tree_digest=np.array(tree_digest(osc, desc)))
        print(name, rays.shape[0], "rays", int((hi[:, 0] >= 0).sum()), "hits", int(occ.sum()), "occluded",
              "image mean", img.mean(0))


if __name__ == "__main__":
    main(sys.argv[1:])
