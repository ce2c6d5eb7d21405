"""Writes the golden fixtures under tests/golden/ from the reference's own
render core (oracle/_ref/libxenon_ref.so, compiled from the reference
sources by oracle/Makefile). Run in a container that has the reference:

    make -C oracle ref && python tests/golden/make_golden.py [name ...]

(no names: every fixture; image_metrics: the output-stage fixture). Each
scene fixture holds the inputs (a scene named by generator + arguments and
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


def image_set():
    """Framebuffers (swapchain layout) for the output-stage fixture: smooth
    shading with noise, hard edges, a constant field, HDR values beyond
    [0, 1] and negatives, the smallest measured size and a degenerate one."""
    rng = np.random.default_rng(21)
    out = []
    h, w = 48, 64
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    smooth = (0.35 + 0.2 * np.sin(xx / 6.0) * np.cos(yy / 9.0))[..., None] * np.array([1.0, 0.85, 0.7], np.float32)
    out.append(smooth + rng.normal(size=(h, w, 3)).astype(np.float32) * 0.02)
    h, w = 60, 80
    yy, xx = np.mgrid[0:h, 0:w]
    blocks = (((yy // 10) + (xx // 10)) % 2).astype(np.float32)[..., None] * np.array([0.8, 0.5, 0.3], np.float32)
    out.append(blocks + 0.05 + rng.normal(size=(h, w, 3)).astype(np.float32) * 0.03)
    out.append(np.full((16, 16, 3), 0.42, np.float32))
    hdr = rng.lognormal(mean=-0.5, sigma=1.2, size=(40, 50, 3)).astype(np.float32)
    hdr[rng.random((40, 50)) < 0.05] *= -1.0
    out.append(hdr)
    out.append(rng.random((3, 3, 3)).astype(np.float32))
    out.append(rng.random((2, 5, 3)).astype(np.float32))
    return [np.ascontiguousarray(a, np.float32) for a in out]


def write_image_fixture(ref):
    arrs = {}
    for k, img in enumerate(image_set()):
        h, w = img.shape[:2]
        st = ref.noise_stats(img, w, h)
        arrs[f"img{k}"] = img
        arrs[f"stats{k}"] = np.array([st[f] for f in pyoracle.NOISE_FIELDS], np.float32)
        arrs[f"counts{k}"] = np.array([st["valid_pixel_count"], st["total_interior_pixel_count"]], np.int32)
        arrs[f"ratio{k}"] = np.array([st["valid_pixel_ratio"]], np.float32)
        arrs[f"rgb8_{k}"] = ref.rgb8(img, w, h)
        print("image", k, w, h, st["valid_pixel_count"], st["rms_score"])
    np.savez_compressed(os.path.join(HERE, "image_metrics.npz"), **arrs)


def main(names):
    ref = pyoracle.Lib("ref")
    if not names or "image_metrics" in names:
        write_image_fixture(ref)
        names = [n for n in names if n != "image_metrics"]
        if not names and len(sys.argv) > 1:
            return
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
