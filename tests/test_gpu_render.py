"""Image parity: the HIP wavefront renderer against the CPU oracle on the same
scene, seeds and sample indices.

Both sides draw the same PCG streams per (pixel, sample), trace bit-identical
primary rays and pick the same lights; only device cos/sin/pow/exp differ from
glibc by ulps, so images agree far inside Monte-Carlo noise. The bar from the
task: relMSE <= 1e-3 and per-channel means within 3 sigma (sigma from batch
means over disjoint sample ranges)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RELMSE_MAX = 1e-3
# Regression bars, far inside the task's: the same seeds give the same paths,
# so images differ only where a device transcendental's last ulp sends a path
# elsewhere. Measured relMSE 2e-13 to 5e-7 and 73-92 % of pixel sums
# bit-identical on the reduced scenes.
RELMSE_REGRESSION = 1e-5
SAME_PIXELS_MIN = 0.6


def relmse(x, r):
    ok = np.isfinite(x) & np.isfinite(r)
    return float(np.mean((x[ok] - r[ok]) ** 2 / (r[ok] ** 2 + 1e-2)))


def _render_pair(oracle, desc, spp, spp_per_launch=0, batches=1):
    from xenon_amd import HipRenderer, HipScene
    hs = HipScene(desc)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces,
                     spp_per_launch=spp_per_launch) as r:
        r.upload_scene(hs)
        cam = hs.camera()
        per = spp // batches
        img = None
        for b in range(batches):
            img = r.render(cam, 1 + b * per, per)
        st = r.stats()
        acc, n = r.read_accum()
    osc = oracle.load(desc)
    ref_sum, ost = osc.render_sum(1, spp)
    return img, acc, ref_sum, st, ost


@pytest.mark.parametrize("name", ["cornell", "big_mesh", "instancing", "big_mesh_L8", "caustic_room"])
def test_image_matches_oracle(oracle, name):
    """Reduced scenes, and two benchmark geometries at full size (the L8
    icosphere; the caustic room's 499,280-triangle delta-glass sphere) at a
    small resolution and a few samples. The full instancing scene has its
    own test at 4096 samples below: among its dense glass and metal
    instances one path in a thousand is sent elsewhere by a last-bit
    difference, and at a few samples a single such path decides the
    relMSE."""
    from xenon_amd import scenes
    desc = {"cornell": lambda: scenes.cornell(48, 48, spp=16),
            "big_mesh": lambda: scenes.big_mesh(width=64, height=36, spp=8, level=6),
            "instancing": lambda: scenes.instancing(width=64, height=36, spp=8, n_inst=512,
                                                    tri_range=(2000, 20000), n_emit=64),
            "big_mesh_L8": lambda: scenes.big_mesh(width=64, height=36, spp=8),
            "caustic_room": lambda: scenes.caustic_room(width=64, height=36, spp=16)}[name]()
    print(f"[render parity] {name} {desc.width}x{desc.height} spp {desc.spp}", flush=True)
    img, acc, ref_sum, st, ost = _render_pair(oracle, desc, desc.spp)
    ref = ref_sum * (np.float32(1) / np.float32(desc.spp))
    e = relmse(img, ref)
    same = float(np.mean(np.all(acc == ref_sum, axis=1)))
    print(f"  relMSE {e:.3e}, bit-identical pixel sums {same:.3f}, closest rays dev {st['closest_rays']} "
          f"cpu {int(ost[0])}, shadow dev {st['shadow_rays']} cpu {int(ost[1])}", flush=True)
    assert e <= RELMSE_MAX
    assert e <= RELMSE_REGRESSION
    assert same >= SAME_PIXELS_MIN
    assert abs(st["closest_rays"] - int(ost[0])) <= 0.01 * int(ost[0])
    assert st["nonfinite"] == int(ost[2])


def test_batching_is_invisible(oracle):
    """Samples split across launches, or packed differently per launch,
    accumulate to bit-identical sums (seeds are a function of the sample)."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(32, 32, spp=12)
    hs = HipScene(desc)
    sums = []
    for spl, batches in ((0, 1), (1, 1), (5, 1), (4, 3)):
        with HipRenderer(32, 32, desc.max_bounces, desc.min_bounces, spp_per_launch=spl) as r:
            r.upload_scene(hs)
            per = 12 // batches
            for b in range(batches):
                r.render(hs.camera(), 1 + b * per, per, want_image=False)
            sums.append(r.read_accum()[0])
    for s in sums[1:]:
        assert np.array_equal(s.view(np.uint32), sums[0].view(np.uint32))


def test_render_frame_contract(oracle):
    """render_frame adds exactly one sample and returns the swapchain image
    (rows flipped), like CudaRenderer::render_frame."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(24, 16, spp=3)
    hs = HipScene(desc)
    with HipRenderer(24, 16, desc.max_bounces, desc.min_bounces) as r:
        r.upload_scene(hs)
        for k in range(3):
            frame = r.render_frame(hs.camera())
            assert r.spp() == k + 1
        assert frame.shape == (16, 24, 3)
    ref = oracle.load(desc).render_average(3).reshape(16, 24, 3)[::-1]
    assert relmse(frame, ref) <= RELMSE_MAX


def test_cornell_4096spp(oracle):
    """The task's bar at 4096 samples (Cornell box, 32x32 pixels): relMSE and
    per-channel means within 3 sigma of the oracle's."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(32, 32, spp=4096)
    hs = HipScene(desc)
    with HipRenderer(32, 32, desc.max_bounces, desc.min_bounces) as r:
        r.upload_scene(hs)
        img = r.render(hs.camera(), 1, 4096)
    osc = oracle.load(desc)
    nb = 8
    parts = [osc.render_sum(1 + k * 512, 512)[0] * np.float32(1 / 512) for k in range(nb)]
    ref = np.mean(parts, axis=0)
    e = relmse(img, ref)
    means = np.array([p.mean(0) for p in parts])
    sigma = means.std(0, ddof=1) / np.sqrt(nb)
    diff = np.abs(img.mean(0) - ref.mean(0))
    print(f"  4096spp relMSE {e:.3e}; channel mean diff {diff} vs 3 sigma {3 * sigma}", flush=True)
    assert e <= RELMSE_MAX
    assert np.all(diff <= 3 * sigma + 1e-6)


def test_caustic_room_full_geometry_512spp(oracle):
    """The guided configuration's scene rendered plain at full geometry
    (499,280-triangle delta-glass sphere, freq 158; the room lit only through
    the wall opening) against the oracle: relMSE and per-channel means within
    3 sigma of the oracle's, sigma from 8 batch means over disjoint sample
    ranges."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.caustic_room(width=48, height=27, spp=512)
    assert desc.num_tris > 499_000
    hs = HipScene(desc)
    print(f"[caustic parity] {desc.width}x{desc.height}, {desc.spp} spp, {desc.num_tris} tris", flush=True)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces) as r:
        r.upload_scene(hs)
        img = r.render(hs.camera(), 1, desc.spp)
        acc, _ = r.read_accum()
    osc = oracle.load(desc)
    nb, per = 8, desc.spp // 8
    sums = [osc.render_sum(1 + k * per, per)[0] for k in range(nb)]
    parts = [s_ * np.float32(1 / per) for s_ in sums]
    ref = np.mean(parts, axis=0)
    e = relmse(img, ref)
    means = np.array([p.mean(0) for p in parts])
    sigma = means.std(0, ddof=1) / np.sqrt(nb)
    diff = np.abs(img.mean(0) - ref.mean(0))
    print(f"  relMSE {e:.3e}; channel mean diff {diff} vs 3 sigma {3 * sigma}", flush=True)
    assert e <= RELMSE_MAX
    assert e <= RELMSE_REGRESSION
    assert np.all(diff <= 3 * sigma)
    assert np.all(ref.mean(0) > 0)


def test_instancing_full_size_4096spp(oracle):
    """The instancing configuration at full size (TLAS over 4096 instances
    of ten 20k-200k-triangle BLAS, 40/30/30 diffuse/GGX/dielectric, 64
    emissive triangles, depth 10) at the task's bar: 4096 samples, relMSE
    <= 1e-3 and per-channel means within 3 sigma of the oracle's (sigma from
    8 batch means over disjoint sample ranges)."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.instancing(width=24, height=14, spp=4096)
    hs = HipScene(desc)
    print(f"[instancing parity] {desc.width}x{desc.height}, {desc.spp} spp, {len(hs.tlas()[1])} instances",
          flush=True)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces) as r:
        r.upload_scene(hs)
        img = r.render(hs.camera(), 1, desc.spp)
    osc = oracle.load(desc)
    nb, per = 8, desc.spp // 8
    parts = [osc.render_sum(1 + k * per, per)[0] * np.float32(1 / per) for k in range(nb)]
    ref = np.mean(parts, axis=0)
    e = relmse(img, ref)
    means = np.array([p.mean(0) for p in parts])
    sigma = means.std(0, ddof=1) / np.sqrt(nb)
    diff = np.abs(img.mean(0) - ref.mean(0))
    print(f"  relMSE {e:.3e}; channel mean diff {diff} vs 3 sigma {3 * sigma}", flush=True)
    assert e <= RELMSE_MAX
    assert np.all(diff <= 3 * sigma)
    assert np.all(ref.mean(0) > 0)


def test_scene_file_renders_like_its_description(tmp_path):
    """A scene written as .xenon + .mat + OBJ and loaded by
    xnh_scene_load_file renders the same image, bit for bit, as the
    description it was written from (the loader's parity with the
    reference's parser is checked on the CPU)."""
    from xenon_amd import sceneio as SF
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(40, 30, spp=8)
    desc.suns.append(scenes.Sun((0.3, 1.0, 0.2), (1.0, 0.9, 0.8), 1.5))
    path = SF.write_scene(desc, str(tmp_path))
    print(f"[scene file] {path}", flush=True)
    info = HipScene.from_file(path).info
    assert info.missing_files == 0 and info.meshes == 1
    imgs = []
    for hs in (HipScene(desc), HipScene.from_file(path)):
        with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces) as r:
            r.upload_scene(hs)
            imgs.append(r.render(hs.camera(desc.width, desc.height), 1, desc.spp).copy())
    assert np.isfinite(imgs[0]).all() and imgs[0].mean() > 0
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
