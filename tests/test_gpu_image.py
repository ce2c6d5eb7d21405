"""Output stage on the device: NoiseStats (math/image.h:35-118) and the
8-bit PNG conversion (main.cpp:26-31) computed by the HIP kernels, against
the reference-written fixture and the oracle. Per-pixel values are the
reference's float expressions in its order, so counts, maximum, median,
95th percentile and the 8-bit bytes must be identical; the four values that
come from double sums (average luminance and the three means, with the
thresholds and rms derived from them) are summed in another order and are
held to 1 ulp."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("mean_squared_score", "rms_score", "mean_absolute", "median_absolute", "p95_absolute", "max_absolute",
          "avg_luminance", "edge_threshold", "epsilon")
SUMMED = {"mean_squared_score", "rms_score", "mean_absolute", "avg_luminance", "edge_threshold", "epsilon"}


def _ulps(a, b):
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32)))


def _same(dev, ref, where):
    for k in FIELDS:
        if k in SUMMED:
            assert _ulps(dev[k], ref[k]) <= 1, (where, k, dev[k], ref[k])
        else:
            assert np.float32(dev[k]).tobytes() == np.float32(ref[k]).tobytes(), (where, k, dev[k], ref[k])
    for k in ("valid_pixel_count", "total_interior_pixel_count"):
        assert int(dev[k]) == int(ref[k]), (where, k)
    assert _ulps(dev["valid_pixel_ratio"], ref["valid_pixel_ratio"]) == 0, where


@pytest.fixture(scope="module")
def renderer():
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(96, 72, spp=8)
    hs = HipScene(desc)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces) as r:
        r.upload_scene(hs)
        yield r, hs, desc


def test_fixture_images(renderer):
    r, _, _ = renderer
    g = np.load(os.path.join(HERE, "golden", "image_metrics.npz"), allow_pickle=False)
    k = 0
    while f"img{k}" in g:
        img = g[f"img{k}"]
        h, w = img.shape[:2]
        print(f"[image] fixture {k}: {w}x{h}", flush=True)
        st = r.noise_stats(img)
        ref = dict(zip(FIELDS, g[f"stats{k}"]))
        ref.update(valid_pixel_count=g[f"counts{k}"][0], total_interior_pixel_count=g[f"counts{k}"][1],
                   valid_pixel_ratio=g[f"ratio{k}"][0])
        _same(st, ref, k)
        if w > 0 and h > 0:
            assert np.array_equal(r.image_rgb8(img), g[f"rgb8_{k}"]), k
        k += 1
    assert k == 6


def test_full_hd_against_oracle(renderer, oracle):
    """1920 x 1080 (2,073,600 pixels, 8 radix passes' worth of keys): device
    against the oracle on a framebuffer with smooth shading, hard edges, noise
    and HDR values."""
    r, _, _ = renderer
    rng = np.random.default_rng(8)
    h, w = 1080, 1920
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    img = (0.4 + 0.3 * np.sin(xx / 40.0) * np.cos(yy / 33.0))[..., None] * np.array([1.0, 0.8, 0.6], np.float32)
    img = img * (1.0 + ((xx // 160 + yy // 120) % 2)[..., None]) + rng.normal(size=(h, w, 3)).astype(np.float32) * 0.05
    img = np.ascontiguousarray(img, np.float32)
    print("[image] 1920x1080 noise stats", flush=True)
    _same(r.noise_stats(img), oracle.noise_stats(img, w, h), "fullhd")
    assert np.array_equal(r.image_rgb8(img), oracle.rgb8(img, w, h))


def test_running_average_source(renderer, oracle):
    """With no framebuffer the calls read the context's running average in
    the swapchain layout: the same results as the frame render_frame returns."""
    r, hs, desc = renderer
    r.reset()
    frame = None
    for _ in range(desc.spp):
        frame = r.render_frame(hs.camera())
    print("[image] running average of cornell 96x72", flush=True)
    a, b = r.noise_stats(), r.noise_stats(frame)
    for k in a:
        assert np.asarray(a[k]).tobytes() == np.asarray(b[k]).tobytes(), k
    _same(a, oracle.noise_stats(frame, desc.width, desc.height), "average")
    assert np.array_equal(r.image_rgb8(), oracle.rgb8(frame, desc.width, desc.height))


def test_save_image_and_progressive(renderer, tmp_path):
    from xenon_amd import TripleSwapchain, imageio, progressive
    r, hs, desc = renderer
    r.reset()
    chain = TripleSwapchain(desc.width, desc.height)
    out = progressive(r, hs.camera(), chain, max_spp=256, noise_target=1e9, check_every=4)
    assert out["spp"] == 4 and out["noise"]["valid_pixel_count"] > 0   # the first check meets any target
    assert chain.frames == 4
    a, b = r.noise_stats(chain.get_read_buffer()), r.noise_stats()   # the newest frame is the running average
    assert all(np.asarray(a[k]).tobytes() == np.asarray(b[k]).tobytes() for k in a)
    p = str(tmp_path / "f.png")
    st = r.save_image(p, report=False)
    assert np.array_equal(imageio.read_png_rgb8(p), r.image_rgb8())
    assert st["valid_pixel_count"] > 0


def test_errors(renderer):
    import ctypes as C
    from xenon_amd import native as N
    r, _, _ = renderer
    st = N.NoiseStats()
    L = N.lib()
    assert L.xnh_image_noise_stats(r.h, None, 5, 5, C.byref(st)) == -1          # not the context's size
    assert L.xnh_image_noise_stats(r.h, None, 8193, 8193, C.byref(st)) == -1    # above 2^26 pixels
    assert L.xnh_image_noise_stats(r.h, None, -1, 5, C.byref(st)) == -1
    with pytest.raises(ValueError):
        r.noise_stats(np.zeros(10, np.float32), width=4, height=4)
    z = r.noise_stats(np.zeros((2, 2, 3), np.float32))
    assert z["total_interior_pixel_count"] == 0
