"""Guided rendering: one-sample MIS of the learned vMF mixture and the BSDF
sampler, with the network trained on the device after every batch.

The guided estimator is unbiased for any network state, so its image mean
must match a converged plain render within the statistical error measured
from independent sample ranges; training must actually run (parameters move,
records are collected) and the image must stay finite."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lum(img):
    return img @ np.array([0.2126, 0.7152, 0.0722], np.float32)


def test_guided_cornell_unbiased():
    from xenon_amd import HipRenderer, HipScene, scenes
    W = H = 64
    desc = scenes.cornell(W, H, spp=1)
    hs = HipScene(desc)
    with HipRenderer(W, H, 8, 3, spp_per_launch=64) as r:
        r.upload_scene(hs)
        cam = hs.camera()
        print("[guided] plain reference 16384 spp", flush=True)
        ref = r.render(cam, 1, 16384)
        r.reset()
        r.guide_enable(t_log2=14, learning_rate=5e-3, seed=1, train_batch=1 << 14, record_paths=1 << 13)
        g0, w0 = r.guide_params()
        means, imgs = [], []
        for k in range(8):
            print(f"[guided] chunk {k}", flush=True)
            r.reset()
            img = r.render(cam, 100001 + 256 * k, 256)
            assert np.all(np.isfinite(img))
            imgs.append(img)
            means.append(float(_lum(img).mean()))
        g1, w1 = r.guide_params()
        assert np.abs(w1 - w0).max() > 1e-3, "the network was not trained"
    m, s = float(np.mean(means)), float(np.std(means, ddof=1) / np.sqrt(len(means)))
    mref = float(_lum(ref).mean())
    print(f"[guided] mean {m:.5f} +- {s:.5f}, plain {mref:.5f}", flush=True)
    assert abs(m - mref) < 4 * s + 2e-3 * mref
    # per-pixel agreement of the guided average with the plain reference
    avg = np.mean(imgs, 0)
    rel = np.mean((avg - ref) ** 2 / (ref ** 2 + 1e-2))
    assert rel < 5e-2


def test_guided_caustic_room_unbiased():
    """The guided configuration's scene at full geometry (499,280-triangle
    delta-glass sphere) and a small resolution: the guided image mean, with
    training after every batch from fresh parameters, matches a plain render
    of 16x the samples within the statistical error of 8 chunk means."""
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.caustic_room(width=64, height=36, spp=1)
    hs = HipScene(desc)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces, spp_per_launch=256) as r:
        r.upload_scene(hs)
        cam = hs.camera()
        print(f"[guided caustic] {desc.num_tris} tris: plain reference 65536 spp", flush=True)
        ref = r.render(cam, 1_000_001, 65536)
        r.reset()
        r.guide_enable(seed=3)
        g0, w0 = r.guide_params()
        means = []
        for k in range(8):
            print(f"[guided caustic] chunk {k}", flush=True)
            r.reset()
            img = r.render(cam, 1 + 512 * k, 512)
            assert np.all(np.isfinite(img))
            means.append(float(_lum(img).mean()))
        g1, w1 = r.guide_params()
        assert np.abs(w1 - w0).max() > 1e-3, "the network was not trained"
        r.guide_disable()
    m, s = float(np.mean(means)), float(np.std(means, ddof=1) / np.sqrt(len(means)))
    mref = float(_lum(ref).mean())
    print(f"[guided caustic] mean {m:.5f} +- {s:.5f}, plain {mref:.5f}", flush=True)
    assert mref > 0
    assert abs(m - mref) < 4 * s + 2e-3 * mref


def test_guided_frozen_network_renders_deterministically():
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(32, 32, spp=1)
    hs = HipScene(desc)
    outs = []
    for _ in range(2):
        with HipRenderer(32, 32, 8, 3, spp_per_launch=8) as r:
            r.upload_scene(hs)
            r.guide_enable(t_log2=12, seed=7)
            r.guide_set_training(False)
            outs.append(r.render(hs.camera(), 1, 32))
    assert np.array_equal(outs[0], outs[1])
