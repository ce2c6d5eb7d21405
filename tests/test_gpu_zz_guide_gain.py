"""What the trained guide gains over the method without it, at equal sample
counts, on the caustic room at full geometry (a closed room lit only through
a wall opening, with a 499,280-triangle delta-glass sphere: light that BSDF
sampling and NEE find only by chance).

The guide is trained on the device by its own kernels (guided rendering of
TRAIN_SPP samples with training after every batch, from fresh parameters),
then frozen, and the compared images are rendered with the frozen network:
a frozen network makes the guided render a deterministic function of the
network, so the measure does not move with the order of the training's
floating-point atomics beyond what the trained network itself varies. The
same sample indices are rendered plain and guided; errors are per-pixel
relative squared errors (channel mean of (x - r)^2 / (r^2 + 1e-2)) against a
plain render of 16x the samples from disjoint sample indices. Guiding must
lower both the median pixel error and the mean without the 0.1 % largest
pixel errors by the factors below (measured 1.29x and 1.23x; the assertion
messages print the values of the run). The full mean is set by single caustic outliers at these counts,
for both methods alike, and is reported only.

(Collected last: a regression here must not hide the parity tests.)"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPP = 1024
REF_SPP = 16 * SPP
TRAIN_SPP = 2048
MEDIAN_GAIN_MIN = 1.18   # measured 1.29x
TRIM_GAIN_MIN = 1.10     # measured 1.23x


def _errors(img, ref):
    e = ((img - ref) ** 2 / (ref ** 2 + 1e-2)).reshape(-1, 3).mean(1)
    e = np.sort(e[np.isfinite(e)])
    return float(np.median(e)), float(e[: int(len(e) * 0.999)].mean()), float(e.mean())


def test_trained_guide_lowers_error_at_equal_samples():
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.caustic_room(width=160, height=90)
    assert desc.num_tris > 499_000
    hs = HipScene(desc)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces, spp_per_launch=8) as r:
        r.upload_scene(hs)
        cam = hs.camera()
        print(f"[guide gain] caustic room 160x90 ({desc.num_tris} tris): reference {REF_SPP} spp", flush=True)
        r.render(cam, 10_000_001, REF_SPP, want_image=False)
        ref_sum, n = r.read_accum()
        ref = ref_sum / n
        r.reset()
        plain = r.render(cam, 1, SPP)
        r.reset()
        r.guide_enable(seed=11)
        print(f"[guide gain] training on {TRAIN_SPP} guided samples", flush=True)
        for k in range(TRAIN_SPP // 8):
            r.render(cam, 20_000_001 + 8 * k, 8, want_image=False)
        r.guide_set_training(False)
        r.reset()
        print(f"[guide gain] frozen guide, {SPP} spp", flush=True)
        guided = r.render(cam, 1, SPP)
        r.guide_disable()
    mp, tp, fp = _errors(plain, ref)
    mg, tg, fg = _errors(guided, ref)
    msg = (f"{SPP} spp: median pixel error plain {mp:.5f} guided {mg:.5f} ({mp / mg:.2f}x); trimmed mean plain "
           f"{tp:.5f} guided {tg:.5f} ({tp / tg:.2f}x); full mean plain {fp:.4f} guided {fg:.4f} ({fp / fg:.2f}x)")
    print("  " + msg, flush=True)
    assert np.isfinite(guided).all()
    assert mp / mg >= MEDIAN_GAIN_MIN, msg
    assert tp / tg >= TRIM_GAIN_MIN, msg
