"""What the trained guide gains over the method without it, at equal sample
counts, on the caustic room (a closed room lit only through a wall opening,
with a delta-glass sphere: caustics that BSDF sampling and NEE find only by
chance). The same samples are rendered plain and guided; the guided run
trains its network on the device between batches of 8 samples, from fresh
parameters, so its image includes the untrained start. Errors are per-pixel
relative squared errors against a plain render of 16x the samples from
disjoint sample indices. Guiding must lower the median pixel error by at
least 5 % (measured 1.11-1.12x over four runs) and must not raise the mean
without the 0.1 % largest pixel errors by more than 10 % (it was 1.01-1.23x
lower over the same runs; training order on the device varies, so that
measure moves most). The full
mean is set by single caustic outliers at these counts, for both methods
alike.

(Collected last: a regression here must not hide the parity tests.)"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPP = 1024
REF_SPP = 16 * SPP


def _errors(img, ref):
    e = ((img - ref) ** 2 / (ref ** 2 + 1e-2)).reshape(-1, 3).mean(1)
    e = np.sort(e[np.isfinite(e)])
    return float(np.median(e)), float(e[: int(len(e) * 0.999)].mean())


def test_guide_lowers_error_at_equal_samples():
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.caustic_room(width=160, height=90, freq=40)
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
        r.guide_enable()
        for k in range(SPP // 8):
            r.render(cam, 1 + 8 * k, 8, want_image=False)
        acc, n = r.read_accum()
        guided = acc / n
        r.guide_disable()
    mp, tp = _errors(plain, ref)
    mg, tg = _errors(guided, ref)
    print(f"  {SPP} spp: median pixel error plain {mp:.5f} guided {mg:.5f} ({mp / mg:.2f}x); trimmed mean plain "
          f"{tp:.5f} guided {tg:.5f} ({tp / tg:.2f}x)", flush=True)
    assert np.isfinite(guided).all()
    assert mg < 0.95 * mp
    assert tg < 1.10 * tp
