"""Output stage, CPU side: the oracle's restatement of NoiseStats
(math/image.h:35-118) and of save_image's 8-bit conversion (main.cpp:26-31)
against the reference-written fixture and, where it is built, the
reference's own estimator; closed forms the estimator must satisfy; the PNG
encoder round trip; the swapchain's publication order."""
import os
import threading

import numpy as np
import pytest

from xenon_amd import imageio as IO
from xenon_amd.swapchain import TripleSwapchain

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("mean_squared_score", "rms_score", "mean_absolute", "median_absolute", "p95_absolute", "max_absolute",
          "avg_luminance", "edge_threshold", "epsilon")


def fixture():
    g = np.load(os.path.join(HERE, "golden", "image_metrics.npz"), allow_pickle=False)
    k = 0
    while f"img{k}" in g:
        yield g[f"img{k}"], g[f"stats{k}"], g[f"counts{k}"], g[f"ratio{k}"], g[f"rgb8_{k}"]
        k += 1


def _as_arrays(st):
    return (np.array([st[f] for f in FIELDS], np.float32),
            np.array([st["valid_pixel_count"], st["total_interior_pixel_count"]], np.int32),
            np.array([st["valid_pixel_ratio"]], np.float32))


def test_oracle_matches_fixture(oracle):
    n = 0
    for img, stats, counts, ratio, rgb8 in fixture():
        h, w = img.shape[:2]
        s, c, r = _as_arrays(oracle.noise_stats(img, w, h))
        assert s.tobytes() == stats.tobytes() and np.array_equal(c, counts) and r.tobytes() == ratio.tobytes()
        assert np.array_equal(oracle.rgb8(img, w, h), rgb8)
        n += 1
    assert n == 6


def test_oracle_matches_reference(oracle, reference):
    rng = np.random.default_rng(5)
    for trial in range(25):
        w, h = int(rng.integers(3, 120)), int(rng.integers(3, 90))
        yy, xx = np.mgrid[0:h, 0:w]
        img = (0.3 + 0.2 * np.sin(xx / 5.0) * np.cos(yy / 4.0))[..., None] * np.ones(3)
        img = img + rng.normal(size=(h, w, 3)) * (0.01 if trial % 2 else 0.3)
        if trial % 4 == 0:
            img *= 20.0
        img = img.astype(np.float32)
        a, b = oracle.noise_stats(img, w, h), reference.noise_stats(img, w, h)
        for k in a:
            assert np.asarray(a[k]).tobytes() == np.asarray(b[k]).tobytes(), (trial, k)
        assert np.array_equal(oracle.rgb8(img, w, h), reference.rgb8(img, w, h))


def test_closed_forms(oracle):
    # below 3 x 3 nothing is measured (image.h:41-42)
    z = oracle.noise_stats(np.ones((2, 9, 3), np.float32), 9, 2)
    assert all(float(z[f]) == 0.0 for f in FIELDS) and z["total_interior_pixel_count"] == 0
    # a black image: every interior pixel counts, every residual is zero,
    # and the thresholds take their floors 1e-6 and 1e-8
    c = oracle.noise_stats(np.zeros((7, 9, 3), np.float32), 9, 7)
    assert c["valid_pixel_count"] == c["total_interior_pixel_count"] == 35
    assert all(float(c[f]) == 0.0 for f in FIELDS[:7])
    assert c["edge_threshold"] == np.float32(1e-6) and c["epsilon"] == np.float32(1e-8)
    # one pixel raised by d in a unit field: the blur gives it 4/16 of its
    # own rise, so its residual is 12/16 d against a blur of about 1; its
    # eight neighbours get 2/16 or 1/16 of d; the gradient there (d / 2)
    # stays below the edge threshold (a tenth of the average), so all 49
    # interior pixels count, and the median is one of the 40 untouched ones
    d = 1.0 / 64.0
    img = np.full((9, 9, 3), 1.0, np.float32)
    img[4, 4] = 1.0 + d
    s = oracle.noise_stats(img, 9, 9)
    assert s["valid_pixel_count"] == 49
    avg = 1.0 + d / 81.0
    eps = avg * avg * 1e-4
    assert abs(float(s["avg_luminance"]) - avg) < 1e-6
    # (residual, blur) of the raised pixel, its 4 edge neighbours, its 4 corners
    rb = [(12 / 16 * d, 1 + 4 / 16 * d), (-2 / 16 * d, 1 + 2 / 16 * d), (-1 / 16 * d, 1 + 1 / 16 * d)]
    q2 = [r * r / (b * b + eps) for r, b in rb]
    q1 = [abs(r) / np.sqrt(b * b + eps) for r, b in rb]
    assert abs(float(s["max_absolute"]) / q1[0] - 1) < 1e-4
    assert abs(float(s["mean_squared_score"]) / ((q2[0] + 4 * q2[1] + 4 * q2[2]) / 49) - 1) < 1e-3
    assert abs(float(s["mean_absolute"]) / ((q1[0] + 4 * q1[1] + 4 * q1[2]) / 49) - 1) < 1e-3
    assert float(s["median_absolute"]) < 1e-6
    # rank floor(0.95 * 48) = 45 of the sorted 49: the third-largest, an edge neighbour
    assert abs(float(s["p95_absolute"]) / q1[1] - 1) < 1e-3
    # 8-bit conversion: the ends, truncation and clamping
    px = np.array([[[0.0, 1.0, 0.5], [-3.0, 7.0, 0.0]]], np.float32)
    b = oracle.rgb8(px, 2, 1)[0]
    assert b[0].tolist() == [0, 255, 186]     # 255.99 * 0.5^(1/2.2) = 186.9
    assert b[1].tolist() == [0, 255, 0]


def test_rgb8_rows_are_flipped(oracle):
    img = np.zeros((4, 3, 3), np.float32)
    img[0] = 1.0   # first framebuffer row (the top of the swapchain buffer)
    out = oracle.rgb8(img, 3, 4)
    assert out[3].min() == 255 and out[:3].max() == 0


def test_png_round_trip_and_conversion(oracle, tmp_path):
    for img, stats, counts, ratio, rgb8 in fixture():
        h, w = img.shape[:2]
        assert np.array_equal(IO.to_srgb8(img[::-1]), rgb8)   # host conversion = the reference's bytes
        p = str(tmp_path / "x.png")
        IO.write_png_rows(p, rgb8)
        assert np.array_equal(IO.read_png_rgb8(p), rgb8)


def test_swapchain_publication():
    ch = TripleSwapchain(4, 2)
    first = ch.get_read_buffer()
    assert not first.any()                    # nothing published yet: the zeroed reader
    ch.get_write_buffer()[:] = 1.0
    ch.swap_writer()
    assert ch.get_read_buffer().min() == 1.0  # the newest frame
    ch.get_write_buffer()[:] = 2.0
    ch.swap_writer()
    ch.get_write_buffer()[:] = 3.0
    ch.swap_writer()
    assert ch.get_read_buffer().min() == 3.0  # frames published in between are skipped
    assert ch.get_read_buffer().min() == 3.0  # and re-reading keeps the frame
    # a writer and a reader on two threads never share a buffer
    seen = []
    stop = threading.Event()

    def reader():
        while not stop.is_set():
            b = ch.get_read_buffer()
            v = float(b[0, 0, 0])
            assert np.all(b == v)
            seen.append(v)

    t = threading.Thread(target=reader)
    t.start()
    for k in range(4, 400):
        ch.get_write_buffer()[:] = float(k)
        ch.swap_writer()
    stop.set()
    t.join()
    assert ch.get_read_buffer().min() == 399.0
    assert seen == sorted(seen)
