"""xnh_image_noise_stats_noalloc: NoiseStats (math/image.h:35-118) of a
framebuffer already in device memory, computed in the library's static
scratch array with no allocation. Its results must equal, field for field
and bit for bit, those of xnh_image_noise_stats (the allocating entry point
on the same kernels) on the same image, over calls in a row on the shared
array (the second call on a smaller image, which must not see the first
call's scratch contents); sizes beyond the static maximum, host memory and
images that run past their device allocation are refused with an error,
before any kernel runs.

Device buffers come from the HIP runtime the library itself links
(libamdhip64.so.7), with exact-size hipMalloc and synchronous copies."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_PIXELS = 1 << 22


def _hip():
    from xenon_amd import native as N
    N.lib()   # loads libamdhip64.so.7 first; the name below resolves to it
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class DevBuf:
    """An exact-size device allocation holding a copy of a host array."""

    def __init__(self, host: np.ndarray = None, nbytes: int = None):
        self.hip = _hip()
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), self.nbytes) == 0
        if host is not None:
            assert self.hip.hipMemcpy(self.p, host.ctypes.data, self.nbytes, 1) == 0   # host to device

    @property
    def ptr(self) -> int:
        return self.p.value

    def free(self):
        if self.p.value:
            assert self.hip.hipFree(self.p) == 0
            self.p = C.c_void_p()


def _image(w, h, seed):
    """A render-like framebuffer: a smooth gradient with noise, a few sharp
    edges (the edge test), zeros and a bright spot."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    base = 0.2 + 0.6 * (x / max(w - 1, 1)) * (0.5 + 0.5 * np.sin(y / 17.0))
    img = np.stack([base, base * 0.8, base * 0.6], -1)
    img *= 1.0 + 0.15 * rng.standard_normal(img.shape).astype(np.float32)
    img[h // 3: h // 3 + 5, :, :] = 3.0
    img[:, w // 4: w // 4 + 2, :] = 0.0
    img[h // 2, w // 2, :] = 50.0
    return np.ascontiguousarray(img.astype(np.float32))


@pytest.fixture(scope="module")
def ctx():
    from xenon_amd import HipRenderer, HipScene, scenes
    desc = scenes.cornell(64, 64, spp=4)
    with HipRenderer(desc.width, desc.height, desc.max_bounces, desc.min_bounces) as r:
        r.upload_scene(HipScene(desc))
        yield r


def _noalloc(r, ptr, w, h):
    from xenon_amd import native as N
    st = N.NoiseStats()
    rc = N.lib().xnh_image_noise_stats_noalloc(r.h, C.c_void_p(ptr), w, h, C.byref(st))
    return rc, st.as_dict()


def _err(r):
    from xenon_amd import native as N
    return N.lib().xnh_last_error(r.h)


def _bits(d):
    out = {}
    for k, v in d.items():
        out[k] = int(v) if isinstance(v, (int, np.integer)) else np.float32(v).view(np.uint32).item()
    return out


def test_noalloc_equals_allocating_entry_point(ctx):
    r = ctx
    for call, (w, h) in enumerate([(1920, 1080), (640, 360), (2048, 2048), (3, 3)]):
        print(f"[noalloc] call {call}: {w}x{h}", flush=True)
        img = _image(w, h, seed=call)
        dev = DevBuf(img)
        try:
            rc, got = _noalloc(r, dev.ptr, w, h)
        finally:
            dev.free()
        assert rc == 0, _err(r)
        want = r.noise_stats(img)
        assert _bits(got) == _bits(want), (w, h, got, want)
        if w > 3:
            assert got["valid_pixel_count"] > 0 and got["median_absolute"] > 0


def test_noalloc_second_smaller_call_sees_no_stale_scratch(ctx):
    """A large image leaves keys and partial sums in the shared array; the
    next, smaller image must give its own result (the allocating entry
    point on it alone is the check)."""
    r = ctx
    print("[noalloc] 2048x1536 then 100x37", flush=True)
    big = DevBuf(_image(2048, 1536, seed=11))
    small_img = _image(100, 37, seed=12)
    small = DevBuf(small_img)
    try:
        rc, _ = _noalloc(r, big.ptr, 2048, 1536)
        assert rc == 0, _err(r)
        rc, got = _noalloc(r, small.ptr, 100, 37)
        assert rc == 0, _err(r)
    finally:
        big.free()
        small.free()
    assert _bits(got) == _bits(r.noise_stats(small_img))


def test_noalloc_refusals(ctx):
    r = ctx
    print("[noalloc] refusals", flush=True)
    # beyond the static maximum: refused before any launch
    big = DevBuf(nbytes=4 * 3 * (MAX_PIXELS + 2048))
    try:
        rc, _ = _noalloc(r, big.ptr, 2048, 2049)
        assert rc == -1
    finally:
        big.free()
    # host memory
    host = np.zeros(3 * 64 * 64, np.float32)
    rc, _ = _noalloc(r, host.ctypes.data, 64, 64)
    assert rc == -1
    # null image
    rc, _ = _noalloc(r, 0, 64, 64)
    assert rc == -1
    # past the end of its allocation
    img = _image(64, 64, seed=3)
    dev = DevBuf(img)
    try:
        rc, _ = _noalloc(r, dev.ptr + 4, 64, 64)
        assert rc == -1
        # a valid call after the refusals: the whole allocation, accepted
        rc, got = _noalloc(r, dev.ptr, 64, 64)
        assert rc == 0, _err(r)
    finally:
        dev.free()
    assert _bits(got) == _bits(r.noise_stats(img))
