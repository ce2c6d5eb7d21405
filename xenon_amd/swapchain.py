"""Progressive preview: the reference's triple-buffered swapchain
(render/swapchain.h:17-91) and the render-thread loop of main.cpp:143-165,
with an optional stop on the device-computed noise estimate.

The render thread writes frames into the write buffer and publishes them
with swap_writer(); a display thread takes the newest published frame with
get_read_buffer() without ever waiting for the renderer. Buffers hold the
running average in the swapchain layout (row 0 at the top), as
HipRenderer.render_frame returns it."""
from __future__ import annotations

import threading

import numpy as np


class TripleSwapchain:
    def __init__(self, width: int, height: int):
        self.width, self.height = int(width), int(height)
        self._buf = [np.zeros((self.height, self.width, 3), np.float32) for _ in range(3)]
        self._lock = threading.Lock()
        self._writer, self._pending, self._reader = 0, 1, 2
        self._dirty = False
        self.frames = 0

    def get_write_buffer(self) -> np.ndarray:
        """Render thread: the buffer to fill."""
        return self._buf[self._writer]

    def swap_writer(self) -> None:
        """Render thread: publish the filled buffer as the newest frame."""
        with self._lock:
            self._writer, self._pending = self._pending, self._writer
            self._dirty = True
            self.frames += 1

    def get_read_buffer(self) -> np.ndarray:
        """Display thread: the newest published frame (the same one again if
        nothing new was published)."""
        with self._lock:
            if self._dirty:
                self._reader, self._pending = self._pending, self._reader
                self._dirty = False
            return self._buf[self._reader]


def progressive(renderer, camera, chain: TripleSwapchain, max_spp: int, noise_target: float | None = None,
                check_every: int = 16, stop: threading.Event | None = None) -> dict:
    """Adds one sample per frame (the render_frame contract) and publishes
    each frame to `chain`, until `max_spp` samples, the `stop` event, or, with
    `noise_target`, until the running average's NoiseStats rms score (on the
    device, every `check_every` frames) falls to the target. Returns the last
    noise statistics (None if never computed) and the sample count."""
    stats = None
    while renderer.spp() < max_spp and not (stop is not None and stop.is_set()):
        renderer.render_frame(camera, chain)
        if noise_target is not None and renderer.spp() % check_every == 0:
            stats = renderer.noise_stats()
            if stats["valid_pixel_count"] > 0 and stats["rms_score"] <= noise_target:
                break
    return {"spp": renderer.spp(), "noise": stats}
