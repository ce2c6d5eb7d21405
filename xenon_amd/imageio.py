"""Image output: the reference's PNG conversion (src/main.cpp:25-52: clamp to
[0, 1], gamma 1/2.2, 8-bit) with a dependency-free PNG encoder, plus PFM
for linear float images."""
from __future__ import annotations

import struct
import zlib

import numpy as np


def to_srgb8(img_hw3: np.ndarray) -> np.ndarray:
    """Linear float (H, W, 3), row 0 at the top -> uint8 with gamma 1/2.2."""
    x = np.nan_to_num(np.asarray(img_hw3, np.float32), nan=0.0, posinf=1.0, neginf=0.0)
    x = np.clip(x, 0.0, 1.0) ** (1.0 / 2.2)
    return (x * 255.0 + 0.5).astype(np.uint8)


def write_png(path: str, img_hw3: np.ndarray) -> None:
    rgb = to_srgb8(img_hw3) if img_hw3.dtype != np.uint8 else img_hw3
    h, w, _ = rgb.shape
    raw = b"".join(b"\x00" + rgb[y].tobytes() for y in range(h))

    def chunk(tag, data):
        c = struct.pack(">I", len(data)) + tag + data
        return c + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def write_pfm(path: str, img_hw3: np.ndarray) -> None:
    """Linear float RGB PFM (rows stored bottom-up, as the format requires)."""
    a = np.ascontiguousarray(np.asarray(img_hw3, np.float32)[::-1])
    h, w, _ = a.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(a.astype("<f4").tobytes())


def pixel_order_to_image(flat: np.ndarray, width: int, height: int) -> np.ndarray:
    """(H*W, 3) in pixel order (y = 0 bottom) -> (H, W, 3) with row 0 at the top."""
    return np.asarray(flat).reshape(height, width, 3)[::-1]
