"""Image output: the reference's PNG conversion (src/main.cpp:25-52: clamp to
[0, 1], gamma 1/2.2, scale by 255.99 and truncate) with a dependency-free
PNG encoder, plus PFM for linear float images. The device computes the same
bytes for large images (HipRenderer.image_rgb8)."""
from __future__ import annotations

import struct
import zlib

import numpy as np


def to_srgb8(img_hw3: np.ndarray) -> np.ndarray:
    """Linear float (H, W, 3) -> uint8 as main.cpp:28-30 converts each
    channel: (uint8)(255.99f * powf(clamp(c, 0, 1), 1/2.2f)); NaN -> 0."""
    x = np.nan_to_num(np.asarray(img_hw3, np.float32), nan=0.0, posinf=1.0, neginf=0.0)
    x = np.clip(x, np.float32(0), np.float32(1))
    g = np.power(x.astype(np.float64), np.float64(np.float32(1.0) / np.float32(2.2))).astype(np.float32)
    return (np.float32(255.99) * g).astype(np.uint8)


def write_png_rows(path: str, rgb8: np.ndarray) -> None:
    """Writes (H, W, 3) uint8 rows as they are, first row at the top."""
    rgb = np.ascontiguousarray(rgb8, np.uint8)
    h, w, _ = rgb.shape
    raw = b"".join(b"\x00" + rgb[y].tobytes() for y in range(h))

    def chunk(tag, data):
        c = struct.pack(">I", len(data)) + tag + data
        return c + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def write_png(path: str, img_hw3: np.ndarray) -> None:
    """A linear float image with row 0 at the top (or uint8 rows) as PNG."""
    write_png_rows(path, to_srgb8(img_hw3) if img_hw3.dtype != np.uint8 else img_hw3)


def read_png_rgb8(path: str) -> np.ndarray:
    """Decodes an 8-bit RGB PNG without interlacing (what write_png_rows
    writes): (H, W, 3) uint8 in file row order."""
    with open(path, "rb") as f:
        data = f.read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG"
    pos, w, h, idat = 8, 0, 0, b""
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2, "only 8-bit RGB"
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    stride = 3 * w
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        if ft == 1:
            for x in range(3, stride):
                line[x] = (line[x] + line[x - 3]) & 255
        elif ft == 2:
            line = (line + prev) & 255
        elif ft == 3:
            for x in range(stride):
                line[x] = (line[x] + ((line[x - 3] if x >= 3 else 0) + prev[x]) // 2) & 255
        elif ft == 4:
            for x in range(stride):
                a = line[x - 3] if x >= 3 else 0
                b, c = prev[x], (prev[x - 3] if x >= 3 else 0)
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                line[x] = (line[x] + (a if pa <= pb and pa <= pc else (b if pb <= pc else c))) & 255
        out[y] = line
        prev = line
    return out.reshape(h, w, 3)


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
