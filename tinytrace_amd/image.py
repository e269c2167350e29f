"""Image output for the display pass: RGBA8 arrays to PNG files (the
viewer's render-to-image step), with the standard library only."""
from __future__ import annotations

import os
import struct
import zlib

import numpy as np


def _chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def encode_png(rgba8: np.ndarray) -> bytes:
    """PNG bytes of an (H, W, 4) uint8 image: 8-bit RGBA, no interlace,
    filter 0 on every row."""
    img = np.ascontiguousarray(rgba8, dtype=np.uint8)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) uint8 image")
    h, w, _ = img.shape
    rows = np.concatenate([np.zeros((h, 1), dtype=np.uint8), img.reshape(h, w * 4)], axis=1)
    header = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", header) + _chunk(b"IDAT", zlib.compress(rows.tobytes(), 6)) +
            _chunk(b"IEND", b""))


def save_png(path, rgba8: np.ndarray) -> None:
    with open(os.fspath(path), "wb") as f:
        f.write(encode_png(rgba8))


def render_to_image(renderer, path) -> np.ndarray:
    """Render one pass, run the display pass and write it as PNG; returns the
    RGBA8 image."""
    renderer.render_to_target()
    image = renderer.display()
    save_png(path, image)
    return image
