"""PFM images (crates/denoise/src/pfm.rs): little-endian float RGB or grey,
rows stored bottom to top. Images here are (H, W, 4) float32 with alpha 1."""
from __future__ import annotations

import numpy as np


def read(data: bytes) -> np.ndarray:
    """pfm::read (pfm.rs:7-56)."""
    fields, at = [], 0
    while len(fields) < 4:
        while at < len(data) and data[at:at + 1].isspace():
            at += 1
        start = at
        while at < len(data) and not data[at:at + 1].isspace():
            at += 1
        if start == at:
            raise ValueError("invalid PFM image")
        fields.append(data[start:at].decode("ascii"))
    at += 1
    channels = {"PF": 3, "Pf": 1}.get(fields[0])
    if channels is None:
        raise ValueError("invalid PFM image")
    width, height, scale = int(fields[1]), int(fields[2]), float(fields[3])
    if scale >= 0.0:
        raise ValueError("big-endian PFM images are not supported")
    count = width * height * channels
    if len(data) < at + count * 4:
        raise ValueError("invalid PFM image: error reading pixel data")
    values = np.frombuffer(data, "<f4", count, at).reshape(height, width, channels) * np.float32(abs(scale))
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = values[::-1][..., [0, min(1, channels - 1), min(2, channels - 1)]]
    return out


def write(image: np.ndarray) -> bytes:
    """pfm::write (pfm.rs:58-69): RGB, rows bottom to top."""
    height, width = image.shape[:2]
    head = f"PF\n{width} {height}\n-1.0\n".encode()
    return head + np.ascontiguousarray(image[::-1, :, :3], "<f4").tobytes()
