"""OpenEXR film output, as the still writes it (crates/still/src/picture.rs):
`write_exr` (picture.rs:524-549) stores a picture as RGBA half floats,
`write_passes` (:551-587) the colour, albedo and normal passes as RGBA
32-bit floats, each a single-part scan-line file with run-length
compression (the `exr` crate's FAST_LOSSLESS encoding), and `read_exr`
(:591-620) reads the first RGBA layer back.

The file layout follows the OpenEXR 2 specification (single part, scan
lines, one line per chunk for RLE): magic 20000630, version 2, the header
attributes (channels, compression, dataWindow, displayWindow, lineOrder,
pixelAspectRatio, screenWindowCenter, screenWindowWidth, name), the chunk
offset table, then per line its y, its byte count and the channels in
name order (A, B, G, R) as the line's samples. RLE chunks are byte-split
into even and odd positions, delta-coded (+128) and run-length coded
(runs of 3..128 as count-1 and the byte; literals as -count and the
bytes); a chunk that would not shrink is stored raw.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

MAGIC = 20000630
HALF, FLOAT = 1, 2
NO_COMPRESSION, RLE_COMPRESSION = 0, 1
CHANNELS = ("A", "B", "G", "R")  # name order; R, G, B, A = 0..3 of a pixel
_INDEX = {"R": 0, "G": 1, "B": 2, "A": 3}


def _attr(name: str, kind: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value


def _rle_encode(data: bytes) -> bytes:
    n = len(data)
    # split into even and odd bytes, then delta-code
    t = bytearray(data[0::2] + data[1::2])
    if n > 1:
        prev = t[0]
        for i in range(1, n):
            cur = t[i]
            t[i] = (cur - prev + 128 + 256) & 0xFF
            prev = cur
    out = bytearray()
    start = 0
    while start < n:
        end = start + 1
        while end < n and t[end] == t[start] and end - start - 1 < 127:
            end += 1
        if end - start >= 3:
            out.append(end - start - 1)
            out.append(t[start])
            start = end
            continue
        # literal run until the next run of three or 127 bytes
        end = start
        while end < n and end - start < 127:
            if end + 2 < n and t[end] == t[end + 1] == t[end + 2]:
                break
            end += 1
        out.append((start - end) & 0xFF)
        out += t[start:end]
        start = end
    return bytes(out)


def _rle_decode(data: bytes, size: int) -> bytes:
    t = bytearray()
    i = 0
    while i < len(data):
        count = struct.unpack("<b", data[i:i + 1])[0]
        i += 1
        if count < 0:
            t += data[i:i - count]
            i -= count
        else:
            t += bytes([data[i]]) * (count + 1)
            i += 1
    if len(t) != size:
        raise ValueError("corrupt RLE chunk")
    for k in range(1, size):
        t[k] = (t[k - 1] + t[k] - 128) & 0xFF
    half = (size + 1) // 2
    out = bytearray(size)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return bytes(out)


def encode(pixels: np.ndarray, pixel_type: int = FLOAT, compression: int = RLE_COMPRESSION,
           name: str | None = None) -> bytes:
    """An (H, W, 4) RGBA image as an OpenEXR file."""
    pixels = np.asarray(pixels)
    h, w = pixels.shape[:2]
    if pixels.shape != (h, w, 4) or h == 0 or w == 0:
        raise ValueError("an EXR picture is a nonempty (H, W, 4) array")
    dtype = np.dtype("<f2") if pixel_type == HALF else np.dtype("<f4")
    chlist = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", pixel_type, 0, 1, 1) for c in CHANNELS) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = struct.pack("<ii", MAGIC, 2)
    header += _attr("channels", "chlist", chlist)
    header += _attr("compression", "compression", bytes([compression]))
    header += _attr("dataWindow", "box2i", box)
    header += _attr("displayWindow", "box2i", box)
    header += _attr("lineOrder", "lineOrder", b"\0")
    if name is not None:
        header += _attr("name", "string", name.encode())
    header += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    planes = [pixels[..., _INDEX[c]].astype(dtype) for c in CHANNELS]
    chunks = []
    for y in range(h):
        raw = b"".join(p[y].tobytes() for p in planes)
        data = raw
        if compression == RLE_COMPRESSION:
            packed = _rle_encode(raw)
            if len(packed) < len(raw):
                data = packed
        chunks.append(struct.pack("<ii", y, len(data)) + data)
    offsets, at = [], len(header) + 8 * h
    for c in chunks:
        offsets.append(at)
        at += len(c)
    return header + struct.pack(f"<{h}Q", *offsets) + b"".join(chunks)


def decode(data: bytes) -> np.ndarray:
    """The first RGBA layer of a single-part scan-line OpenEXR file (no or
    RLE compression) as (H, W, 4) float32; missing channels read as 0
    (alpha as 1)."""
    magic, version = struct.unpack_from("<ii", data, 0)
    if magic != MAGIC or (version & 0xFF) != 2 or version & 0x200:
        raise ValueError("not a single-part scan-line OpenEXR file")
    at, attrs = 8, {}
    while data[at] != 0:
        end = data.index(b"\0", at)
        name = data[at:end].decode()
        end2 = data.index(b"\0", end + 1)
        kind = data[end + 1:end2].decode()
        size = struct.unpack_from("<i", data, end2 + 1)[0]
        attrs[name] = (kind, data[end2 + 5:end2 + 5 + size])
        at = end2 + 5 + size
    at += 1
    channels, cl = [], attrs["channels"][1]
    i = 0
    while cl[i] != 0:
        end = cl.index(b"\0", i)
        ptype = struct.unpack_from("<i", cl, end + 1)[0]
        channels.append((cl[i:end].decode(), ptype))
        i = end + 17
    compression = attrs["compression"][1][0]
    if compression not in (NO_COMPRESSION, RLE_COMPRESSION):
        raise ValueError(f"EXR compression {compression} is not read here")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", data, at)
    out = np.zeros((h, w, 4), np.float32)
    out[..., 3] = 1.0
    sizes = [w * (2 if p == HALF else 4) for _, p in channels]
    for off in offsets:
        y, n = struct.unpack_from("<ii", data, off)
        chunk = data[off + 8:off + 8 + n]
        raw = chunk if n == sum(sizes) or compression == NO_COMPRESSION else _rle_decode(chunk, sum(sizes))
        pos = 0
        for (cname, ptype), sz in zip(channels, sizes):
            vals = np.frombuffer(raw[pos:pos + sz], "<f2" if ptype == HALF else "<f4")
            if cname in _INDEX:
                out[y - y0, :, _INDEX[cname]] = vals
            pos += sz
    return out


def write_exr(path, pixels: np.ndarray, alpha: bool = True) -> None:
    """picture::write_exr: RGBA half floats (alpha forced to 1 unless kept)."""
    p = np.array(pixels, np.float32)
    if not alpha:
        p[..., 3] = 1.0
    Path(path).write_bytes(encode(p, HALF, name="pfx render --recipe"))


def write_passes(directory, name: str, color, albedo, normal) -> list[Path]:
    """picture::write_passes: NAME.color/.albedo/.normal.exr as RGBA 32-bit floats."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    paths = []
    for pass_name, img in (("color", color), ("albedo", albedo), ("normal", normal)):
        path = d / f"{name}.{pass_name}.exr"
        path.write_bytes(encode(np.asarray(img, np.float32), FLOAT, name=pass_name))
        paths.append(path)
    return paths


def read_exr(path) -> np.ndarray:
    """picture::read_exr: the first RGBA layer as (H, W, 4) float32."""
    return decode(Path(path).read_bytes())
