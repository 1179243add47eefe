"""OpenEXR film output (crates/still/src/picture.rs:524-620): round trips
through our reader for RGBA half (write_exr) and float (write_passes) files,
with and without RLE; the RLE coder against its specification on runs and
literals at the length limits and on a hand-worked example; the header's
required attributes and the chunk table. No EXR library is installed here
to read the files independently, so agreement with the `exr` crate's bytes
is unpinned; the layout follows the OpenEXR 2 specification."""
import struct

import numpy as np
import pytest

from pfx_amd import exr


def image(seed=1, h=7, w=13):
    rng = np.random.default_rng(seed)
    img = rng.exponential(1.0, (h, w, 4)).astype(np.float32)
    img[2:5, :, :3] = 0.25  # constant rows compress
    return img


@pytest.mark.parametrize("compression", [exr.NO_COMPRESSION, exr.RLE_COMPRESSION])
def test_float_round_trip(compression):
    img = image()
    out = exr.decode(exr.encode(img, exr.FLOAT, compression))
    assert out.tobytes() == img.tobytes()


def test_half_round_trip_and_alpha():
    img = image(2)
    data = exr.encode(img, exr.HALF)
    np.testing.assert_array_equal(exr.decode(data), img.astype(np.float16).astype(np.float32))


def test_rle_coder_runs_and_literals():
    cases = [b"", b"a", b"ab", b"aaa", b"a" * 300, bytes(range(256)) * 2, b"ab" * 200 + b"c" * 129 + b"xyz"]
    rng = np.random.default_rng(3)
    cases.append(rng.integers(0, 4, 1000).astype(np.uint8).tobytes())
    for raw in cases:
        if not raw:
            continue
        assert exr._rle_decode(exr._rle_encode(raw), len(raw)) == raw


def test_rle_known_answer():
    # 4 zero bytes: split (0,0 | 0,0), delta-coded to 0, 128, 128, 128; a
    # literal of one byte (count -1, 0x00) then a run of three 128s (count 2)
    assert exr._rle_encode(b"\0\0\0\0") == bytes([0xFF, 0x00, 2, 128])
    assert exr._rle_decode(bytes([0xFF, 0x00, 2, 128]), 4) == b"\0\0\0\0"


def test_header_and_chunk_table():
    img = image(4, h=5, w=6)
    data = exr.encode(img, exr.HALF, name="pfx render --recipe")
    magic, version = struct.unpack_from("<ii", data, 0)
    assert (magic, version) == (20000630, 2)
    for attr in (b"channels\0chlist\0", b"compression\0compression\0", b"dataWindow\0box2i\0",
                 b"displayWindow\0box2i\0", b"lineOrder\0lineOrder\0", b"pixelAspectRatio\0float\0",
                 b"screenWindowCenter\0v2f\0", b"screenWindowWidth\0float\0", b"name\0string\0"):
        assert attr in data
    table = data.index(b"screenWindowWidth\0float\0") + len(b"screenWindowWidth\0float\0") + 4 + 4 + 1
    offsets = struct.unpack_from("<5Q", data, table)
    for y, off in enumerate(offsets):
        assert struct.unpack_from("<i", data, off)[0] == y
    assert offsets[-1] + 8 + struct.unpack_from("<i", data, offsets[-1] + 4)[0] == len(data)


def test_passes_and_picture_files(tmp_path):
    c, a, n = image(5), image(6), image(7)
    paths = exr.write_passes(tmp_path, "still", c, a, n)
    assert [p.name for p in paths] == ["still.color.exr", "still.albedo.exr", "still.normal.exr"]
    for p, img in zip(paths, (c, a, n)):
        assert exr.read_exr(p).tobytes() == img.tobytes()
    exr.write_exr(tmp_path / "pic.exr", c, alpha=False)
    back = exr.read_exr(tmp_path / "pic.exr")
    np.testing.assert_array_equal(back[..., :3], c[..., :3].astype(np.float16).astype(np.float32))
    assert (back[..., 3] == 1.0).all()
    with pytest.raises(ValueError):
        exr.decode(b"\0" * 16)
