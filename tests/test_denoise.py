"""The repo denoiser on the host: the weights parse into the RT filter's
layers (tza.rs:139-147), tiles cover every pixel once as OIDN splits them
(tile.rs:104-129), PFM fixtures read as the reference reads them, and the
oracle restatement matches OIDN 2.5.1's own output on the desk fixture at the
reference's tolerance (tests/oidn.rs: 1e-3 * (1 + |v|))."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import denoise as D
from pfx_amd import pfm

GOLDEN = Path(__file__).resolve().parent / "golden" / "denoise"


def desk(part):
    return pfm.read((GOLDEN / f"desk.{part}.pfm").read_bytes())


def worst_gap(a, b):
    return float(np.max(np.abs(a[..., :3] - b[..., :3]) / (1.0 + np.abs(b[..., :3]))))


def test_the_shipped_weights_parse_with_their_layers(native_lib):
    data = D.WEIGHTS.read_bytes()
    assert D.layers(data) == (16, 919811)
    assert O.lib().or_tza_count(data, len(data)) == 32
    with pytest.raises(ValueError):
        D.layers(data[:100])
    with pytest.raises(ValueError):
        D.layers(b"PNG" + data[3:])


def covered(tiles, h, w):
    hits = np.zeros((h, w), np.int32)
    for t in tiles:
        (y, x), (th, tw) = t["out_dst"], t["out_size"]
        hits[y:y + th, x:x + tw] += 1
    return bool((hits == 1).all())


def test_tiles_cover_every_pixel_once_as_oidn_splits_them(native_lib):
    th, tw, tiles = D.plan(1080, 1920)
    assert (th, tw, len(tiles)) == (1088, 1920, 1) and covered(tiles, 1080, 1920)
    assert covered(D.plan(24, 32)[2], 24, 32)
    th, tw, tiles = D.plan(2160, 3840)
    assert (th, tw, len(tiles)) == (2160, 2016, 2) and covered(tiles, 2160, 3840)
    for h, w, m in ((1000, 3000, 1 << 20), (777, 1333, 600_000), (5000, 4000, 2160 * 2160)):
        th, tw, tiles = D.plan(h, w, m)
        assert covered(tiles, h, w) and len(tiles) > 1, (h, w)
        # the oracle's restatement plans the same tiles
        rows = (C.c_uint32 * (12 * len(tiles)))()
        a, b = C.c_uint32(), C.c_uint32()
        n = O.lib().or_denoise_plan(h, w, m, C.byref(a), C.byref(b), rows, len(tiles))
        assert (n, a.value, b.value) == (len(tiles), th, tw)
        flat = [v for t in tiles for k in ("src", "dst", "size", "out_src", "out_dst", "out_size") for v in t[k]]
        assert list(rows) == flat


def test_pfm_reads_rows_bottom_up_and_round_trips():
    color = desk("color")
    assert color.shape == (24, 32, 4) and (color[..., 3] == 1.0).all()
    assert np.array_equal(pfm.read(pfm.write(color)), color)


def test_oracle_matches_oidn_on_the_desk():
    color, albedo, normal, oidn = (desk(p) for p in ("color", "albedo", "normal", "oidn"))
    h, w = color.shape[:2]
    data = D.WEIGHTS.read_bytes()
    out = np.zeros_like(color)
    rc = O.lib().or_denoise(data, len(data), w, h, color.ctypes.data, albedo.ctypes.data, normal.ctypes.data,
                            out.ctypes.data, 2160 * 2160)
    assert rc == 0
    gap = worst_gap(out, oidn)
    print(f"oracle vs OIDN 2.5.1 on the desk: worst gap {gap:.2e}")
    assert gap <= 1e-3
