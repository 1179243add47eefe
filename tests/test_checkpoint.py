"""PFXCKPT1 checkpoint bytes (crates/trace/src/checkpoint.rs), host only.

Mirrors the reference's own test (checkpoint.rs:219-279): a checkpoint reads
back as written only under its own key, a cut or over-long checkpoint is
refused, mismatched buffers are refused on write, and other files are not
checkpoints. The byte layout itself is pinned against a hand-assembled
checkpoint built from the format's definition."""
import struct

import pytest

from pfx_amd.checkpoint import ACCUM, SNAP, TEXEL, Rounds, Saved


def saved(adaptive: bool) -> Saved:
    size = (3, 2)
    pixels = size[0] * size[1]

    def fill(per, seed):
        return bytes(((n * 7) + seed) & 0xFF for n in range(pixels * per))

    return Saved(size, 21, fill(ACCUM, 1), fill(TEXEL, 2), fill(TEXEL, 3), fill(TEXEL, 4),
                 Rounds([(8, 6), (16, 2)], fill(SNAP, 5)) if adaptive else None)


@pytest.mark.parametrize("adaptive", [False, True])
def test_a_checkpoint_reads_back_as_written_only_under_its_own_key(native_lib, adaptive):
    s = saved(adaptive)
    data = s.write_to("clip frame 4")
    assert Saved.read_from("clip frame 4", data) == s
    with pytest.raises(ValueError, match="another render"):
        Saved.read_from("clip frame 5", data)
    with pytest.raises(ValueError, match="cut short"):
        Saved.read_from("clip frame 4", data[:-1])
    with pytest.raises(ValueError):
        Saved.read_from("clip frame 4", data + b"\0")


def test_mismatched_buffers_and_foreign_files_are_refused(native_lib):
    wrong = saved(False)
    wrong.color = wrong.color[:-1]
    with pytest.raises(ValueError):
        wrong.write_to("k")
    with pytest.raises(ValueError, match="not a pfx checkpoint"):
        Saved.read_from("k", b"PNG.....")
    with pytest.raises(ValueError, match="cut short"):
        Saved.read_from("k", b"PFX")


def test_byte_layout(native_lib):
    """checkpoint.rs:85-110: magic, u64 key length, key, u32 w/h/samples,
    accum, color, albedo, normal, then u32::MAX or the rounds and snaps, all
    little-endian."""
    for adaptive in (False, True):
        s = saved(adaptive)
        key = "still"
        expect = b"PFXCKPT1" + struct.pack("<Q", len(key)) + key.encode() + struct.pack("<III", 3, 2, 21)
        expect += s.accum + s.color + s.albedo + s.normal
        if adaptive:
            expect += struct.pack("<I", 2) + struct.pack("<IQ", 8, 6) + struct.pack("<IQ", 16, 2) + s.adaptive.snaps
        else:
            expect += struct.pack("<I", 0xFFFFFFFF)
        assert s.write_to(key) == expect
