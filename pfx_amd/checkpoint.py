"""PFXCKPT1 checkpoints (crates/trace/src/checkpoint.rs) over libpfxhip.

`Saved` mirrors `checkpoint::Saved` (checkpoint.rs:13-28): the frame size,
sample count, the 48-byte-per-pixel accumulator, the three RGBA32F images
and, for adaptive renders, the reached rounds and 32-byte-per-pixel
snapshots. `write_to`/`read_from` produce and parse the same bytes as the
reference (checkpoint.rs:85-158), through the library's host-only format
functions; `Trace.save`/`Trace.restore` move a trace's state in and out.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

from . import _native as N

ACCUM, TEXEL, SNAP = 48, 16, 32


@dataclass
class Rounds:
    """checkpoint::Rounds: reached (samples, active) pairs and snapshots."""

    reached: list = field(default_factory=list)
    snaps: bytes = b""


@dataclass
class Saved:
    size: tuple
    samples: int
    accum: bytes
    color: bytes
    albedo: bytes
    normal: bytes
    adaptive: Optional[Rounds] = None

    def write_to(self, key: str) -> bytes:
        """Saved::write_to (checkpoint.rs:85-110)."""
        pixels = self.size[0] * self.size[1]
        fits = (len(self.accum) == pixels * ACCUM and all(len(b) == pixels * TEXEL for b in (self.color, self.albedo, self.normal))
                and (self.adaptive is None or len(self.adaptive.snaps) == pixels * SNAP))
        if not fits:
            raise ValueError("the checkpoint's buffers do not match its size")
        keep = [C.create_string_buffer(b, len(b)) for b in (self.accum, self.color, self.albedo, self.normal)]
        c = N.CheckpointC()
        c.width, c.height = self.size
        c.samples = self.samples
        c.accum, c.color, c.albedo, c.normal = [C.addressof(b) for b in keep]
        c.round_count = N.NO_ROUNDS
        if self.adaptive is not None:
            n = len(self.adaptive.reached)
            rs = (C.c_uint32 * max(n, 1))(*[r[0] for r in self.adaptive.reached])
            ra = (C.c_uint64 * max(n, 1))(*[r[1] for r in self.adaptive.reached])
            sn = C.create_string_buffer(self.adaptive.snaps, len(self.adaptive.snaps))
            keep += [rs, ra, sn]
            c.round_count = n
            c.round_samples, c.round_active, c.snaps = C.addressof(rs), C.addressof(ra), C.addressof(sn)
        length = C.c_size_t()
        err = N.err_buffer()
        k = key.encode()
        st = N.lib().pfxh_checkpoint_write(C.byref(c), k, None, 0, C.byref(length), err, len(err))
        if st != N.OK:
            raise ValueError(err.value.decode(errors="replace"))
        out = C.create_string_buffer(length.value)
        st = N.lib().pfxh_checkpoint_write(C.byref(c), k, out, length.value, C.byref(length), err, len(err))
        if st != N.OK:
            raise ValueError(err.value.decode(errors="replace"))
        return out.raw[: length.value]

    @staticmethod
    def read_from(key: str, data: bytes) -> "Saved":
        """Saved::read_from (checkpoint.rs:112-158)."""
        buf = C.create_string_buffer(data, len(data))
        c = N.CheckpointC()
        err = N.err_buffer()
        cap = 4096
        rs = (C.c_uint32 * cap)()
        ra = (C.c_uint64 * cap)()
        st = N.lib().pfxh_checkpoint_read(buf, len(data), key.encode(), C.byref(c), rs, ra, cap, err, len(err))
        if st != N.OK:
            raise ValueError(err.value.decode(errors="replace"))
        base = C.addressof(buf)
        pixels = c.width * c.height

        def block(ptr, n):
            return bytes(data[ptr - base: ptr - base + n]) if n else b""

        adaptive = None
        if c.round_count != N.NO_ROUNDS:
            if c.round_count > cap:
                raise ValueError("the checkpoint holds more adaptive rounds than this reader keeps")
            adaptive = Rounds([(int(rs[i]), int(ra[i])) for i in range(c.round_count)], block(c.snaps, pixels * SNAP))
        return Saved((c.width, c.height), c.samples, block(c.accum, pixels * ACCUM), block(c.color, pixels * TEXEL),
                     block(c.albedo, pixels * TEXEL), block(c.normal, pixels * TEXEL), adaptive)
