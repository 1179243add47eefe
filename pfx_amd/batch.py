"""`Batch`: many stills (own scene, size, sample count and seed each) rendered
with one launch per stage, the bounce loop replayed from a HIP graph, tile
and queue sizes from an autotuner. Each still's images equal those of its
own `Trace` at the same seed, byte for byte.

The reference renders one still per `Trace`; its only batched tracing is the
bake's probe batch (crates/bake/src/batch.rs:10-14,79-100), many camera
faces of one scene.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _native as N
from .scene import Detail, Scene
from .trace import Trace


@dataclass
class Still:
    scene: Scene
    width: int
    height: int
    spp: int
    seed: int


@dataclass
class Tuning:
    wave_paths: int
    refill: int
    trace_blocks_per_cu: int
    shade_blocks_per_cu: int
    graph: bool = True

    @staticmethod
    def from_c(t: N.BatchTuningC) -> "Tuning":
        return Tuning(t.wave_paths, t.refill, t.trace_blocks_per_cu, t.shade_blocks_per_cu, bool(t.graph))

    def to_c(self) -> N.BatchTuningC:
        return N.BatchTuningC(self.wave_paths, self.refill, self.trace_blocks_per_cu, self.shade_blocks_per_cu,
                              1 if self.graph else 0)


class _Lent(Trace):
    """A still of a batch seen as a trace; the batch owns it."""

    def __init__(self, handle, width: int, height: int, owner: "Batch"):
        super().__init__(handle, width, height)
        self._owner = owner

    def close(self) -> None:
        self._h = None


class Batch:
    def __init__(self, stills: list[Still], detail: Detail):
        lib = N.lib()
        keep, rows = [], (N.BatchStillC * len(stills))()
        for row, s in zip(rows, stills):
            sc, k = s.scene.to_c()
            keep.append((sc, k))
            row.scene = C.cast(C.pointer(sc), C.c_void_p)
            row.width, row.height, row.spp, row.seed = s.width, s.height, s.spp, s.seed
        d = detail.to_c()
        out = C.c_void_p()
        err = N.err_buffer()
        st = lib.pfxh_batch_new(rows, len(stills), C.byref(d), C.byref(out), err, len(err))
        del keep
        if st != N.OK:
            raise RuntimeError(err.value.decode(errors="replace"))
        self._h = out
        self.stills = list(stills)

    def _check(self, st: int, err) -> None:
        if st != N.OK:
            msg = err.value.decode(errors="replace")
            raise (ValueError if st == N.INVALID else RuntimeError)(msg)

    def render(self, counts: list[int] | None = None) -> None:
        err = N.err_buffer()
        if counts is None:
            self._check(N.lib().pfxh_batch_render(self._h, err, len(err)), err)
        else:
            arr = (C.c_uint32 * len(self.stills))(*counts)
            self._check(N.lib().pfxh_batch_render_samples(self._h, arr, err, len(err)), err)

    def tune(self) -> Tuning:
        t = N.BatchTuningC()
        err = N.err_buffer()
        self._check(N.lib().pfxh_batch_tune(self._h, C.byref(t), err, len(err)), err)
        return Tuning.from_c(t)

    @property
    def tuning(self) -> Tuning:
        t = N.BatchTuningC()
        N.lib().pfxh_batch_get_tuning(self._h, C.byref(t))
        return Tuning.from_c(t)

    @tuning.setter
    def tuning(self, value: Tuning) -> None:
        t = value.to_c()
        err = N.err_buffer()
        self._check(N.lib().pfxh_batch_set_tuning(self._h, C.byref(t), err, len(err)), err)

    def __len__(self) -> int:
        return int(N.lib().pfxh_batch_count(self._h))

    def trace(self, index: int) -> Trace:
        h = N.lib().pfxh_batch_trace(self._h, index)
        if not h:
            raise IndexError(index)
        s = self.stills[index]
        return _Lent(C.c_void_p(h), s.width, s.height, self)

    def stats(self) -> dict:
        s = N.StatsC()
        N.lib().pfxh_batch_stats(self._h, C.byref(s))
        return {name: getattr(s, name) for name, _ in N.StatsC._fields_}

    def reset_stats(self) -> None:
        N.lib().pfxh_batch_reset_stats(self._h)

    def close(self) -> None:
        if self._h is not None:
            N.lib().pfxh_batch_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _main() -> None:
    """`python -m pfx_amd.batch [SEED]`: build the 32-still uneven mix
    (scenes.batch_mix), tune the batch, render it twice and print the
    times."""
    import sys
    import time

    from . import Bounces, scenes

    seed = int(sys.argv[1]) if len(sys.argv) > 1 else 3
    mix = scenes.batch_mix(seed=seed)
    with Batch([Still(*m) for m in mix], Detail(bounces=Bounces.deep(6))) as b:
        t = time.perf_counter()
        tuning = b.tune()
        print(f"tuned in {time.perf_counter() - t:.1f} s: {tuning}", flush=True)
        for k in range(2):
            b.reset_stats()
            t = time.perf_counter()
            b.render()
            dt = time.perf_counter() - t
            s = b.stats()
            print(f"batch render {k}: {dt * 1e3:.0f} ms, {(s['closest_rays'] + s['shadow_rays']) / dt / 1e6:.0f} "
                  f"Mrays/s", flush=True)


if __name__ == "__main__":
    _main()
