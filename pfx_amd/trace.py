"""`Trace`: the still-render surface of the reference tracer over libpfxhip.

Mirrors `pfx_trace::Trace` (crates/trace/src/gpu.rs:309-342): build once
from a scene, `sample` more passes (they accumulate; every setter resets the
count), `readback` the colour/albedo/normal images (RGBA32F, row-major, the
same layout as `Output`, gpu.rs:405-410), `probe` primary hits. Errors are
`TraceError` with `refused` set for limits and memory, as
`TraceError::{Refused, Setup}` (gpu.rs:268-307); runtime failures raise
`RuntimeError`, the reference's `Err(String)`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .scene import Adaptive, Camera, Detail, Scene, Sun


class TraceError(Exception):
    def __init__(self, message: str, refused: bool):
        super().__init__(message)
        self.refused = refused


@dataclass
class Output:
    """gpu::Output: three (H, W, 4) float32 images."""

    color: np.ndarray
    albedo: np.ndarray
    normal: np.ndarray


@dataclass
class Round:
    samples: int
    active: int


@dataclass
class Converged:
    """adaptive::Converged (crates/trace/src/adaptive.rs:66-100)."""

    rounds: list
    pixels: int

    def max_samples(self) -> int:
        return self.rounds[-1].samples if self.rounds else 0

    def mean_samples(self) -> float:
        if self.pixels == 0:
            return 0.0
        total, before, running = 0.0, 0, self.pixels
        for r in self.rounds:
            total += float(r.samples - before) * running
            before, running = r.samples, r.active
        return total / self.pixels

    def converged(self) -> bool:
        return bool(self.rounds) and self.rounds[-1].active == 0


@dataclass
class Probe:
    """probe::Probe (crates/trace/src/probe.rs:10-18) plus primitive ids."""

    width: int
    height: int
    depth: np.ndarray
    distance: np.ndarray
    id: np.ndarray
    normal: np.ndarray
    primitive: np.ndarray
    instance: np.ndarray


class Trace:
    def __init__(self, handle, width: int, height: int):
        self._h = handle
        self.width = width
        self.height = height

    # Trace::new_detailed (crates/trace/src/gpu.rs:1926-1934)
    @classmethod
    def new_detailed(cls, scene: Scene, detail: Detail, width: int, height: int) -> "Trace":
        lib = N.lib()
        sc, keep = scene.to_c()
        d = detail.to_c()
        out = C.c_void_p()
        err = N.err_buffer()
        st = lib.pfxh_trace_new(C.byref(sc), C.byref(d), width, height, C.byref(out), err, len(err))
        del keep
        if st != N.OK:
            raise TraceError(err.value.decode(errors="replace"), refused=st == N.REFUSED)
        return cls(out, width, height)

    @classmethod
    def new(cls, scene: Scene, width: int, height: int) -> "Trace":
        return cls.new_detailed(scene, Detail(), width, height)

    def _check(self, st: int, err) -> None:
        if st != N.OK:
            raise RuntimeError(err.value.decode(errors="replace") or f"pfxhip status {st}")

    def _handle(self):
        if self._h is None:
            raise RuntimeError("the trace is closed")
        return self._h

    def sample(self, count: int, seed: int) -> None:
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_sample(self._handle(), count, seed, err, len(err)), err)

    def sample_adaptive(self, adaptive: Adaptive, seed: int) -> Converged:
        err = N.err_buffer()
        a = adaptive.to_c()
        out = N.ConvergedC()
        self._check(
            N.lib().pfxh_trace_sample_adaptive(self._handle(), C.byref(a), seed, C.byref(out), err, len(err)), err
        )
        rounds = [Round(int(out.samples[i]), int(out.active[i])) for i in range(out.round_count)]
        return Converged(rounds, int(out.pixels))

    def readback(self) -> Output:
        shape = (self.height, self.width, 4)
        color, albedo, normal = (np.empty(shape, np.float32) for _ in range(3))
        err = N.err_buffer()
        st = N.lib().pfxh_trace_readback(
            self._handle(), color.ctypes.data, albedo.ctypes.data, normal.ctypes.data, err, len(err)
        )
        self._check(st, err)
        return Output(color, albedo, normal)

    def readback_color(self) -> np.ndarray:
        color = np.empty((self.height, self.width, 4), np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_readback(self._handle(), color.ctypes.data, None, None, err, len(err)), err)
        return color

    def readback_deviation(self) -> np.ndarray:
        """The standard error of each pixel's mean luminance ((r + g + b) / 3),
        (H, W) float32: the learned denoiser's noise channel."""
        out = np.empty((self.height, self.width), np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_readback_deviation(self._handle(), out.ctypes.data, err, len(err)), err)
        return out

    def set_camera(self, camera: Camera) -> None:
        err = N.err_buffer()
        c = camera.to_c()
        self._check(N.lib().pfxh_trace_set_camera(self._handle(), C.byref(c), err, len(err)), err)

    def set_window(self, full: tuple, origin: tuple) -> None:
        err = N.err_buffer()
        st = N.lib().pfxh_trace_set_window(self._handle(), full[0], full[1], origin[0], origin[1], err, len(err))
        self._check(st, err)

    def set_sun(self, sun: Sun) -> None:
        err = N.err_buffer()
        s = sun.to_c()
        self._check(N.lib().pfxh_trace_set_sun(self._handle(), C.byref(s), err, len(err)), err)

    def set_sun_radius(self, radius_deg: float) -> None:
        """Trace::set_sun_radius (gpu.rs:2903-2910); resets the sample count."""
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_set_sun_radius(self._handle(), float(radius_deg), err, len(err)), err)

    def set_environment(self, sky) -> None:
        """Trace::set_environment (gpu.rs:2810-2815): an (H, W, 4) float32
        lat-long sky, or None for black; resets the sample count."""
        err = N.err_buffer()
        if sky is None or not np.asarray(sky).size:
            env = N.Environment(0, 0, None)
            keep = None
        else:
            keep = np.ascontiguousarray(sky, np.float32)
            if keep.ndim != 3 or keep.shape[2] != 4:
                raise ValueError(f"a sky is an (H, W, 4) array, not {keep.shape}")
            env = N.Environment(keep.shape[1], keep.shape[0], keep.ctypes.data)
        self._check(N.lib().pfxh_trace_set_environment(self._handle(), C.byref(env), err, len(err)), err)
        del keep

    def set_lens(self, shift=(0.0, 0.0), aperture: float = 0.0, focus_distance: float = 0.0) -> None:
        err = N.err_buffer()
        st = N.lib().pfxh_trace_set_lens(self._handle(), shift[0], shift[1], aperture, focus_distance, err, len(err))
        self._check(st, err)

    def set_projection(self, kind: int, width: float = 0.0, height: float = 0.0) -> None:
        """Trace::set_projection (gpu.rs:2872-2890): PERSPECTIVE, ORTHOGRAPHIC
        (view width/height) or EQUIRECT; resets the sample count."""
        err = N.err_buffer()
        st = N.lib().pfxh_trace_set_projection(self._handle(), int(kind), float(width), float(height), err, len(err))
        self._check(st, err)

    def save(self, key: str) -> bytes:
        """Trace::save + Saved::write_to (checkpoint.rs:85-110, 163-174): the
        PFXCKPT1 bytes of this trace's accumulator, images and sample count."""
        err = N.err_buffer()
        length = C.c_size_t()
        k = key.encode()
        self._check(N.lib().pfxh_trace_save(self._handle(), k, None, 0, C.byref(length), err, len(err)), err)
        out = C.create_string_buffer(length.value)
        self._check(N.lib().pfxh_trace_save(self._handle(), k, out, length.value, C.byref(length), err, len(err)), err)
        return out.raw[: length.value]

    def restore(self, key: str, data: bytes) -> None:
        """Saved::read_from + Trace::restore (checkpoint.rs:112-200)."""
        err = N.err_buffer()
        buf = C.create_string_buffer(data, len(data))
        self._check(N.lib().pfxh_trace_restore(self._handle(), key.encode(), buf, len(data), err, len(err)), err)

    def denoise(self, denoiser) -> np.ndarray:
        """The repo denoiser on this trace's colour, albedo and normal, on the
        device (Denoiser::denoise_textures): the denoised (H, W, 4) image."""
        out = np.empty((self.height, self.width, 4), np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_denoise(self._handle(), denoiser._h, out.ctypes.data, err, len(err)), err)
        return out

    def learned_denoise(self, learner) -> np.ndarray:
        """The learned denoiser (`pfx_amd.learn.Learner`) on this trace's colour,
        albedo, normal and probe depth, on the device: the (H, W, 4) image."""
        out = np.empty((self.height, self.width, 4), np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_learned_denoise(self._handle(), learner._h, out.ctypes.data, err, len(err)),
                    err)
        return out

    def learned_denoise_checked(self, learner, denoiser, threshold: float, tile: int = 32) -> tuple[np.ndarray, int]:
        """The learned denoiser with the repo denoiser as per-tile fallback:
        tiles whose removed variance (mean squared log difference between the
        noisy and the learned image) exceeds `threshold` take the repo
        denoiser's output. Returns the image and the number of such tiles."""
        out = np.empty((self.height, self.width, 4), np.float32)
        n = C.c_uint32()
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_learned_denoise_checked(self._handle(), learner._h, denoiser._h,
                                                               float(threshold), int(tile), out.ctypes.data,
                                                               C.byref(n), err, len(err)), err)
        return out, int(n.value)

    def atrous(self) -> np.ndarray:
        """denoise::filter (crates/trace/src/denoise.rs:1-62) on this trace's
        colour, albedo and normal at its sample count, on the device."""
        out = np.empty((self.height, self.width, 4), np.float32)
        err = N.err_buffer()
        st = N.lib().pfxh_trace_atrous(self._handle(), out.ctypes.data, err, len(err))
        if st == N.INVALID:
            raise ValueError(err.value.decode(errors="replace"))
        self._check(st, err)
        return out

    def samples(self) -> int:
        return int(N.lib().pfxh_trace_samples(self._handle()))

    def gpu_bytes(self) -> int:
        return int(N.lib().pfxh_trace_gpu_bytes(self._handle()))

    def node_bytes(self) -> int:
        """Bytes a closest-hit node visit fetches: 64 (quantized nodes) or
        112 (4-wide FP32 nodes)."""
        return int(N.lib().pfxh_trace_node_bytes(self._handle()))

    def cast(self, origins, directions, limit: float = 1e20):
        """Closest hits of caller rays: (caller triangle ids, UINT32_MAX on a
        miss; distances, 0 on a miss)."""
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        if o.shape != d.shape or o.size % 3 != 0:
            raise ValueError(f"origins {o.shape} and directions {d.shape} must be equal (n, 3) arrays")
        o, d = o.reshape(-1, 3), d.reshape(-1, 3)
        if len(o) >= 2**32:
            raise ValueError(f"{len(o)} rays do not fit a uint32 count")
        ids = np.zeros(len(o), np.uint32)
        t = np.zeros(len(o), np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_cast(self._handle(), o.ctypes.data, d.ctypes.data, len(o), limit,
                                            ids.ctypes.data, t.ctypes.data, err, len(err)), err)
        return ids, t

    def probe(self) -> Probe:
        hits = np.zeros(self.width * self.height, dtype=N.PROBE_HIT)
        err = N.err_buffer()
        self._check(N.lib().pfxh_trace_probe(self._handle(), hits.ctypes.data, err, len(err)), err)
        shape = (self.height, self.width)
        return Probe(
            self.width,
            self.height,
            hits["depth"].reshape(shape),
            hits["distance"].reshape(shape),
            hits["material_id"].reshape(shape),
            hits["normal"].reshape(shape + (3,)),
            hits["primitive"].reshape(shape),
            hits["instance"].reshape(shape),
        )

    def stats(self) -> dict:
        s = N.StatsC()
        st = N.lib().pfxh_trace_stats(self._handle(), C.byref(s))
        if st != N.OK:
            raise RuntimeError(f"pfxhip stats failed ({st})")
        return {name: getattr(s, name) for name, _ in N.StatsC._fields_}

    def reset_stats(self) -> None:
        N.lib().pfxh_trace_reset_stats(self._handle())

    def set_counting(self, on: bool) -> None:
        N.lib().pfxh_trace_set_counting(self._handle(), 1 if on else 0)

    def close(self) -> None:
        if self._h is not None:
            N.lib().pfxh_trace_free(self._h)
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


def hbm_probe(nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Achievable HBM bandwidth in GB/s: the best of `reps` device-to-device
    streaming copies of `nbytes` (read + write bytes over the copy time)."""
    err = N.err_buffer()
    gbs = C.c_double()
    st = N.lib().pfxh_hbm_probe(nbytes, reps, C.byref(gbs), err, len(err))
    if st != 0:
        raise TraceError(err.value.decode(errors="replace"))
    return float(gbs.value)


def animated_seed(seed: int, frame: int) -> int:
    """sobol::animated_seed (crates/trace/src/sobol.rs:105-107): the seed an
    animation rendered with `seed` passes to `sample` for frame `frame`
    (still::Seed::Animated, crates/still/src/trace.rs:77-82)."""
    return int(N.lib().pfxh_sobol_animated_seed(int(seed) & 0xFFFFFFFF, int(frame) & 0xFFFFFFFF))
