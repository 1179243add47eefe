"""`Learner`: the learned denoiser over libpfxhip -- a 5-level encoder-decoder
conv net (16 3x3 convolutions with skip connections, 32->128 channels,
1,239,651 parameters) trained on the GPU by the library's own forward and
backward kernels from pairs the tracer renders, BF16 operands on matrix
cores with FP32 sums.

The reference has no trained denoiser: it runs OIDN's fixed RT filter
(crates/denoise/src/lib.rs:15-17), which `pfx_amd.denoise.Denoiser`
mirrors and which stays the fallback. This network keeps the U-Net shape of
that filter (crates/denoise/src/net.rs:202-232) and adds a depth guide.

Weights files: b"PFXLNET1", then the parameter count as little-endian u64,
then the parameters as little-endian f32 in the canonical order (per layer
weights [out][in][3][3], then bias [out]).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _native as N

MAGIC = b"PFXLNET1"
WEIGHTS = Path(__file__).resolve().parent / "weights" / "learned_terrain.pfxl"


def layers() -> list[dict]:
    out = []
    for k in range(N.LEARN_LAYERS):
        info = N.LearnLayerC()
        if N.lib().pfxh_learn_layer(k, C.byref(info)) != N.OK:
            raise RuntimeError(f"no layer {k}")
        out.append({"name": info.name.decode(), "in0": info.in0, "in1": info.in1, "in_channels": info.in_channels,
                    "out_channels": info.out_channels, "level": info.level, "upsample": bool(info.upsample),
                    "pool": bool(info.pool), "relu": bool(info.relu)})
    return out


def param_count() -> int:
    return int(N.lib().pfxh_learn_param_count())


def init_params(seed: int) -> np.ndarray:
    """He-normal weights (std sqrt(2 / (9 * in))) and zero biases, from
    numpy's PCG64 with `seed`."""
    rng = np.random.default_rng(seed)
    parts = []
    for l in layers():
        ci, co = l["in_channels"], l["out_channels"]
        parts.append(rng.normal(0.0, (2.0 / (9.0 * ci)) ** 0.5, co * ci * 9))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


def save_params(path, params: np.ndarray) -> None:
    p = np.ascontiguousarray(params, "<f4")
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.uint64(len(p)).astype("<u8").tobytes())
        f.write(p.tobytes())


def load_params(path) -> np.ndarray:
    data = Path(path).read_bytes()
    if data[:8] != MAGIC or len(data) < 16:
        raise ValueError(f"{path}: not a learned-denoiser weights file")
    n = int(np.frombuffer(data[8:16], "<u8")[0])
    if n != param_count() or len(data) != 16 + 4 * n:
        raise ValueError(f"{path}: {n} parameters for a network of {param_count()}")
    return np.frombuffer(data[16:], "<f4").astype(np.float32)


def _img(a, h: int, w: int, channels: int | None = 4) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    want = (h, w, channels) if channels else (h, w)
    if a.shape != want:
        raise ValueError(f"image shape {a.shape}, expected {want}")
    return a


class Learner:
    """The network, its optimiser state and a device-resident training set."""

    def __init__(self, fp32: bool = False, params: np.ndarray | None = None, weights: str | Path | None = None):
        err = N.err_buffer()
        h = C.c_void_p()
        st = N.lib().pfxh_learner_new(N.LEARN_FP32 if fp32 else 0, C.byref(h), err, len(err))
        if st != N.OK:
            raise RuntimeError(err.value.decode(errors="replace"))
        self._h = h
        if weights is not None:
            params = load_params(weights)
        if params is not None:
            self.set_params(params)

    def _check(self, st: int, err) -> None:
        if st != N.OK:
            msg = err.value.decode(errors="replace")
            raise (ValueError if st == N.INVALID else RuntimeError)(msg)

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_learner_set_params(self._h, p.ctypes.data, len(p), err, len(err)), err)

    def params(self) -> np.ndarray:
        out = np.zeros(param_count(), np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_learner_params(self._h, out.ctypes.data, len(out), err, len(err)), err)
        return out

    def grads(self) -> np.ndarray:
        out = np.zeros(param_count(), np.float32)
        err = N.err_buffer()
        self._check(N.lib().pfxh_learner_grads(self._h, out.ctypes.data, len(out), err, len(err)), err)
        return out

    def save(self, path) -> None:
        save_params(path, self.params())

    def reset_optimizer(self) -> None:
        N.lib().pfxh_learner_reset_optimizer(self._h)

    def set_loss(self, mode: str) -> None:
        """'l1_log' (L1 on log radiance, the default) or 'rel_l2' (relative
        L2 of linear radiance, which tolerates noisy targets)."""
        code = {"l1_log": 0, "rel_l2": 1}[mode]
        if N.lib().pfxh_learner_set_loss(self._h, code) != N.OK:
            raise ValueError(mode)

    def set_passes(self, passes: int) -> None:
        """Inference passes over mirror images of the frame, averaged: 1, 2
        (also mirrored in x) or 4 (also in y, and in both)."""
        if N.lib().pfxh_learner_set_passes(self._h, int(passes)) != N.OK:
            raise ValueError(f"passes must be 1, 2 or 4, not {passes}")

    def reserve(self, side: int, capacity: int) -> None:
        err = N.err_buffer()
        self._check(N.lib().pfxh_learner_reserve(self._h, side, capacity, err, len(err)), err)
        self.side = side

    def add(self, color, albedo, normal, depth, target, deviation=None) -> None:
        """One training crop; depth and deviation (the standard error of each
        pixel's mean luminance, Trace.readback_deviation) are (side, side)
        or None."""
        s = self.side
        c, a, n, t = (_img(x, s, s) for x in (color, albedo, normal, target))
        d = None if depth is None else _img(depth, s, s, None)
        v = None if deviation is None else _img(deviation, s, s, None)
        err = N.err_buffer()
        st = N.lib().pfxh_learner_add(self._h, c.ctypes.data, a.ctypes.data, n.ctypes.data,
                                      None if d is None else d.ctypes.data, None if v is None else v.ctypes.data,
                                      t.ctypes.data, err, len(err))
        self._check(st, err)

    @property
    def crops(self) -> int:
        return int(N.lib().pfxh_learner_crops(self._h))

    def gradient(self, picks) -> float:
        p = np.ascontiguousarray(picks, np.uint32)
        loss = C.c_float()
        err = N.err_buffer()
        self._check(N.lib().pfxh_learner_gradient(self._h, p.ctypes.data, len(p), C.byref(loss), err, len(err)), err)
        return float(loss.value)

    def train(self, steps: int, batch: int = 16, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        err = N.err_buffer()
        st = N.lib().pfxh_learner_train(self._h, steps, batch, lr, seed, losses.ctypes.data, err, len(err))
        self._check(st, err)
        return losses[:steps]

    def denoise(self, color, albedo, normal, depth=None, deviation=None) -> np.ndarray:
        h, w = np.shape(color)[:2]
        c, a, n = (_img(x, h, w) for x in (color, albedo, normal))
        d = None if depth is None else _img(depth, h, w, None)
        v = None if deviation is None else _img(deviation, h, w, None)
        out = np.empty_like(c)
        err = N.err_buffer()
        st = N.lib().pfxh_learned_denoise(self._h, w, h, c.ctypes.data, a.ctypes.data, n.ctypes.data,
                                          None if d is None else d.ctypes.data, None if v is None else v.ctypes.data,
                                          out.ctypes.data, err, len(err))
        self._check(st, err)
        return out

    def close(self) -> None:
        if self._h is not None:
            N.lib().pfxh_learner_free(self._h)
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
    """`python -m pfx_amd.learn [WIDTH HEIGHT [RUNS]]`: trace the benchmark
    terrain at 4 spp and run the shipped learned denoiser on it RUNS times,
    printing the time of each run."""
    import sys
    import time

    from . import Bounces, Detail, Trace, scenes

    w, h = (int(sys.argv[1]), int(sys.argv[2])) if len(sys.argv) > 2 else (3840, 2160)
    runs = int(sys.argv[3]) if len(sys.argv) > 3 else 3
    scene = scenes.terrain(width=w, height=h)
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(6)), w, h) as t, Learner(weights=WEIGHTS) as L:
        t.sample(4, 1)
        for k in range(runs):
            s = time.perf_counter()
            img = t.learned_denoise(L)
            print(f"learned denoise {w}x{h} run {k}: {(time.perf_counter() - s) * 1e3:.1f} ms, "
                  f"finite {bool(np.isfinite(img).all())}", flush=True)


if __name__ == "__main__":
    _main()
