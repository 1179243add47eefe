"""`Denoiser`: the repo denoiser (crates/denoise, OIDN 2.5.1's RT filter) over
libpfxhip. Mirrors `pfx_denoise::Denoiser` (crates/denoise/src/lib.rs:65-144):
built once from OIDN's RT weights (pfx_amd/weights/rt_hdr_alb_nrm.tza,
Apache-2.0, see pfx_amd/weights/NOTICE), then `denoise` colour with its
albedo and normal guides; `Trace.denoise` runs it on a trace's images on the
device."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _native as N

WEIGHTS = Path(__file__).resolve().parent / "weights" / "rt_hdr_alb_nrm.tza"


def _f32(image: np.ndarray, height: int, width: int) -> np.ndarray:
    a = np.ascontiguousarray(image, np.float32)
    if a.shape != (height, width, 4):
        raise ValueError("image size mismatch")
    return a


class Denoiser:
    def __init__(self, weights: bytes | None = None):
        data = weights if weights is not None else WEIGHTS.read_bytes()
        err = N.err_buffer()
        h = C.c_void_p()
        st = N.lib().pfxh_denoiser_new(data, len(data), C.byref(h), err, len(err))
        if st != N.OK:
            raise ValueError(err.value.decode(errors="replace"))
        self._h = h

    def set_max_tile(self, pixels: int) -> None:
        """Largest tile in pixels (0: one tile per image)."""
        N.lib().pfxh_denoiser_set_max_tile(self._h, int(pixels))

    def denoise(self, color: np.ndarray, albedo: np.ndarray, normal: np.ndarray) -> np.ndarray:
        """Denoiser::denoise: (H, W, 4) float32 images in, the denoised colour out."""
        height, width = color.shape[:2]
        c, a, n = (_f32(x, height, width) for x in (color, albedo, normal))
        out = np.empty_like(c)
        err = N.err_buffer()
        st = N.lib().pfxh_denoise(self._h, width, height, c.ctypes.data, a.ctypes.data, n.ctypes.data,
                                  out.ctypes.data, err, len(err))
        if st != N.OK:
            raise RuntimeError(err.value.decode(errors="replace"))
        return out

    def close(self) -> None:
        if self._h is not None:
            N.lib().pfxh_denoiser_free(self._h)
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


def layers(weights: bytes | None = None) -> tuple[int, int]:
    """Layer and parameter counts of a weights blob (host only)."""
    data = weights if weights is not None else WEIGHTS.read_bytes()
    count, params = C.c_uint32(), C.c_uint64()
    err = N.err_buffer()
    st = N.lib().pfxh_denoiser_layers(data, len(data), C.byref(count), C.byref(params), err, len(err))
    if st != N.OK:
        raise ValueError(err.value.decode(errors="replace"))
    return count.value, params.value


def plan(height: int, width: int, max_tile_pixels: int = 2160 * 2160):
    """tile::plan (crates/denoise/src/tile.rs:31-101): (tile_h, tile_w, tiles)."""
    th, tw, count = C.c_uint32(), C.c_uint32(), C.c_uint32()
    N.lib().pfxh_denoise_plan(height, width, max_tile_pixels, C.byref(th), C.byref(tw), None, 0, C.byref(count))
    tiles = (N.DenoiseTile * max(count.value, 1))()
    N.lib().pfxh_denoise_plan(height, width, max_tile_pixels, C.byref(th), C.byref(tw), tiles, count.value,
                              C.byref(count))
    return th.value, tw.value, [{k: tuple(getattr(t, k)) for k, _ in N.DenoiseTile._fields_} for t in tiles[:count.value]]


def _main() -> None:
    """`python -m pfx_amd.denoise [WIDTH HEIGHT [RUNS]]`: denoise a rendered
    terrain frame (4 spp) RUNS times and print the time of each run."""
    import sys
    import time

    from . import Bounces, Detail, Trace, scenes

    w, h = (int(sys.argv[1]), int(sys.argv[2])) if len(sys.argv) > 2 else (3840, 2160)
    runs = int(sys.argv[3]) if len(sys.argv) > 3 else 3
    scene = scenes.terrain(width=w, height=h)
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(6)), w, h) as t, Denoiser() as d:
        t.sample(4, 1)
        for k in range(runs):
            s = time.perf_counter()
            img = t.denoise(d)
            print(f"denoise {w}x{h} run {k}: {(time.perf_counter() - s) * 1e3:.1f} ms, finite {bool(np.isfinite(img).all())}",
                  flush=True)


if __name__ == "__main__":
    _main()


def atrous(color: np.ndarray, albedo: np.ndarray, normal: np.ndarray, samples: int) -> np.ndarray:
    """The tracer's own denoiser, `pfx_trace::denoise::filter`
    (crates/trace/src/denoise.rs:1-62): three edge-aware a-trous passes on
    the GPU. (H, W, 4) float32 images in, the filtered colour out."""
    height, width = color.shape[:2]
    c, a, n = (_f32(x, height, width) for x in (color, albedo, normal))
    out = np.empty_like(c)
    err = N.err_buffer()
    st = N.lib().pfxh_atrous(width, height, c.ctypes.data, a.ctypes.data, n.ctypes.data, int(samples),
                             out.ctypes.data, err, len(err))
    if st != N.OK:
        raise (ValueError if st == N.INVALID else RuntimeError)(err.value.decode(errors="replace"))
    return out
