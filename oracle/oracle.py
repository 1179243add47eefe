"""ctypes wrapper of oracle/_build/liboracle.so -- TEST INFRASTRUCTURE ONLY.

`Oracle` follows the same calls as `pfx_amd.Trace` so parity tests read the
same on both sides.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
_lib = None

_P, _U, _F = C.c_void_p, C.c_uint32, C.c_float


def build() -> Path:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        sig = {
            "or_scene_new": (_P, [_P, _P, _U, _U, C.c_char_p, C.c_size_t]),
            "or_tza_count": (C.c_int, [C.c_char_p, C.c_size_t]),
            "or_denoise_plan": (C.c_int, [_U, _U, C.c_uint64, C.POINTER(_U), C.POINTER(_U), _P, _U]),
            "or_denoise": (C.c_int, [C.c_char_p, C.c_size_t, _U, _U, _P, _P, _P, _P, C.c_uint64]),
            "or_atrous": (C.c_int, [_U, _U, _P, _P, _P, _U, _P]),
            "or_learn_param_count": (C.c_uint64, []),
            "or_learn_layer": (None, [_U, _P]),
            "or_learn_grad": (C.c_double, [_P, _U, _U, _U, _P, _P, _P, _P, _P, _P, _P]),
            "or_learn_set_loss": (None, [C.c_int]),
            "or_learn_forward": (None, [_P, _U, _U, _P, _P, _P, _P, _P, _P]),
            "or_scene_free": (None, [_P]),
            "or_set_camera": (None, [_P, _P]),
            "or_set_sun_radius": (None, [_P, _F]),
            "or_set_environment": (None, [_P, _P]),
            "or_set_window": (C.c_int, [_P, _U, _U, _U, _U]),
            "or_sample": (None, [_P, _U, _U, _U, _U]),
            "or_samples": (_U, [_P]),
            "or_readback": (None, [_P, _P, _P, _P]),
            "or_deviation": (None, [_P, _P]),
            "or_counters": (None, [_P, _P]),
            "or_probe": (None, [_P, _P]),
            "or_sample_adaptive": (C.c_int, [_P, _P, _U, _P]),
            "or_tree_pick": (None, [_P, _P, _P, _F, _P]),
            "or_tree_pick_many": (None, [_P, _P, _P, _P, _U, _P]),
            "or_sobol_pattern": (None, [_P, _U, _U]),
            "or_sobol_point": (None, [_P, _U, _U, _U, _U, _P]),
            "or_light_meta": (_U, [_P, C.POINTER(_U), C.POINTER(_U)]),
            "or_trace_counts": (None, [_P, _P, _P, _P]),
            "or_cast": (None, [_P, _P, _P, _U, C.c_int, _F, _P, _P]),
            "or_pcg": (_U, [_U]),
            "or_offset_axis": (_F, [_F, _F]),
            "or_spawn_point": (None, [_P, _P, _P, _F, _P]),
            "or_fm": (_F, [C.c_int, _F, _F]),
            "or_fm_many": (None, [C.c_int, _P, _P, _P, C.c_size_t]),
            "or_bvh": (_U, [_P, _U, _P, _P]),
            "or_bvh_intersect": (_U, [_P, _U, _P, _U, _P, _P, _P]),
            "or_triangle_hit": (_F, [_P, _P, _P, _F]),
            "or_levels": (C.c_int, [_P, _P, C.POINTER(_U), _P, C.POINTER(_U), _P, C.c_char_p, C.c_size_t]),
            "or_lights": (_U, [_P, _U, _P, C.POINTER(_U), C.POINTER(_U)]),
            "or_set_tree_split": (None, [C.c_int]),
            "or_light_pack": (None, [_P, _P]),
            "or_light_falloff": (_F, [_P, _F]),
            "or_light_cone": (_F, [_P, _P]),
            "or_light_valid": (C.c_int, [_P]),
            "or_emitters": (_U, [_P, _U, _U, _U, _P, _P]),
            "or_sky_cdf": (None, [_P, _U, _U, _U, _U, _P, _P]),
            "or_rounds": (_U, [_P, _P, _U]),
            "or_select_nth": (C.c_size_t, [_P, _P, C.c_size_t, C.c_size_t, C.c_int]),
            "or_adaptive_error": (_F, [_P, _F, _P, _F]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data if a is not None and a.size else None


class Oracle:
    """The CPU restatement with the `Trace` call surface."""

    def __init__(self, scene, detail, width: int, height: int, threads: int | None = None):
        from pfx_amd import _native as N  # struct layouts only

        if threads is not None:
            os.environ["OMP_NUM_THREADS"] = str(threads)
        self._N = N
        sc, keep = scene.to_c()
        d = detail.to_c()
        err = C.create_string_buffer(1024)
        self._h = lib().or_scene_new(C.byref(sc), C.byref(d), width, height, err, len(err))
        del keep
        if not self._h:
            raise ValueError(err.value.decode())
        self.width, self.height = width, height

    def sample(self, count: int, seed: int, rows: tuple | None = None) -> None:
        r0, r1 = rows if rows else (0, 0)
        lib().or_sample(self._h, count, seed, r0, r1)

    def samples(self) -> int:
        return int(lib().or_samples(self._h))

    def readback(self):
        shape = (self.height, self.width, 4)
        c, a, n = (np.empty(shape, np.float32) for _ in range(3))
        lib().or_readback(self._h, c.ctypes.data, a.ctypes.data, n.ctypes.data)
        return c, a, n

    def deviation(self):
        out = np.empty((self.height, self.width), np.float32)
        lib().or_deviation(self._h, out.ctypes.data)
        return out

    def set_camera(self, camera) -> None:
        cc = camera.to_c()
        lib().or_set_camera(self._h, C.byref(cc))

    def set_sun_radius(self, radius_deg: float) -> None:
        lib().or_set_sun_radius(self._h, float(radius_deg))

    def set_environment(self, sky) -> None:
        from pfx_amd import _native as N

        if sky is None:
            env = N.Environment(0, 0, None)
            self._sky = None
        else:
            self._sky = np.ascontiguousarray(sky, np.float32)
            env = N.Environment(self._sky.shape[1], self._sky.shape[0], self._sky.ctypes.data)
        lib().or_set_environment(self._h, C.byref(env))

    def set_window(self, full, origin) -> None:
        if lib().or_set_window(self._h, full[0], full[1], origin[0], origin[1]):
            raise ValueError("the window must lie inside the full frame")

    def probe(self) -> np.ndarray:
        hits = np.zeros(self.width * self.height, dtype=self._N.PROBE_HIT)
        lib().or_probe(self._h, hits.ctypes.data)
        return hits.reshape(self.height, self.width)

    def sample_adaptive(self, adaptive, seed: int):
        out = self._N.ConvergedC()
        a = adaptive.to_c()
        if lib().or_sample_adaptive(self._h, C.byref(a), seed, C.byref(out)):
            raise ValueError("invalid adaptive settings")
        return [(int(out.samples[i]), int(out.active[i])) for i in range(out.round_count)], int(out.pixels)

    def counters(self) -> dict:
        out = np.zeros(3, np.uint64)
        lib().or_counters(self._h, out.ctypes.data)
        return {"closest_rays": int(out[0]), "shadow_rays": int(out[1]), "paths": int(out[2])}

    def tree_pick(self, p, n, u: float):
        p = np.asarray(p, np.float32)
        n = np.asarray(n, np.float32)
        out = np.zeros(2, np.float32)
        lib().or_tree_pick(self._h, p.ctypes.data, n.ctypes.data, u, out.ctypes.data)
        return float(out[0]), float(out[1])

    def tree_picks(self, p, n, u):
        """tree_pick at every uniform of `u`: (light rows, pdfs)."""
        p = np.asarray(p, np.float32)
        n = np.asarray(n, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        out = np.zeros((len(u), 2), np.float32)
        lib().or_tree_pick_many(self._h, p.ctypes.data, n.ctypes.data, u.ctypes.data, len(u), out.ctypes.data)
        return out[:, 0].astype(np.int64), out[:, 1]

    def light_meta(self):
        looped, tree = _U(), _U()
        rows = lib().or_light_meta(self._h, C.byref(looped), C.byref(tree))
        return int(rows), int(looped.value), int(tree.value)

    def cast(self, origins, directions, any_hit: bool = False, limit: float = 1e20):
        """Closest-hit caller triangle ids (UINT32_MAX on a miss) and distances
        of a ray list; with any_hit, 1 where blocked before `limit`."""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        ids = np.zeros(len(o), np.uint32)
        t = np.zeros(len(o), np.float32)
        lib().or_cast(self._h, _ptr(o), _ptr(d), len(o), 1 if any_hit else 0, limit, _ptr(ids), _ptr(t))
        return ids, t

    def trace_counts(self, o, d):
        o = np.asarray(o, np.float32)
        d = np.asarray(d, np.float32)
        out = np.zeros(2, np.uint64)
        lib().or_trace_counts(self._h, o.ctypes.data, d.ctypes.data, out.ctypes.data)
        return int(out[0]), int(out[1])

    def close(self):
        if self._h:
            lib().or_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- the learned denoiser (FP64) ---------------------------------------------------------

def learn_param_count() -> int:
    return int(lib().or_learn_param_count())


def learn_layers() -> list[tuple[int, ...]]:
    """(in0, in1, in_channels, out_channels, level, upsample, pool, relu) per layer."""
    out = []
    for k in range(16):
        v = np.zeros(8, np.int32)
        lib().or_learn_layer(k, v.ctypes.data)
        out.append(tuple(int(x) for x in v))
    return out


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, np.float32)


def learn_grad(params, color, albedo, normal, depth, target, deviation=None, want_grad: bool = True,
               loss: str = "l1_log"):
    """Loss and gradient for crops shaped (batch, h, w, 4) (depth and
    deviation (batch, h, w) or None); loss 'l1_log' or 'rel_l2'."""
    lib().or_learn_set_loss({"l1_log": 0, "rel_l2": 1}[loss])
    p = np.ascontiguousarray(params, np.float64)
    color, albedo, normal, depth, deviation, target = (_f32(a) for a in (color, albedo, normal, depth, deviation,
                                                                           target))
    b, h, w = color.shape[:3]
    grad = np.zeros(len(p), np.float64) if want_grad else None
    loss = lib().or_learn_grad(p.ctypes.data, b, h, w, color.ctypes.data, albedo.ctypes.data, normal.ctypes.data,
                               _ptr(depth), _ptr(deviation), target.ctypes.data, _ptr(grad))
    return float(loss), grad


def learn_forward(params, color, albedo, normal, depth, deviation=None):
    """Denoise one (h, w, 4) image; returns (h, w, 4) float32."""
    p = np.ascontiguousarray(params, np.float64)
    color, albedo, normal, depth, deviation = (_f32(a) for a in (color, albedo, normal, depth, deviation))
    h, w = color.shape[:2]
    out = np.zeros((h, w, 4), np.float32)
    lib().or_learn_forward(p.ctypes.data, h, w, color.ctypes.data, albedo.ctypes.data, normal.ctypes.data,
                           _ptr(depth), _ptr(deviation), out.ctypes.data)
    return out


def atrous(color, albedo, normal, samples: int):
    """denoise::filter on (h, w, 4) float32 images; returns (h, w, 4)."""
    color, albedo, normal = (np.ascontiguousarray(a, np.float32) for a in (color, albedo, normal))
    h, w = color.shape[:2]
    out = np.zeros_like(color)
    if lib().or_atrous(w, h, color.ctypes.data, albedo.ctypes.data, normal.ctypes.data, samples, out.ctypes.data):
        raise ValueError("invalid denoise inputs")
    return out
