"""The allocation-free a-trous entry point (pfxh_atrous_noalloc) against the
allocating one (pfxh_atrous) through the C ABI: device images in and out,
two calls in a row on one GPU (the second smaller, so the shared static
middle image holds the first call's pixels beyond the second's), both bit
for bit equal to pfxh_atrous on the same inputs; sizes beyond the static
array, aliased outputs and zero samples are refused before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from pfx_amd import _native as N

pytestmark = pytest.mark.gpu


def images(rng, h, w):
    color = rng.exponential(0.6, (h, w, 4)).astype(np.float32)
    albedo = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    normal = rng.normal(size=(h, w, 4)).astype(np.float32)
    normal[..., :3] /= np.linalg.norm(normal[..., :3], axis=-1, keepdims=True)
    return color, albedo, normal


def host_atrous(color, albedo, normal, spp):
    h, w = color.shape[:2]
    out = np.zeros_like(color)
    err = N.err_buffer()
    st = N.lib().pfxh_atrous(w, h, color.ctypes.data, albedo.ctypes.data, normal.ctypes.data, spp, out.ctypes.data,
                             err, len(err))
    assert st == N.OK, err.value
    return out


class Hip:
    """The HIP runtime the library itself links (device buffers for the
    test's inputs and outputs; the entry point under test allocates none)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, *ptrs):
        for p in ptrs:
            self.rt.hipFree(p)


def device_atrous(hip, color, albedo, normal, spp):
    h, w = color.shape[:2]
    dev = [hip.upload(a) for a in (color, albedo, normal)]
    out = hip.upload(np.full_like(color, np.nan))
    err = N.err_buffer()
    st = N.lib().pfxh_atrous_noalloc(w, h, dev[0], dev[1], dev[2], spp, out, None, err, len(err))
    assert st == N.OK, err.value
    assert hip.rt.hipDeviceSynchronize() == 0
    got = hip.download(out, color)
    hip.free(out, *dev)
    return got


def test_no_alloc_atrous_equals_the_allocating_one():
    rng = np.random.default_rng(41)
    hip = Hip()
    for (h, w), spp in (((720, 1280), 16), ((200, 320), 4)):
        print(f"no-alloc a-trous {w}x{h} spp={spp}", flush=True)
        data = images(rng, h, w)
        want = host_atrous(*data, spp)
        got = device_atrous(hip, *data, spp)
        assert got.tobytes() == want.tobytes(), float(np.nanmax(np.abs(got - want)))


def test_no_alloc_atrous_refuses_bad_arguments():
    lib = N.lib()
    hip = Hip()
    err = N.err_buffer()
    p, q = hip.alloc(256), hip.alloc(256)
    try:
        assert lib.pfxh_atrous_noalloc(2048, 1025, p, p, p, 4, q, None, err, len(err)) == N.INVALID
        assert b"2097152" in err.value
        assert lib.pfxh_atrous_noalloc(2, 2, p, p, p, 0, q, None, err, len(err)) == N.INVALID
        assert lib.pfxh_atrous_noalloc(2, 2, p, p, p, 4, p, None, err, len(err)) == N.INVALID
        assert lib.pfxh_atrous_noalloc(0, 0, p, p, p, 4, q, None, err, len(err)) == N.OK
    finally:
        hip.free(p, q)
