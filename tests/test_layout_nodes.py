"""The quantized traversal nodes (pfxh_layout_nodes): every child box a
64-byte node decodes to contains the exact box of its 128-byte node, with
the same child references, so a walk on them tests every triangle the walk
on the exact boxes would (the hits are the same; the device parity tests
check that end to end).

The decode is restated here without a fused multiply-add: the scale is a
power of two and the byte at most 255, so q * scale is exact in FP32 and
one FP32 addition of the origin rounds the same value fmaf(q, scale,
origin) rounds."""
import ctypes as C

import numpy as np
import pytest

from pfx_amd import _native as N
from pfx_amd import scenes

NODE4 = np.dtype([("lo", "<f4", (3, 4)), ("hi", "<f4", (3, 4)), ("ref", "<u4", 4), ("pad", "<u4", 4)])
QNODE4 = np.dtype([("origin", "<f4", 3), ("scale", "<f4", 3), ("q_lo", "<u4", 3), ("q_hi", "<u4", 3),
                   ("ref", "<u4", 4)])
REF_NONE = 0xFFFFFFFF


def layout_nodes(scene):
    sc, keep = scene.to_c()
    count, quantized = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    assert N.lib().pfxh_layout_nodes(C.byref(sc), None, None, C.byref(count), C.byref(quantized), err, 512) == 0
    nodes = np.zeros(count.value, NODE4)
    qnodes = np.zeros(count.value, QNODE4)
    assert N.lib().pfxh_layout_nodes(C.byref(sc), nodes.ctypes.data, qnodes.ctypes.data, C.byref(count),
                                     C.byref(quantized), err, 512) == 0, err.value
    return nodes, qnodes, bool(quantized.value)


def test_struct_sizes():
    assert NODE4.itemsize == 128 and QNODE4.itemsize == 64


@pytest.mark.parametrize("which", ["box", "terrain"])
def test_quantized_boxes_contain_the_exact_ones(which):
    if which == "box":
        scene = scenes.box_spheres(width=32, height=32)
    else:
        scene = scenes.terrain(seed=4, grid=(96, 48), size=(48.0, 24.0), meshes=4, mesh_tris=(200, 2000),
                               instances=24, lights=8, width=32, height=18)
    nodes, q, quantized = layout_nodes(scene)
    assert quantized and len(nodes) > 10
    np.testing.assert_array_equal(q["ref"], nodes["ref"])
    scale = q["scale"].astype(np.float32)
    # scales are powers of two (normal floats)
    mant = scale.view(np.uint32) & 0x7FFFFF
    assert np.all(mant == 0) and np.all(scale > 0) and np.all(np.isfinite(scale))
    valid = nodes["ref"] != REF_NONE  # (n, 4)
    for a in range(3):
        for c in range(4):
            ql = ((q["q_lo"][:, a] >> (8 * c)) & 0xFF).astype(np.float32)
            qh = ((q["q_hi"][:, a] >> (8 * c)) & 0xFF).astype(np.float32)
            lo = ql * scale[:, a] + q["origin"][:, a]
            hi = qh * scale[:, a] + q["origin"][:, a]
            v = valid[:, c]
            assert np.all(np.isfinite(lo[v])) and np.all(np.isfinite(hi[v]))
            assert np.all(lo[v] <= nodes["lo"][v, a, c]), (a, c)
            assert np.all(hi[v] >= nodes["hi"][v, a, c]), (a, c)
            # and the grid is tight: one step of the scale at most on each side
            # (beyond the rounding of the decode)
            assert np.all(nodes["lo"][v, a, c] - lo[v] <= 2 * scale[v, a])
            assert np.all(hi[v] - nodes["hi"][v, a, c] <= 2 * scale[v, a])
