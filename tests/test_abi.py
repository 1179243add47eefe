"""The C ABI: libpfxhip.so loads without a GPU, exports every function
include/pfxhip.h declares, and its struct layouts match the Python view.
No device compute runs here."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from pfx_amd import _native as N

HEADER = Path(__file__).resolve().parent.parent / "include" / "pfxhip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(pfxh_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(native_lib):
    names = declared_functions()
    assert "pfxh_trace_new" in names and "pfxh_layout_bvh" in names
    for name in names:
        assert hasattr(native_lib, name), name
    assert set(names) == set(N.SIGNATURES), set(names) ^ set(N.SIGNATURES)


def test_version_string(native_lib):
    assert native_lib.pfxh_version().startswith(b"pfxhip")


def test_struct_sizes_match_the_header():
    # sizes of the C structs as laid out by the x86-64 SysV ABI
    assert C.sizeof(N.Camera) == 48
    assert C.sizeof(N.Sun) == 28
    assert C.sizeof(N.Environment) == 16
    assert C.sizeof(N.DetailC) == 76
    assert C.sizeof(N.AdaptiveC) == 16
    assert C.sizeof(N.ConvergedC) == 4 + 4 * 64 + 4 + 8 * 64 + 8
    assert C.sizeof(N.SceneC) == 7 * 16 + 16 + 48 + 28 + 4
    assert N.PROBE_HIT.itemsize == 32 and N.BVH_NODE.itemsize == 48


def test_invalid_inputs_are_refused_without_a_device(native_lib):
    err = C.create_string_buffer(256)
    tris = np.zeros(1, N.TRIANGLE)
    tris["vertices"] = [0, 0, 0, 1, 0, 0, np.nan, 1, 0]
    count = C.c_uint32()
    nodes = np.zeros(2, N.BVH_NODE)
    order = np.zeros(1, np.uint32)
    st = native_lib.pfxh_layout_bvh(tris.ctypes.data, 1, nodes.ctypes.data, C.byref(count), order.ctypes.data,
                                    err, 256)
    assert st == N.SETUP and b"invalid" in err.value
    out = C.c_void_p()
    assert native_lib.pfxh_trace_new(None, None, 4, 4, C.byref(out), err, 256) == N.INVALID
    assert native_lib.pfxh_trace_samples(None) == 0
    native_lib.pfxh_trace_free(None)


def test_invalid_projection_is_refused_before_any_device_work(native_lib):
    """Projection::valid / fits (gpu.rs:191-208): an orthographic view needs a
    positive finite size, and orthographic or equirect cameras need nonzero
    forward, right and up."""
    from pfx_amd import EQUIRECT, ORTHOGRAPHIC, Detail, scenes

    scene = scenes.box_spheres(width=8, height=8)
    err = C.create_string_buffer(256)
    out = C.c_void_p()
    sc, keep = scene.to_c()
    bad = Detail(projection=ORTHOGRAPHIC, ortho_size=(0.0, 1.0)).to_c()
    assert native_lib.pfxh_trace_new(C.byref(sc), C.byref(bad), 8, 8, C.byref(out), err, 256) == N.SETUP
    assert b"invalid trace detail" in err.value
    unknown = Detail(projection=7).to_c()
    assert native_lib.pfxh_trace_new(C.byref(sc), C.byref(unknown), 8, 8, C.byref(out), err, 256) == N.SETUP
    scene.camera.up = (0.0, 0.0, 0.0)
    sc, keep = scene.to_c()
    flat = Detail(projection=EQUIRECT).to_c()
    assert native_lib.pfxh_trace_new(C.byref(sc), C.byref(flat), 8, 8, C.byref(out), err, 256) == N.SETUP
    assert b"needs forward, right and up" in err.value
    assert not out.value


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct the bindings mirror has the header's size, as a C
    compiler lays it out."""
    import subprocess
    from pathlib import Path

    pairs = {"pfxh_camera": N.Camera, "pfxh_sun": N.Sun, "pfxh_environment": N.Environment,
             "pfxh_detail": N.DetailC, "pfxh_adaptive": N.AdaptiveC, "pfxh_converged": N.ConvergedC,
             "pfxh_scene": N.SceneC, "pfxh_stats": N.StatsC, "pfxh_denoise_tile": N.DenoiseTile,
             "pfxh_checkpoint": N.CheckpointC, "pfxh_learn_layer_info": N.LearnLayerC,
             "pfxh_batch_still": N.BatchStillC, "pfxh_batch_tuning": N.BatchTuningC}
    header = Path(__file__).resolve().parent.parent / "include" / "pfxhip.h"
    src = tmp_path / "sizes.c"
    src.write_text(f'#include "{header}"\n#include <stdio.h>\nint main(void) {{\n'
                   + "".join(f'    printf("{k} %zu\\n", sizeof({k}));\n' for k in pairs) + "    return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for k, t in pairs.items():
        assert int(out[k]) == C.sizeof(t), (k, out[k], C.sizeof(t))


def test_host_side_argument_errors(native_lib):
    """Calls the library refuses before touching a device."""
    lib = native_lib
    err = N.err_buffer()
    img = (C.c_float * 16)()
    assert lib.pfxh_atrous(2, 2, img, img, img, 0, img, err, len(err)) == N.INVALID
    assert b"invalid denoise inputs" in err.value
    assert lib.pfxh_atrous(2, 2, None, img, img, 4, img, err, len(err)) == N.INVALID
    out = C.c_void_p()
    assert lib.pfxh_batch_new(None, 0, None, C.byref(out), err, len(err)) == N.INVALID
    assert lib.pfxh_batch_trace(None, 0) is None
    assert lib.pfxh_batch_count(None) == 0
    assert lib.pfxh_learner_set_params(None, None, 0, err, len(err)) == N.INVALID
    assert lib.pfxh_learner_train(None, 1, 1, C.c_float(1e-3), 1, None, err, len(err)) == N.INVALID
    assert lib.pfxh_trace_learned_denoise_checked(None, None, None, C.c_float(1.0), 32, None, None, err,
                                                  len(err)) == N.INVALID
    info = N.LearnLayerC()
    assert lib.pfxh_learn_layer(N.LEARN_LAYERS, C.byref(info)) == N.INVALID
    gbs = C.c_double()
    assert lib.pfxh_hbm_probe(1 << 10, 3, C.byref(gbs), err, len(err)) == N.INVALID  # under 1 MiB
    assert lib.pfxh_hbm_probe((1 << 20) + 4, 3, C.byref(gbs), err, len(err)) == N.INVALID  # not whole float4s
    assert lib.pfxh_hbm_probe(1 << 20, 0, C.byref(gbs), err, len(err)) == N.INVALID
    f3 = (C.c_float * 3)()
    ids = (C.c_uint32 * 1)()
    assert lib.pfxh_trace_cast(None, f3, f3, 1, C.c_float(1.0), ids, f3, err, len(err)) == N.INVALID


def test_cast_refuses_unequal_ray_arrays():
    """Trace.cast checks that origins and directions are equal (n, 3)
    arrays before the C call, which reads 3*count floats from each."""
    from pfx_amd.trace import Trace

    o = np.zeros((4, 3), np.float32)
    for d in (np.zeros((3, 3), np.float32), np.zeros((4, 2), np.float32), np.zeros(11, np.float32)):
        with pytest.raises(ValueError):
            Trace.cast(object(), o, d)


def test_device_limits_round_trip(native_lib):
    got = N.DeviceLimitsC()
    native_lib.pfxh_set_device_limits(C.byref(N.DeviceLimitsC(123, 456)))
    native_lib.pfxh_get_device_limits(C.byref(got))
    assert (got.max_buffer_size, got.max_memory) == (123, 456)
    native_lib.pfxh_set_device_limits(None)
    native_lib.pfxh_get_device_limits(C.byref(got))
    assert (got.max_buffer_size, got.max_memory) == (0, 0)
