"""The watertight ray-triangle test on the device (PFXH_TRI_WATERTIGHT):
bit-exact against the oracle's restatement (primary-hit ids and distances,
images at equal spp), closed meshes seen from inside with no leaks, and a
batch of watertight stills byte-identical to their own traces. The test's
watertightness property itself is pinned on the CPU (test_watertight.py)."""
import numpy as np
import pytest

from oracle import Oracle
from pfx_amd import EQUIRECT, Bounces, Detail, Trace, scenes
from pfx_amd.batch import Batch
from test_gpu_batch import mix, same, single
from test_gpu_parity import CASES, rel_mse
from test_watertight import ico_scene

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


@pytest.mark.parametrize("case", sorted(CASES))
def test_probe_ids_bit_exact(case):
    make, bounces, w, h = CASES[case]
    scene = make(w, h)
    detail = Detail(bounces=bounces, watertight=True)
    print(f"watertight probe parity {case} {w}x{h}", flush=True)
    ref = Oracle(scene, detail, w, h).probe()
    with Trace.new_detailed(scene, detail, w, h) as t:
        got = t.probe()
    assert np.array_equal(got.primitive, ref["primitive"])
    assert np.array_equal(got.instance, ref["instance"])
    np.testing.assert_array_equal(got.distance, ref["distance"])
    np.testing.assert_allclose(got.normal, ref["normal"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("case", sorted(CASES))
def test_images_match_oracle_at_equal_spp(case):
    make, bounces, w, h = CASES[case]
    scene = make(w, h)
    detail = Detail(bounces=bounces, watertight=True)
    print(f"watertight image parity {case} {w}x{h}", flush=True)
    o = Oracle(scene, detail, w, h)
    o.sample(8, 7)
    rc, ra, rn = o.readback()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(8, 7)
        out = t.readback()
    exact = float(np.mean(out.color == rc))
    print(f"  exact fraction {exact:.6f}")
    assert rel_mse(out.color[..., :3], rc[..., :3]) <= 1e-4
    assert exact > 0.99


@pytest.mark.parametrize("instanced", [False, True])
def test_closed_mesh_from_inside_has_no_leaks(instanced):
    """An equirect probe from inside a closed icosphere: every one of the
    512x256 rays hits it with the watertight test, identically to the
    oracle; the count the reference's test lets through is printed."""
    scene, _ = ico_scene(sub=4, instanced=instanced)
    w, h = 512, 256
    print(f"closed icosphere instanced={instanced}", flush=True)
    counts = {}
    for wt in (False, True):
        detail = Detail(bounces=Bounces.deep(1), projection=EQUIRECT, watertight=wt)
        with Trace.new_detailed(scene, detail, w, h) as t:
            got = t.probe()
        counts[wt] = int(np.sum(got.primitive == MISS))
        if wt:
            ref = Oracle(scene, detail, w, h).probe()
            assert np.array_equal(got.primitive, ref["primitive"])
    print(f"  leaks: Moller-Trumbore {counts[False]}, watertight {counts[True]} of {w * h}")
    assert counts[True] == 0


def test_batch_of_watertight_stills_equals_single_traces():
    detail = Detail(bounces=Bounces.deep(4), watertight=True)
    stills = mix()
    with Batch(stills, detail) as b:
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
    for still, out in zip(stills, outs):
        same(out, single(still, detail))
