"""The Sobol sampler on the device (Sampler::Sobol, sobol.wgsl): images at
equal spp against the oracle on the parity scenes (camera jitter, DOF,
light slots, BSDF and Russian roulette all read the patterns), a batch of
Sobol stills byte-identical to their own traces, and the direct-term
floor's error below PCG's at equal spp on the device too."""
import numpy as np
import pytest

import stat_scenes as S
from oracle import Oracle
from pfx_amd import Bounces, Detail, Trace
from pfx_amd.batch import Batch
from test_gpu_batch import mix, same, single
from test_gpu_parity import CASES, rel_mse

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", sorted(CASES))
def test_sobol_images_match_the_oracle(case):
    make, bounces, w, h = CASES[case]
    scene = make(w, h)
    detail = Detail(bounces=bounces, sampler="sobol", sobol_samples=16, aperture=0.02, focus_distance=3.0)
    print(f"sobol image parity {case} {w}x{h}", flush=True)
    o = Oracle(scene, detail, w, h)
    o.sample(8, 7)
    o.sample(8, 8)
    rc, _, _ = o.readback()
    o.close()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(8, 7)
        t.sample(8, 8)
        out = t.readback()
    exact = float(np.mean(out.color == rc))
    rel = rel_mse(out.color[..., :3], rc[..., :3])
    print(f"  exact fraction {exact:.6f} relMSE {rel:.3e}")
    assert rel <= 1e-4 and exact > 0.99


def test_sobol_batch_equals_single_traces():
    detail = Detail(bounces=Bounces.deep(4), sampler="sobol")
    stills = mix()
    with Batch(stills, detail) as b:
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
    for still, out in zip(stills, outs):
        same(out, single(still, detail))


def test_sobol_error_below_pcg_on_the_device():
    scene, lights = S.lit_floor_area()
    size = (48, 48)
    expected = S.lit_floor_area_expected(scene, lights, size)
    errs = {}
    for sampler in ("pcg", "sobol"):
        with Trace.new_detailed(scene, Detail(bounces=Bounces(0, 0, 0, 0), sampler=sampler, sobol_samples=256),
                                *size) as t:
            t.sample(256, 3)
            img = t.readback().color[..., :3].astype(np.float64)
        errs[sampler] = float(np.sqrt(np.mean((img - expected) ** 2)) / expected.mean())
        print(f"  {sampler} 256 spp: relative RMS error {errs[sampler]:.5f}, mean ratio {img.mean() / expected.mean():.5f}")
        assert abs(img.mean() / expected.mean() - 1.0) < 0.005
    assert errs["sobol"] < 0.8 * errs["pcg"]
