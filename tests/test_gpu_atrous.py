"""The HIP a-trous filter against the oracle, bit for bit: random images of
an odd size, the reference's edge-and-grain case, and a trace's own images
at its sample count."""
import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import Bounces, Detail, Trace, scenes
from pfx_amd import denoise as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,spp", [(1, 4), (2, 64), (3, 1)])
def test_atrous_bit_exact_vs_oracle(seed, spp):
    print(f"atrous seed={seed} spp={spp}", flush=True)
    rng = np.random.default_rng(seed)
    h, w = 37, 53
    color = rng.exponential(0.5, (h, w, 4)).astype(np.float32)
    albedo = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    normal = rng.uniform(-1, 1, (h, w, 4)).astype(np.float32)
    out = D.atrous(color, albedo, normal, spp)
    assert out.tobytes() == O.atrous(color, albedo, normal, spp).tobytes()


def test_atrous_reference_case_and_trace_images():
    print("atrous trace", flush=True)
    h, w = 16, 32
    yy, xx = np.indices((h, w))
    side = np.where(xx < w // 2, 0.2, 0.8)
    grain = np.where((xx + yy) % 2 == 0, 0.08, -0.08)
    one = np.ones((h, w))
    color = np.stack([side + grain] * 3 + [one], -1).astype(np.float32)
    albedo = np.stack([side] * 3 + [one], -1).astype(np.float32)
    normal = np.stack([0 * one, 0 * one, one, one], -1).astype(np.float32)
    out = D.atrous(color, albedo, normal, 64)
    assert out.tobytes() == O.atrous(color, albedo, normal, 64).tobytes()
    assert out[8, 16, 0] - out[8, 15, 0] > 0.5
    with pytest.raises(ValueError):
        D.atrous(color, albedo, normal, 0)
    scene = scenes.box_spheres(seed=2, spheres=8, width=96, height=64)
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), 96, 64) as t:
        with pytest.raises(ValueError):
            t.atrous()
        t.sample(6, 1)
        img = t.readback()
        filtered = t.atrous()
    assert filtered.tobytes() == O.atrous(img.color, img.albedo, img.normal, 6).tobytes()
