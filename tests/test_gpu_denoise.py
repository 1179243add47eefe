"""The HIP denoiser (FP32 matrix-core convolutions) against OIDN 2.5.1's own
output on the desk fixture at the reference's tolerance (tests/oidn.rs:
1e-3 * (1 + |v|)), against the CPU oracle on a rendered frame, tiled against
untiled, and on a trace's images without a host round trip."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import Bounces, Detail, Trace, scenes
from pfx_amd import denoise as D
from pfx_amd import pfm

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "denoise"


def desk(part):
    return pfm.read((GOLDEN / f"desk.{part}.pfm").read_bytes())


def worst_gap(a, b):
    return float(np.max(np.abs(a[..., :3] - b[..., :3]) / (1.0 + np.abs(b[..., :3]))))


def oracle_denoise(color, albedo, normal, max_tile=2160 * 2160):
    h, w = color.shape[:2]
    data = D.WEIGHTS.read_bytes()
    out = np.zeros_like(color)
    assert O.lib().or_denoise(data, len(data), w, h, color.ctypes.data, albedo.ctypes.data, normal.ctypes.data,
                              out.ctypes.data, max_tile) == 0
    return out


def test_the_denoiser_matches_oidn_on_the_desk():
    color, albedo, normal, oidn = (desk(p) for p in ("color", "albedo", "normal", "oidn"))
    print("desk 32x24 against OIDN", flush=True)
    with D.Denoiser() as d:
        first = d.denoise(color, albedo, normal)
        again = d.denoise(color, albedo, normal)
    assert np.array_equal(first, again)
    gap = worst_gap(first, oidn)
    print(f"  worst gap {gap:.2e}")
    assert gap <= 1e-3
    assert np.array_equal(first[..., 3], color[..., 3])


def test_the_denoiser_matches_the_oracle_on_a_render():
    w, h = 96, 64
    scene = scenes.box_spheres(width=w, height=h)
    print("box_spheres 96x64 at 4 spp, denoised", flush=True)
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), w, h) as t:
        t.sample(4, 3)
        out = t.readback()
        with D.Denoiser() as d:
            ours = d.denoise(out.color, out.albedo, out.normal)
            on_device = t.denoise(d)
    assert np.array_equal(ours, on_device)
    ref = oracle_denoise(out.color, out.albedo, out.normal)
    gap = worst_gap(ours, ref)
    print(f"  worst gap against the oracle {gap:.2e}")
    assert gap <= 1e-4


def test_tiles_do_not_change_the_result():
    """Tiles overlap by more than half the receptive field and start on the
    16-pixel grid, so a tiled frame equals the untiled one."""
    w, h = 1000, 800
    scene = scenes.box_spheres(width=w, height=h)
    print("box_spheres 1000x800, one tile against 600k-pixel tiles", flush=True)
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), w, h) as t:
        t.sample(2, 1)
        out = t.readback()
    assert len(D.plan(h, w, 600_000)[2]) > 1
    with D.Denoiser() as d:
        whole = d.denoise(out.color, out.albedo, out.normal)
        d.set_max_tile(600_000)
        tiled = d.denoise(out.color, out.albedo, out.normal)
    assert np.isfinite(whole).all()
    assert float(np.max(np.abs(whole - tiled))) <= 1e-6 * max(1.0, float(np.max(np.abs(whole))))
