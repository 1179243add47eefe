"""Rect and disc area lights on the device (LightShape::Rect/Disc,
crates/trace/src/lights.rs:10-99; lights.wgsl:57-91, 149-170): images at
equal spp against the oracle on a scene where camera and bounce rays see
the lights (two-sided, one-sided and unshadowed area lights beside the
light tree's point and spot lights), and the reference's direct-term test
with its rect light at 48x48 and 65,536 spp against the closed form."""
import numpy as np
import pytest

import stat_scenes as S
from oracle import Oracle
from pfx_amd import Bounces, Detail, Trace
from pfx_amd.scene import disc_light, rect_light
from test_gpu_parity import rel_mse, small_terrain

pytestmark = pytest.mark.gpu


def area_terrain(w=96, h=54):
    scene = small_terrain(w, h)
    extra = np.concatenate([
        rect_light((0.0, 6.0, 0.0), (1.5, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.9, 0.7), 60.0, two_sided=True),
        rect_light((-6.0, 4.0, 2.0), (0.0, 1.0, 0.0), (0.0, 0.0, 0.8), (0.6, 0.8, 1.0), 40.0),
        disc_light((5.0, 5.0, -2.0), (0.0, -1.0, 0.2), 0.8, (1.0, 0.6, 0.4), 50.0),
        disc_light((3.0, 3.0, 3.0), (1.0, -1.0, 0.0), 0.5, (0.5, 1.0, 0.5), 20.0, two_sided=True, shadow=False),
    ])
    scene.lights = np.concatenate([extra, scene.lights])
    return scene


@pytest.mark.parametrize("left", [False, True])
def test_area_lights_match_the_oracle_at_equal_spp(left):
    w, h = 96, 54
    scene = area_terrain(w, h)
    detail = Detail(bounces=Bounces.deep(6), left_first=left)
    print(f"area-lit terrain {w}x{h} 8 spp left_first={left}", flush=True)
    o = Oracle(scene, detail, w, h)
    rows, looped, tree = o.light_meta()
    assert looped == 4 and tree == 24
    o.sample(8, 17)
    rc, _, _ = o.readback()
    o.close()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(8, 17)
        out = t.readback()
    exact = float(np.mean(out.color == rc))
    rel = rel_mse(out.color[..., :3], rc[..., :3])
    print(f"  exact fraction {exact:.6f} relMSE {rel:.3e}")
    assert rel <= 1e-4
    assert exact > 0.99


def test_area_lights_seen_by_camera_rays():
    """An equal-spp render of the area-lit terrain without the tree's lights
    and with the sky black: the area lights alone light the frame; a camera
    ray through the two-sided rect (seen from above) picks up its radiance
    and goes on to the ground behind it (an area light is not an occluder,
    lights.wgsl:64-91), so those pixels hold at least that radiance."""
    w, h = 64, 64
    scene = area_terrain(w, h)
    scene.lights = scene.lights[:4]
    from pfx_amd import Camera
    scene.camera = Camera.look((0.0, 12.0, 0.0), (0.0, 0.0, 0.01), 20.0, 1.0)
    detail = Detail(bounces=Bounces.deep(3))
    o = Oracle(scene, detail, w, h)
    o.sample(4, 5)
    rc, _, _ = o.readback()
    o.close()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(4, 5)
        out = t.readback()
    assert rel_mse(out.color[..., :3], rc[..., :3]) <= 1e-4
    radiance = np.float32(60.0) / np.float32(4.0 * 1.5 * 1.0) * np.float32([1.0, 0.9, 0.7])
    centre = out.color[h // 2 - 2:h // 2 + 2, w // 2 - 2:w // 2 + 2, :3].reshape(-1, 3)
    assert np.all(centre >= radiance * np.float32(1.0 - 1e-6))
    assert float(out.color[..., :3].mean()) > 0.0


def test_area_lights_match_the_direct_term_on_the_gpu():
    """gpu.rs:4443-4560 with its rect light (stat_scenes.lit_floor_area), at
    the reference's 48x48 and 65,536 spp: the image mean within 0.3% of the
    closed form and every pixel with at least 2% of the peak within 2%
    (the reference's own bar)."""
    scene, lights = S.lit_floor_area()
    size = (48, 48)
    print("area-lit floor 48x48 65536 spp", flush=True)
    with Trace.new_detailed(scene, Detail(bounces=Bounces(0, 0, 0, 0)), *size) as t:
        for _ in range(16):
            t.sample(4096, 3)
        img = t.readback().color[..., :3].astype(np.float64)
    expected = S.lit_floor_area_expected(scene, lights, size, sub=4)
    ratio = img.mean() / expected.mean()
    bright = expected > 0.02 * expected.max()
    worst = float(np.max(np.abs(img[bright] - expected[bright]) / expected[bright]))
    print(f"  mean / direct term {ratio:.5f}, worst bright pixel {worst:.4f}")
    assert abs(ratio - 1.0) < 0.003
    assert worst < 0.02
