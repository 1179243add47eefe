"""GPU parity: the HIP wavefront tracer against the CPU oracle on the same
scenes and seeds, through the C ABI.

Bars (for integer/index outputs and, with identical float operation order on
both sides, for the images): primary-hit primitive ids and instance ids are
bit-exact; images at equal spp match within relMSE <= 1e-4 (tests also report
the exact-match fraction). Full-size properties are checked in
test_gpu_properties.py.
"""
import numpy as np
import pytest

from oracle import Oracle
from pfx_amd import EQUIRECT, ORTHOGRAPHIC, PERSPECTIVE, Bounces, Detail, Trace, scenes
from pfx_amd.scene import ACCUM_COMPENSATED, ACCUM_PLAIN

pytestmark = pytest.mark.gpu


def rel_mse(x, r):
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def small_terrain(width=96, height=54):
    return scenes.terrain(seed=3, grid=(64, 32), size=(32.0, 16.0), amplitude=4.0, meshes=3,
                          mesh_tris=(200, 600), instances=12, lights=24, width=width, height=height,
                          light_range=12.0)


CASES = {
    "box_spheres": (lambda w, h: scenes.box_spheres(width=w, height=h), Bounces.deep(4), 64, 64),
    "terrain": (lambda w, h: small_terrain(w, h), Bounces.deep(6), 96, 54),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_probe_primitive_ids_bit_exact(case):
    make, bounces, w, h = CASES[case]
    scene = make(w, h)
    detail = Detail(bounces=bounces)
    print(f"probe parity {case} {w}x{h}", flush=True)
    ref = Oracle(scene, detail, w, h).probe()
    with Trace.new_detailed(scene, detail, w, h) as t:
        got = t.probe()
    assert np.array_equal(got.primitive, ref["primitive"])
    assert np.array_equal(got.instance, ref["instance"])
    assert np.array_equal(got.id, ref["material_id"])
    np.testing.assert_array_equal(got.distance, ref["distance"])
    np.testing.assert_allclose(got.normal, ref["normal"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("accumulate", [ACCUM_PLAIN, ACCUM_COMPENSATED])
@pytest.mark.parametrize("case", sorted(CASES))
def test_images_match_oracle_at_equal_spp(case, accumulate):
    make, bounces, w, h = CASES[case]
    scene = make(w, h)
    detail = Detail(bounces=bounces, accumulate=accumulate)
    spp, seed = 8, 7
    print(f"image parity {case} {w}x{h} spp={spp} accumulate={accumulate}", flush=True)
    o = Oracle(scene, detail, w, h)
    o.sample(spp, seed)
    rc, ra, rn = o.readback()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(spp, seed)
        out = t.readback()
        assert t.samples() == spp
    exact = float(np.mean(out.color == rc))
    print(f"  exact fraction {exact:.6f} relMSE {rel_mse(out.color[..., :3], rc[..., :3]):.3e}")
    assert rel_mse(out.color[..., :3], rc[..., :3]) <= 1e-4
    assert rel_mse(out.albedo, ra) <= 1e-6
    np.testing.assert_allclose(out.normal, rn, rtol=0, atol=1e-4)
    assert exact > 0.99


def test_accumulation_across_calls_and_reset():
    """sample(4)+sample(4) equals sample(8); a setter resets the count
    (crates/trace/src/gpu.rs:2817-2940)."""
    scene = scenes.box_spheres(width=48, height=48)
    detail = Detail(bounces=Bounces.deep(4), accumulate=ACCUM_PLAIN)
    with Trace.new_detailed(scene, detail, 48, 48) as t:
        t.sample(4, 11)
        t.sample(4, 11)
        split = t.readback().color
        assert t.samples() == 8
        t.set_camera(scene.camera)
        assert t.samples() == 0
        t.sample(8, 11)
        whole = t.readback().color
    np.testing.assert_array_equal(split, whole)


def test_window_equals_full_frame():
    """A window renders the same bytes as that region of the full frame
    (docs/trace.md:539-541)."""
    W, H = 64, 48
    scene = scenes.box_spheres(width=W, height=H)
    detail = Detail(bounces=Bounces.deep(4))
    with Trace.new_detailed(scene, detail, W, H) as t:
        t.sample(4, 5)
        full = t.readback().color
    with Trace.new_detailed(scene, detail, 32, 16) as t:
        t.set_window((W, H), (16, 20))
        t.sample(4, 5)
        part = t.readback().color
    np.testing.assert_array_equal(part, full[20:36, 16:48])


def test_runs_are_bit_identical():
    scene = small_terrain()
    detail = Detail(bounces=Bounces.deep(6))
    imgs = []
    for _ in range(2):
        with Trace.new_detailed(scene, detail, 96, 54) as t:
            t.sample(4, 3)
            imgs.append(t.readback().color)
    np.testing.assert_array_equal(imgs[0], imgs[1])


def test_adaptive_rounds_match_oracle():
    from pfx_amd import Adaptive

    scene = scenes.box_spheres(width=48, height=48)
    detail = Detail(bounces=Bounces.deep(4), accumulate=ACCUM_PLAIN)
    ad = Adaptive(threshold=0.05, min_samples=8, max_samples=32, growth=1.5)
    o = Oracle(scene, detail, 48, 48)
    ref_rounds, pixels = o.sample_adaptive(ad, 9)
    rc, _, _ = o.readback()
    with Trace.new_detailed(scene, detail, 48, 48) as t:
        conv = t.sample_adaptive(ad, 9)
        got = t.readback().color
    assert [(r.samples, r.active) for r in conv.rounds] == ref_rounds
    assert conv.pixels == pixels
    assert rel_mse(got[..., :3], rc[..., :3]) <= 1e-4


PROJECTIONS = {"orthographic": (ORTHOGRAPHIC, (0.9, 0.9)), "equirect": (EQUIRECT, (0.0, 0.0))}


def tied_scene(w, h):
    """Every hit is an exact tie: each world triangle appears twice (with
    different materials) and one mesh is placed twice with the same
    transform, so equal distances must resolve to the triangle and the
    instance the reference's left-first walk meets first."""
    from pfx_amd import _native as N
    from pfx_amd.scene import Camera, Scene, material

    rng = np.random.default_rng(4)
    balls = []
    for _ in range(12):
        c = rng.uniform(-2.0, 2.0, 3)
        balls.append(scenes.icosphere(1) * rng.uniform(0.3, 0.8) + c)
    world = np.concatenate(balls).astype(np.float32)
    verts = np.concatenate([world, world])
    mats = np.concatenate([np.zeros(len(world)), np.ones(len(world))])
    scene = Scene(triangles=scenes._triangles(verts, mats),
                  materials=np.concatenate([material((0.5, 0.5, 0.5)), material((0.2, 0.6, 0.2)),
                                            material((0.6, 0.2, 0.2))]),
                  camera=Camera.look((0.0, 0.5, -7.0), (0.0, 0.0, 0.0), 60.0, w / h))
    unit = scenes.icosphere(2).astype(np.float32)
    rows = np.zeros(len(unit), N.TRIANGLE_DETAIL)
    rows["vertices"] = unit.reshape(len(unit), 9)
    rows["material"] = 2
    scene.mesh_triangles = rows
    scene.meshes = np.array([(0, len(unit))], N.MESH)
    inst = np.zeros(4, N.INSTANCE)
    for k, (pos, yaw) in enumerate([((0.5, -1.0, 1.0), 0.3), ((0.5, -1.0, 1.0), 0.3),
                                    ((-1.5, 1.0, 0.0), 1.1), ((-1.5, 1.0, 0.0), 1.1)]):
        inst[k]["mesh"] = 0
        inst[k]["transform"] = scenes.transform(pos, yaw, 0.9)
    scene.instances = inst
    return scene


def test_equal_distance_ties_resolve_like_the_reference():
    w, h = 160, 120
    scene = tied_scene(w, h)
    for watertight in (False, True):
        detail = Detail(bounces=Bounces.deep(2), watertight=watertight)
        print(f"ties watertight={watertight}", flush=True)
        ref = Oracle(scene, detail, w, h).probe()
        with Trace.new_detailed(scene, detail, w, h) as t:
            got = t.probe()
        hit = ref["primitive"] != 0xFFFFFFFF
        assert hit.mean() > 0.1
        assert np.array_equal(got.primitive, ref["primitive"])
        assert np.array_equal(got.instance, ref["instance"])
        assert np.array_equal(got.id, ref["material_id"])
        np.testing.assert_array_equal(got.distance, ref["distance"])


@pytest.mark.parametrize("name", sorted(PROJECTIONS))
def test_projections_match_oracle(name):
    """Orthographic and equirect cameras (trace.wgsl:2546-2553, probe.wgsl:18-34):
    probe ids and distances bit-exact, images at equal spp as for perspective,
    and set_projection on a perspective trace equals a trace built with it."""
    kind, size = PROJECTIONS[name]
    w, h = 64, 48
    scene = scenes.box_spheres(width=w, height=h)
    detail = Detail(bounces=Bounces.deep(4), accumulate=ACCUM_PLAIN, projection=kind, ortho_size=size)
    print(f"projection parity {name} {w}x{h}", flush=True)
    o = Oracle(scene, detail, w, h)
    ref = o.probe()
    o.sample(4, 5)
    rc, _, _ = o.readback()
    with Trace.new_detailed(scene, detail, w, h) as t:
        got = t.probe()
        t.sample(4, 5)
        color = t.readback().color
    assert np.array_equal(got.primitive, ref["primitive"])
    np.testing.assert_array_equal(got.distance, ref["distance"])
    np.testing.assert_array_equal(got.depth, ref["depth"])
    assert rel_mse(color[..., :3], rc[..., :3]) <= 1e-4
    assert float(np.mean(color == rc)) > 0.99
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4), accumulate=ACCUM_PLAIN), w, h) as t:
        t.sample(1, 1)
        t.set_projection(kind, *size)
        assert t.samples() == 0
        t.sample(4, 5)
        assert np.array_equal(t.readback().color, color)
        t.set_projection(PERSPECTIVE)
        assert t.samples() == 0


@pytest.mark.parametrize("accumulate", [ACCUM_PLAIN, ACCUM_COMPENSATED])
def test_checkpoint_resume_continues_the_render(accumulate):
    """Trace::save / Trace::restore (checkpoint.rs:163-200): 4 samples, save,
    restore into a fresh trace, 4 more samples equals 8 uninterrupted ones
    (byte for byte with the plain sum; the compensated sum restarts its
    compensation from the saved FP32 sum)."""
    w, h = 48, 40
    scene = scenes.box_spheres(width=w, height=h)
    detail = Detail(bounces=Bounces.deep(4), accumulate=accumulate)
    print(f"checkpoint resume accumulate={accumulate}", flush=True)
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(4, 3)
        data = t.save("still")
        first = t.readback()
        t.sample(4, 3)
        full = t.readback()
        dev_full = t.readback_deviation()
    from pfx_amd.checkpoint import Saved

    s = Saved.read_from("still", data)
    assert s.size == (w, h) and s.samples == 4
    assert s.color == first.color.tobytes() and s.normal == first.normal.tobytes()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.restore("still", data)
        assert t.samples() == 4
        assert np.array_equal(t.readback().color, first.color) or accumulate != ACCUM_PLAIN
        # PFXCKPT1 holds no moments: the restored samples count at their mean
        # with no spread, so the noise channel reads 0 (to rounding) ...
        lum = first.color[..., :3].mean(axis=-1)
        assert np.all(t.readback_deviation() <= 1e-3 * lum + 1e-7)
        t.sample(4, 3)
        resumed = t.readback()
        # ... and after more samples it cannot exceed the uninterrupted one:
        # 4 mean^2 <= the sum of the 4 restored samples' squares
        dev_resumed = t.readback_deviation()
        assert np.isfinite(dev_resumed).all() and (dev_resumed >= 0).all() and dev_resumed.max() > 0
        assert np.all(dev_resumed <= dev_full * (1 + 1e-4) + 1e-6)
    if accumulate == ACCUM_PLAIN:
        assert np.array_equal(resumed.color, full.color)
    else:
        assert rel_mse(resumed.color[..., :3], full.color[..., :3]) <= 1e-10
    with Trace.new_detailed(scene, detail, 40, 40) as t:
        with pytest.raises(RuntimeError, match="the checkpoint is 48x40"):
            t.restore("still", data)
        with pytest.raises(RuntimeError, match="another render"):
            t.restore("other", data)


@pytest.mark.parametrize("case", sorted(CASES))
def test_deviation_matches_oracle(case):
    """The learned denoiser's noise channel (pfxh_trace_readback_deviation):
    the standard error of each pixel's mean luminance, from the luminance
    moments both sides sum in sample order, over two calls of 4 and 12 spp."""
    make, bounces, w, h = CASES[case]
    scene = make(w, h)
    detail = Detail(bounces=bounces)
    print(f"deviation parity {case} {w}x{h}", flush=True)
    o = Oracle(scene, detail, w, h)
    try:
        o.sample(4, 3)
        o.sample(12, 4)
        ref = o.deviation()
    finally:
        o.close()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(4, 3)
        t.sample(12, 4)
        got = t.readback_deviation()
    print(f"  exact {float(np.mean(got == ref)):.4f}, max |gap| {float(np.max(np.abs(got - ref))):.3g}", flush=True)
    assert np.all(np.isfinite(got)) and np.all(got >= 0.0)
    np.testing.assert_array_equal(got, ref)  # the same sums in the same order
