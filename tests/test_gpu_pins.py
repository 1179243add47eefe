"""Device parity at the configs' own sizes and on the reference's pins:

* the two-level test of levels.rs:591-681 cast on the device
  (pfxh_trace_cast): 3,000 rays against brute force over the flattened
  triangles, and, with the reference's left-first walk, id for id and
  distance for distance against the oracle;
* a render under a non-uniform HDR sky at equal spp (the sky slot's
  sampling, pdf and balance-heuristic MIS: trace.wgsl:2420-2448, sky.rs
  SKY_SAMPLING_WGSL);
* the small plain-method still at its full size, 512x512 x 16 spp with 4
  bounces (probe ids bit-exact, image within relMSE 1e-4);
* the 3840x2160 terrain's primary hits (8.3M rays): bit-exact with the
  left-first walk; with the default nearest-first walk the rays whose id
  differs are counted and bounded;
* batched stills against the oracle directly, and the 32-still uneven mix
  (256^2..2048^2, 10k-500k triangles, 1-64 lights, 4/16/64 spp) rendered as
  one batch, byte-identical to each still's own trace.
"""
import numpy as np
import pytest

import levels_case as L
from oracle import Oracle
from pfx_amd import Bounces, Detail, Trace, scenes
from pfx_amd.batch import Batch, Still
from test_gpu_batch import mix, same, single
from test_gpu_parity import rel_mse, small_terrain

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def test_two_level_pin_on_the_device():
    scene, placed, origins, directions = L.case()
    brute = np.array([L.triangle_hits(placed, o, d).min() for o, d in zip(origins, directions)])
    for left in (False, True):
        detail = Detail(left_first=left)
        print(f"levels pin on the device left_first={left}", flush=True)
        with Trace.new_detailed(scene, detail, 4, 4) as t:
            ids, dist = t.cast(origins, directions)
        hit = np.isfinite(brute)
        assert np.array_equal(ids != MISS, hit)
        gap = np.abs(dist[hit] - brute[hit])
        assert np.all(gap <= 1e-4 * np.maximum(brute[hit], 1.0)), float(gap.max())
        if left:
            ref = Oracle(scene, detail, 4, 4)
            rid, rt = ref.cast(origins, directions)
            ref.close()
            np.testing.assert_array_equal(ids, rid)
            np.testing.assert_array_equal(dist, rt)


def hdr_sky(w=64, h=32):
    """A lat-long sky with a gradient and a small bright patch (radiance up
    to 40): the sky CDF is far from uniform."""
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    vv, uu = np.meshgrid(v, u, indexing="ij")
    sky = np.zeros((h, w, 4), np.float32)
    up = np.clip(1.0 - vv * 1.6, 0.05, 1.0)
    sky[..., 0] = 0.3 * up + 0.05
    sky[..., 1] = 0.45 * up + 0.05
    sky[..., 2] = 0.9 * up + 0.08
    spot = np.exp(-(((uu - 0.3) / 0.03) ** 2 + ((vv - 0.25) / 0.04) ** 2))
    sky[..., :3] += (40.0 * spot)[..., None] * np.array([1.0, 0.9, 0.7])
    sky[..., 3] = 1.0
    return sky


@pytest.mark.parametrize("spp", [8])
def test_sky_lit_render_matches_the_oracle(spp):
    w, h = 96, 54
    scene = small_terrain(w, h)
    scene.sky = hdr_sky()
    detail = Detail(bounces=Bounces.deep(6))
    print(f"sky-lit terrain {w}x{h} spp={spp}", flush=True)
    o = Oracle(scene, detail, w, h)
    o.sample(spp, 11)
    rc, ra, rn = o.readback()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(spp, 11)
        out = t.readback()
    exact = float(np.mean(out.color == rc))
    rel = rel_mse(out.color[..., :3], rc[..., :3])
    print(f"  exact fraction {exact:.6f} relMSE {rel:.3e} mean {float(rc[..., :3].mean()):.4f}")
    assert float(rc[..., :3].mean()) > 0.05  # the sky does light the scene
    assert rel <= 1e-4
    assert exact > 0.99
    np.testing.assert_allclose(out.normal, rn, rtol=0, atol=1e-4)


def test_small_plain_still_at_full_size():
    """512x512, 16 spp, Bounces::deep(4), box of 12 triangles + 64
    icospheres of 320 + the emissive ceiling, scene seed 1, trace seed 1."""
    w = h = 512
    scene = scenes.box_spheres(seed=1, width=w, height=h)
    detail = Detail(bounces=Bounces.deep(4))
    print("box_spheres 512x512 x 16 spp against the oracle", flush=True)
    o = Oracle(scene, detail, w, h)
    ref_probe = o.probe()
    o.sample(16, 1)
    rc, _, _ = o.readback()
    with Trace.new_detailed(scene, detail, w, h) as t:
        probe = t.probe()
        t.sample(16, 1)
        out = t.readback()
    assert np.array_equal(probe.primitive, ref_probe["primitive"])
    np.testing.assert_array_equal(probe.distance, ref_probe["distance"])
    exact = float(np.mean(out.color == rc))
    rel = rel_mse(out.color[..., :3], rc[..., :3])
    print(f"  exact fraction {exact:.6f} relMSE {rel:.3e}")
    assert rel <= 1e-4
    assert exact > 0.99


@pytest.mark.timeout(900)
def test_terrain_4k_primary_hits():
    """The 3840x2160 terrain (1,048,576 heightfield triangles, 256 instances
    of 8 meshes, 128 lights): probe ids, instances and distances of all
    8,294,400 primary rays. Left first they equal the oracle's everywhere;
    nearest first the differing ids are counted (the reference's own answer
    depends on its walk order only for hits within an ulp of a box face)."""
    w, h = 3840, 2160
    scene = scenes.terrain(seed=2, width=w, height=h)
    print("4K terrain probe: oracle", flush=True)
    ref = Oracle(scene, Detail(bounces=Bounces.deep(6)), w, h)
    rp = ref.probe()
    ref.close()
    for left in (True, False):
        print(f"4K terrain probe: device left_first={left}", flush=True)
        with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(6), left_first=left), w, h) as t:
            got = t.probe()
        differ = got.primitive != rp["primitive"]
        print(f"  ids differ on {int(differ.sum())} of {w * h} rays")
        assert np.array_equal(got.primitive == MISS, rp["primitive"] == MISS)
        if left:
            assert not differ.any()
            assert np.array_equal(got.instance, rp["instance"])
            np.testing.assert_array_equal(got.distance, rp["distance"])
        else:
            assert int(differ.sum()) <= (w * h) // 100_000
            same_ = ~differ
            np.testing.assert_array_equal(got.distance[same_], rp["distance"][same_])
            a = got.distance[differ].view(np.int32).astype(np.int64)
            b = rp["distance"][differ].view(np.int32).astype(np.int64)
            assert np.all(np.abs(a - b) <= 4)


def test_batched_stills_match_the_oracle():
    """Each still of a batch against the oracle at its own seed and spp."""
    detail = Detail(bounces=Bounces.deep(4))
    stills = mix()
    print(f"batch of {len(stills)} stills against the oracle", flush=True)
    with Batch(stills, detail) as b:
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
    for k, (still, out) in enumerate(zip(stills, outs)):
        o = Oracle(still.scene, detail, still.width, still.height)
        o.sample(still.spp, still.seed)
        rc, ra, _ = o.readback()
        o.close()
        exact = float(np.mean(out.color == rc))
        rel = rel_mse(out.color[..., :3], rc[..., :3])
        print(f"  still {k} {still.width}x{still.height} x {still.spp}: exact {exact:.6f} relMSE {rel:.3e}")
        assert rel <= 1e-4
        assert exact > 0.99


@pytest.mark.timeout(900)
def test_uneven_32_still_batch_equals_single_traces():
    """The 32-still uneven mix at its own sizes, triangle counts, light
    counts and spp, autotuned and graph-captured, against each still's own
    trace byte for byte; the smallest still also against the oracle."""
    detail = Detail(bounces=Bounces.deep(6))
    print("32-still uneven mix: building scenes", flush=True)
    stills = [Still(*s) for s in scenes.batch_mix(seed=3)]
    print(f"  {sum(s.width * s.height for s in stills)} pixels, "
          f"{sum(s.width * s.height * s.spp for s in stills)} paths", flush=True)
    with Batch(stills, detail) as b:
        b.tune()
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
    for k, (still, out) in enumerate(zip(stills, outs)):
        print(f"  still {k}: {still.width}x{still.height} x {still.spp} spp", flush=True)
        same(out, single(still, detail))
    k = min(range(len(stills)), key=lambda i: stills[i].width * stills[i].height * stills[i].spp)
    still = stills[k]
    o = Oracle(still.scene, detail, still.width, still.height)
    o.sample(still.spp, still.seed)
    rc, _, _ = o.readback()
    o.close()
    rel = rel_mse(outs[k].color[..., :3], rc[..., :3])
    print(f"  oracle on still {k}: relMSE {rel:.3e}")
    assert rel <= 1e-4


def test_set_environment_and_sun_radius_match_the_oracle():
    """Trace::set_environment (gpu.rs:2810-2815) on a trace built under a
    black sky (its shadow-ray buffers grow by the sky's light slot), then
    Trace::set_sun_radius (:2903-2910) with a sun: each render equals the
    oracle's after the same setters, and each setter resets the count."""
    from pfx_amd import Sun

    w, h = 96, 54
    scene = small_terrain(w, h)
    scene.sun = Sun(direction=(0.3, 1.0, 0.2), color=(1.0, 0.95, 0.9), intensity=2.0)
    detail = Detail(bounces=Bounces.deep(6))
    o = Oracle(scene, detail, w, h)
    with Trace.new_detailed(scene, detail, w, h) as t:
        for step, (sky, radius) in enumerate([(hdr_sky(), None), (hdr_sky(), 3.0), (None, 0.5)]):
            print(f"setters step {step}: sky {'lit' if sky is not None else 'black'} radius {radius}", flush=True)
            t.sample(2, 1)
            t.set_environment(sky)
            o.set_environment(sky)
            if radius is not None:
                t.set_sun_radius(radius)
                o.set_sun_radius(radius)
            assert t.samples() == 0
            t.sample(4, 21 + step)
            o.sample(4, 21 + step)
            out = t.readback()
            rc, _, _ = o.readback()
            exact = float(np.mean(out.color == rc))
            rel = rel_mse(out.color[..., :3], rc[..., :3])
            print(f"  exact fraction {exact:.6f} relMSE {rel:.3e}")
            assert rel <= 1e-4 and exact > 0.99
        with pytest.raises(RuntimeError):
            t.set_sun_radius(90.0)
    o.close()


def test_batch_still_setters_reach_the_batch():
    """A setter on a batch still's lent trace (here set_camera) is what the
    next batch render uses: the still equals its own trace with that
    camera."""
    from pfx_amd import Camera

    detail = Detail(bounces=Bounces.deep(4))
    stills = mix()[:3]
    cam = Camera.look((2.0, 6.0, -9.0), (0.0, 0.0, 0.0), 50.0, stills[0].width / stills[0].height)
    with Batch(stills, detail) as b:
        b.trace(0).set_camera(cam)
        b.render()
        out = b.trace(0).readback()
    with Trace.new_detailed(stills[0].scene, detail, stills[0].width, stills[0].height) as t:
        t.set_camera(cam)
        t.sample(stills[0].spp, stills[0].seed)
        ref = t.readback()
    same(out, ref)


def test_device_limits_refuse_with_named_shortfalls():
    """gpu.rs:5138-5200 on this device: under a 380 MB buffer limit (the
    reference's test uses wgpu's 128 MiB binding limit; here the largest
    buffer of a small trace is its 352 MB of spilled traversal stacks, sized
    for every thread the GPU holds) a 64x64 trace builds and samples, and a
    3840x2160 one is refused (PFXH_REFUSED) naming the limit, the
    48-byte-per-pixel accumulator it needed and the limit it met; a memory
    limit is named the same way; the device's own limits come back
    afterwards."""
    import ctypes as C

    from pfx_amd import _native as N
    from pfx_amd.trace import TraceError

    lib = N.lib()
    scene = scenes.box_spheres(width=64, height=64)
    try:
        lib.pfxh_set_device_limits(C.byref(N.DeviceLimitsC(380_000_000, 0)))
        with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), 64, 64) as t:
            t.sample(2, 3)
            assert t.samples() == 2
        with pytest.raises(TraceError) as info:
            Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), 3840, 2160)
        assert info.value.refused
        msg = str(info.value)
        print(f"  refused: {msg}")
        assert msg.startswith(f"limit max_buffer_size of {3840 * 2160 * 48} is beyond the device's 380000000")
        lib.pfxh_set_device_limits(C.byref(N.DeviceLimitsC(0, 64 << 20)))
        with pytest.raises(TraceError) as info:
            Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), 64, 64)
        assert info.value.refused and str(info.value).startswith("limit max_memory of ")
    finally:
        lib.pfxh_set_device_limits(None)
    got = N.DeviceLimitsC()
    lib.pfxh_get_device_limits(C.byref(got))
    assert got.max_buffer_size == 0 and got.max_memory == 0
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), 3840, 2160) as t:
        assert t.gpu_bytes() > 3840 * 2160 * 48


def test_staged_scene_traces_like_the_oracle():
    """Staged::trace (stage.rs:383-387): a staged scene of alpha-cut
    icospheres in two materials and three placements, traced through
    pfxh_staged_trace's scene and detail, equals the oracle at equal spp."""
    from pfx_amd import Camera, point_light
    from pfx_amd.scene import material
    from pfx_amd.scenes import icosphere
    from pfx_amd.stage import IDENTITY, Placement, StageMesh, stage

    ico = icosphere(2).astype(np.float32)
    pos = ico.reshape(-1, 3)
    alpha = np.clip(pos[:, 1] * 2.0 + 0.6, 0.0, 1.0).astype(np.float32)  # the bottom is cut away
    mesh = StageMesh(pos, pos, np.arange(len(pos), dtype=np.uint32), alpha)
    moved, lifted = IDENTITY.copy(), IDENTITY.copy()
    moved[3] = [2.5, 0.0, 0.0, 1.0]
    lifted[3] = [-2.5, 0.5, 1.0, 1.0]
    mats = np.concatenate([material((0.8, 0.3, 0.2)), material((0.7, 0.7, 0.75), roughness=0.2, metalness=1.0)])
    lights = np.concatenate([point_light((0.0, 4.0, -3.0), (1, 1, 1), 40.0, 30.0),
                             point_light((3.0, 2.0, -2.0), (1, 0.8, 0.6), 10.0, 30.0)])
    w, h = 96, 64
    staged = stage([mesh], [Placement(0, IDENTITY, 0), Placement(0, moved, 1), Placement(0, lifted, 0, two_sided=False)],
                   mats, lights=lights, sky=hdr_sky(),
                   camera=Camera.look((0.0, 1.5, -7.0), (0.0, 0.0, 0.0), 50.0, w / h))
    detail = staged.detail
    detail.bounces = Bounces.deep(4)
    o = Oracle(staged.scene, detail, w, h)
    o.sample(8, 3)
    rc, _, _ = o.readback()
    o.close()
    with Trace.new_detailed(staged.scene, detail, w, h) as t:
        t.sample(8, 3)
        out = t.readback()
    rel = rel_mse(out.color[..., :3], rc[..., :3])
    print(f"  staged scene relMSE {rel:.3e}, exact {float(np.mean(out.color == rc)):.6f}")
    assert rel <= 1e-4 and float(np.mean(out.color == rc)) > 0.99
