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
from test_watertight import CENTRE, edge_targets, ico_scene

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


def moller_trumbore(tri, o, d):
    """trace.wgsl:780-804 in float32 numpy, in the oracle's operation order
    (cross products as written, dot = (x x + y y) + z z, no fused
    multiply-add): (hit, t) of each ray against its triangle. It reproduces
    the oracle's distance bit for bit on every hit of this test's rays."""
    f = np.float32

    def cross(a, b):
        return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                         a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], -1)

    def dot(a, b):
        return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]

    a = tri[:, 0]
    e1, e2 = tri[:, 1] - a, tri[:, 2] - a
    p = cross(d, e2)
    det = dot(e1, p)
    ok = np.abs(det) >= f(1e-9)
    inv = f(1) / np.where(ok, det, f(1))
    s = o - a
    u = dot(s, p) * inv
    q = cross(s, e1)
    v = dot(d, q) * inv
    t = dot(e2, q) * inv
    return ok & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t > f(1e-4)), t


def oracle_cast(scene, detail, o, d):
    orc = Oracle(scene, detail, 4, 4)
    try:
        return orc.cast(o, d)
    finally:
        orc.close()


@pytest.mark.parametrize("instanced", [False, True])
def test_edge_and_vertex_rays_against_the_oracle(instanced):
    """Rays aimed at every vertex and at points along every edge of a closed
    icosphere, from its centre and from scattered inner points, cast on the
    device (pfxh_trace_cast) with either triangle test.

    The reference's closest hit depends on its walk order for such rays: its
    box entry quotient and its triangle distance round differently, so a box
    whose face holds the hit point can be culled by a hit 1-2 ulps farther.
    Of these 46080 world rays the oracle's answer is not the nearest hit for
    170 (Moller-Trumbore) and 70 (watertight); at exact ties the same
    culling decides which of the tied triangles is kept. The near-first
    device walk may then keep another triangle of the group (world scene:
    171 and 109 rays; instanced: 204 and 75, with quantized nodes). Its box test is also conservative (it accepts every
    box the exact quotients accept, and a few more within rounding), so where
    Moller-Trumbore lets a ray leak between two triangles the device may test
    a neighbour the oracle's walk culled, and hit it (instanced: 42 rays).
    What holds, and is
    asserted: every ray the oracle hits, the device hits (watertight: no
    misses at all), and a ray only the device hits hits the surface at the
    aimed point (t within 1e-5 of 1); of the rays both hit, at least 98% of
    ids are equal, with equal distances; every other device distance is
    within 4 ulps of the oracle's and, for world triangles with
    Moller-Trumbore, is that triangle's own distance bit for bit."""
    scene, world = ico_scene(instanced=instanced)
    targets = edge_targets(world)
    rng = np.random.default_rng(5)
    o = np.concatenate([np.tile(CENTRE, (len(targets), 1)),
                        (CENTRE + rng.uniform(-0.8, 0.8, (len(targets), 3))).astype(np.float32)])
    d = (np.concatenate([targets, targets]) - o).astype(np.float32)
    leaks = {}
    for wt in (False, True):
        print(f"edge rays instanced={instanced} watertight={wt} n={len(o)}", flush=True)
        detail = Detail(watertight=wt)
        with Trace.new_detailed(scene, detail, 4, 4) as t:
            ids, dist = t.cast(o, d)
        rid, rt = oracle_cast(scene, detail, o, d)
        assert not np.any((ids == MISS) & (rid != MISS))
        rescued = (ids != MISS) & (rid == MISS)
        print(f"  hit only on the device: {int(rescued.sum())} rays")
        assert np.all(np.abs(dist[rescued] - 1.0) <= 1e-5)
        both = (ids != MISS) & (rid != MISS)
        differ = np.flatnonzero(both & (ids != rid))
        print(f"  ids differ on {len(differ)} rays")
        assert len(differ) <= len(o) // 50
        if not instanced and not wt:
            hit = rid != MISS
            ok, mt_t = moller_trumbore(world[rid[hit]], o[hit], d[hit])
            assert np.all(ok) and np.array_equal(mt_t, rt[hit])  # the restatement is the oracle's
            ok, mt_t = moller_trumbore(world[ids[differ]], o[differ], d[differ])
            assert np.all(ok)
            np.testing.assert_array_equal(dist[differ], mt_t)
        same = ids == rid
        np.testing.assert_array_equal(dist[same], rt[same])
        ulps = np.abs(dist[differ].view(np.int32).astype(np.int64) - rt[differ].view(np.int32).astype(np.int64))
        assert np.all(ulps <= 4), int(ulps.max())
        leaks[wt] = int(np.sum(ids == MISS))
    print(f"  leaks: Moller-Trumbore {leaks[False]}, watertight {leaks[True]} of {len(o)}")
    assert leaks[True] == 0 and leaks[False] > 0


def test_cast_limit_and_empty_list():
    scene, world = ico_scene(sub=2)
    o = np.tile(CENTRE, (64, 1))
    d = np.random.default_rng(3).normal(size=(64, 3)).astype(np.float32)
    with Trace.new_detailed(scene, Detail(watertight=True), 4, 4) as t:
        ids, dist = t.cast(o, d)
        assert np.all(ids != MISS)
        near, ndist = t.cast(o, d, limit=float(np.min(dist)) * 0.5)
        assert np.all(near == MISS) and np.all(ndist == 0)
        e_ids, e_t = t.cast(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
        assert len(e_ids) == 0 and len(e_t) == 0


def test_batch_of_watertight_stills_equals_single_traces():
    detail = Detail(bounces=Bounces.deep(4), watertight=True)
    stills = mix()
    with Batch(stills, detail) as b:
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
    for still, out in zip(stills, outs):
        same(out, single(still, detail))


@pytest.mark.parametrize("instanced", [False, True])
def test_left_first_walk_keeps_the_oracles_hit_on_every_edge_and_vertex_ray(instanced):
    """With the reference's left-first walk (PFXH_WALK_LEFT_FIRST) the device
    tests the reference's triangles in the reference's order, so on the
    same 46,080 edge and vertex rays where the nearest-first walk keeps
    another triangle for a few hundred, every id and every distance equals
    the oracle's (trace.wgsl:999-1075, push order :1049-1054), with either
    triangle test."""
    scene, world = ico_scene(instanced=instanced)
    targets = edge_targets(world)
    rng = np.random.default_rng(5)
    o = np.concatenate([np.tile(CENTRE, (len(targets), 1)),
                        (CENTRE + rng.uniform(-0.8, 0.8, (len(targets), 3))).astype(np.float32)])
    d = (np.concatenate([targets, targets]) - o).astype(np.float32)
    for wt in (False, True):
        print(f"left-first edge rays instanced={instanced} watertight={wt} n={len(o)}", flush=True)
        detail = Detail(watertight=wt, left_first=True)
        with Trace.new_detailed(scene, detail, 4, 4) as t:
            ids, dist = t.cast(o, d)
        rid, rt = oracle_cast(scene, detail, o, d)
        print(f"  ids differ on {int(np.sum(ids != rid))} rays")
        np.testing.assert_array_equal(ids, rid)
        np.testing.assert_array_equal(dist, rt)


@pytest.mark.parametrize("case", sorted(CASES))
def test_left_first_images_match_the_oracle(case):
    make, bounces, w, h = CASES[case]
    scene = make(w, h)
    detail = Detail(bounces=bounces, left_first=True)
    print(f"left-first image parity {case} {w}x{h}", flush=True)
    o = Oracle(scene, detail, w, h)
    o.sample(8, 7)
    rc, _, _ = o.readback()
    with Trace.new_detailed(scene, detail, w, h) as t:
        t.sample(8, 7)
        out = t.readback()
    exact = float(np.mean(out.color == rc))
    print(f"  exact fraction {exact:.6f}")
    assert rel_mse(out.color[..., :3], rc[..., :3]) <= 1e-4
    assert exact > 0.99
