"""The watertight ray-triangle test (Woop, Benthin and Wald, JCGT 2(1),
2013), PFXH_TRI_WATERTIGHT, on the CPU oracle.

What pins it: the property that defines it -- a ray that crosses a closed
mesh through a shared vertex or edge still hits the mesh -- on a closed
icosphere (world and instanced) and on a heightfield, where the reference's
Moller-Trumbore test (trace.wgsl:780-804) lets some of the same rays through;
agreement with Moller-Trumbore (same triangle, distance and barycentrics to
rounding) on rays away from edges; the BVH walk against a brute-force loop
over every triangle with the same test. The device runs the same test bit
for bit (tests/test_gpu_watertight.py)."""
import numpy as np
import pytest

from oracle import Oracle
from pfx_amd import Detail, scenes
from pfx_amd import _native as N
from pfx_amd.scene import Camera, Scene, material

MISS = 0xFFFFFFFF
CENTRE = np.array([0.3, -0.2, 0.1], np.float32)


def ico_scene(sub=3, radius=1.7, instanced=False):
    tris = scenes.icosphere(sub)
    mats = material((0.5, 0.5, 0.5), roughness=1.0, specular=0.0)
    cam = Camera.look((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 60.0, 1.0)
    if not instanced:
        world = (tris * radius + CENTRE.astype(np.float64)).astype(np.float32)
        return Scene(triangles=scenes._triangles(world, np.zeros(len(tris))), materials=mats, camera=cam), world
    unit = tris.astype(np.float32)
    scene = Scene(materials=mats, camera=cam)
    rows = np.zeros(len(unit), N.TRIANGLE_DETAIL)
    rows["vertices"] = unit.reshape(len(unit), 9)
    scene.mesh_triangles = rows
    scene.meshes = np.array([(0, len(unit))], N.MESH)
    inst = np.zeros(1, N.INSTANCE)
    inst[0]["mesh"] = 0
    inst[0]["transform"] = scenes.transform(tuple(float(c) for c in CENTRE), 0.7, radius)
    scene.instances = inst
    m = np.asarray(inst[0]["transform"], np.float64).reshape(4, 4)  # m[column][row]
    world = (unit.astype(np.float64) @ m[:3, :3] + m[3, :3]).astype(np.float32)
    return scene, world


def edge_targets(tris):
    """Every vertex and points along every edge, in float32."""
    out = []
    for t in tris:
        for i in range(3):
            a, b = t[i], t[(i + 1) % 3]
            out.append(a)
            for s in (0.5, 0.25, 1.0 / 3.0, 0.1, 0.9):
                out.append(a + np.float32(s) * (b - a))
    return np.array(out, np.float32)


def cast(scene, o, d, watertight):
    orc = Oracle(scene, Detail(watertight=watertight), 4, 4)
    try:
        return orc.cast(o, d)
    finally:
        orc.close()


@pytest.mark.parametrize("instanced", [False, True])
def test_rays_through_shared_edges_and_vertices_never_leak(instanced):
    scene, world = ico_scene(instanced=instanced)
    targets = edge_targets(world)
    rng = np.random.default_rng(5)
    origins = [np.tile(CENTRE, (len(targets), 1)),
               (CENTRE + rng.uniform(-0.8, 0.8, (len(targets), 3))).astype(np.float32)]
    leaks = {}
    for o in origins:
        d = (targets - o).astype(np.float32)
        ids, _ = cast(scene, o, d, watertight=True)
        assert np.all(ids != MISS)
        mt, _ = cast(scene, o, d, watertight=False)
        leaks.setdefault("mt", 0)
        leaks["mt"] += int(np.sum(mt == MISS))
    print(f"Moller-Trumbore lets {leaks['mt']} of {2 * len(targets)} edge and vertex rays through")
    # the reference's test has cracks on this mesh; the watertight one has none
    assert leaks["mt"] > 0


def test_heightfield_has_no_cracks():
    nx, nz = 24, 16
    rng = np.random.default_rng(8)
    h = rng.uniform(-0.5, 0.5, (nz + 1, nx + 1))
    verts = scenes.heightfield_triangles(h, 12.0, 8.0).astype(np.float32)
    tris = scenes._triangles(verts, np.zeros(len(verts)))
    mats = material((0.5, 0.5, 0.5), roughness=1.0, specular=0.0)
    cam = Camera.look((0.0, 5.0, -1.0), (0.0, 0.0, 0.0), 60.0, 1.0)
    scene = Scene(triangles=tris, materials=mats, camera=cam)
    targets = edge_targets(verts)
    inner = (np.abs(targets[:, 0]) < 5.9) & (np.abs(targets[:, 2]) < 3.9)
    targets = targets[inner]
    # Rays from above within 20 degrees of vertical: the steepest triangle
    # here is under 64 degrees, so every triangle faces every ray and each
    # ray crosses the surface (a ray that only touches a silhouette edge
    # may rightly miss, with any test).
    leaks = 0
    for dist in (6.0, 2.0):
        tilt = np.radians(rng.uniform(0.0, 20.0, len(targets)))
        az = rng.uniform(0.0, 2.0 * np.pi, len(targets))
        up = np.column_stack([np.sin(tilt) * np.cos(az), np.cos(tilt), np.sin(tilt) * np.sin(az)])
        o = (targets + dist * up).astype(np.float32)
        d = (targets - o).astype(np.float32)
        ids, _ = cast(scene, o, d, watertight=True)
        assert np.all(ids != MISS), int(np.sum(ids == MISS))
        mt, _ = cast(scene, o, d, watertight=False)
        leaks += int(np.sum(mt == MISS))
    print(f"Moller-Trumbore lets {leaks} of {2 * len(targets)} heightfield edge and vertex rays through")


def random_triangles(rng, n):
    v = rng.uniform(-4, 4, (n, 1, 3)) + rng.uniform(-1, 1, (n, 3, 3))
    return v.astype(np.float32)


def test_agrees_with_moller_trumbore_away_from_edges():
    rng = np.random.default_rng(21)
    verts = random_triangles(rng, 300)
    mats = material((0.5, 0.5, 0.5), roughness=1.0, specular=0.0)
    scene = Scene(triangles=scenes._triangles(verts, np.zeros(len(verts))), materials=mats,
                  camera=Camera.look((0, 0, -10), (0, 0, 0), 60.0, 1.0))
    o = rng.uniform(-8, 8, (4000, 3)).astype(np.float32)
    d = (rng.uniform(-3, 3, (4000, 3)) - o).astype(np.float32)
    wid, wt = cast(scene, o, d, True)
    mid, mtt = cast(scene, o, d, False)
    both = (wid != MISS) & (mid != MISS)
    assert both.sum() > 1000
    same = wid[both] == mid[both]
    # the two tests differ only for rays within rounding of an edge
    assert same.mean() > 0.999
    np.testing.assert_allclose(wt[both][same], mtt[both][same], rtol=2e-5)
    assert abs(int((wid != MISS).sum()) - int((mid != MISS).sum())) <= 4


def test_bvh_walk_equals_brute_force():
    """The closest watertight hit found through the BVH is the nearest over
    all triangles tested one by one (each triangle alone in a scene)."""
    rng = np.random.default_rng(22)
    verts = random_triangles(rng, 40)
    mats = material((0.5, 0.5, 0.5), roughness=1.0, specular=0.0)
    cam = Camera.look((0, 0, -10), (0, 0, 0), 60.0, 1.0)
    o = rng.uniform(-8, 8, (500, 3)).astype(np.float32)
    d = (rng.uniform(-3, 3, (500, 3)) - o).astype(np.float32)
    whole = Scene(triangles=scenes._triangles(verts, np.zeros(len(verts))), materials=mats, camera=cam)
    ids, t = cast(whole, o, d, True)
    best = np.full(len(o), np.inf, np.float32)
    for k in range(len(verts)):
        one = Scene(triangles=scenes._triangles(verts[k:k + 1], np.zeros(1)), materials=mats, camera=cam)
        hid, ht = cast(one, o, d, True)
        best = np.where((hid != MISS) & (ht < best), ht, best)
    hit = ids != MISS
    assert np.array_equal(hit, np.isfinite(best))
    np.testing.assert_array_equal(t[hit], best[hit])


def test_unknown_triangle_test_is_refused(native_lib):
    import ctypes as C

    scene, _ = ico_scene(sub=1)
    sc, keep = scene.to_c()
    d = Detail().to_c()
    d.triangle_test = 2
    err = N.err_buffer()
    out = C.c_void_p()
    assert native_lib.pfxh_trace_new(C.byref(sc), C.byref(d), 8, 8, C.byref(out), err, len(err)) == N.SETUP
    assert b"invalid trace detail" in err.value
