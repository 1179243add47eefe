"""Staging (pfxh_stage_build), against the reference's own stage tests
(crates/trace/src/stage.rs:889-969) restated for the parts built here, and
the staged scene against a scene built by hand."""
import numpy as np
import pytest

from pfx_amd import _native as N
from pfx_amd.scene import material
from pfx_amd.stage import IDENTITY, Placement, StageMesh, stage

POSITIONS = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], np.float32)  # stage.rs:865
NORMALS = np.array([[0.0, 0.0, 1.0]] * 3, np.float32)


def triangle(alpha=None):
    return StageMesh(POSITIONS, NORMALS, np.array([0, 1, 2], np.uint32), alpha)


def materials(n):
    return np.concatenate([material() for _ in range(n)])


def area(v9):
    a, b, c = v9[0:3], v9[3:6], v9[6:9]
    e1, e2 = b - a, c - a
    return 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0])


def placed_triangles(staged, k):
    inst = staged.scene.instances[k]
    m = staged.scene.meshes[inst["mesh"]]
    return staged.scene.mesh_triangles[m["first"]:m["first"] + m["count"]], inst


def test_instances_share_a_mesh_with_their_own_material_and_transform():
    """stage.rs:889-908"""
    moved = IDENTITY.copy()
    moved[3] = [2.0, 0.0, -1.0, 1.0]
    staged = stage([triangle()], [Placement(0, IDENTITY, 1), Placement(0, moved, 0)], materials(2))
    sc = staged.scene
    assert len(sc.triangles) == 0 and len(sc.world_details) == 0
    assert len(sc.materials) == 2 and len(sc.instances) == 2
    t0, _ = placed_triangles(staged, 0)
    t1, i1 = placed_triangles(staged, 1)
    assert t0[0]["material"] == 1 and t1[0]["material"] == 0
    m = i1["transform"].reshape(4, 4)  # columns
    b = t1[0]["vertices"][3:6]
    placed = m[3, :3] + m[0, :3] * b[0] + m[1, :3] * b[1] + m[2, :3] * b[2]
    np.testing.assert_array_equal(placed, [3.0, 0.0, -1.0])


def test_placements_of_one_mesh_and_material_share_the_staged_mesh():
    staged = stage([triangle()], [Placement(0, IDENTITY, 0), Placement(0, IDENTITY, 0), Placement(0, IDENTITY, 1),
                                  Placement(0, IDENTITY, 0, alpha_cutoff=0.25)], materials(2))
    meshes = staged.scene.instances["mesh"].tolist()
    assert meshes[0] == meshes[1] and len(set(meshes)) == 3


@pytest.mark.parametrize("alpha,count,expected", [([1.0, 1.0, 1.0], 1, 0.5), ([0.0, 0.0, 0.0], 0, 0.0),
                                                  ([1.0, 0.0, 0.0], 1, 0.125), ([1.0, 1.0, 0.0], 2, 0.375)])
def test_alpha_below_the_cutoff_is_cut(alpha, count, expected):
    """stage.rs:910-933: the kept triangles' count and area; each kept vertex
    lies in the triangle's plane, and the interpolated normals stay the
    (constant) corner normal."""
    staged = stage([triangle(np.array(alpha, np.float32))], [Placement(0, IDENTITY, 0)], materials(1))
    kept, _ = placed_triangles(staged, 0)
    assert len(kept) == count, alpha
    total = sum(area(t["vertices"]) for t in kept)
    assert abs(total - expected) < 1e-6, (alpha, total)
    for t in kept:
        v = t["vertices"].reshape(3, 3)
        assert np.all(v[:, 2] == 0.0) and np.all(v[:, 0] >= 0) and np.all(v[:, 1] >= 0)
        assert np.all(v[:, 0] + v[:, 1] <= 1.0 + 1e-7)
        np.testing.assert_array_equal(t["normals"].reshape(3, 3), NORMALS)


def test_broken_input_is_refused():
    """stage.rs:935-969"""
    good = [Placement(0, IDENTITY, 0)]
    short = StageMesh(POSITIONS, NORMALS, np.array([0, 1], np.uint32))
    outside = StageMesh(POSITIONS, NORMALS, np.array([0, 1, 3], np.uint32))
    for mesh in (short, outside):
        with pytest.raises(ValueError):
            stage([mesh], good, materials(1))
    with pytest.raises(ValueError, match="missing mesh"):
        stage([triangle()], [Placement(1, IDENTITY, 0)], materials(1))
    with pytest.raises(ValueError, match="missing material"):
        stage([triangle()], [Placement(0, IDENTITY, 1)], materials(1))
    with pytest.raises(ValueError, match="at least one material"):
        stage([triangle()], good, np.zeros(0, N.MATERIAL))
    with pytest.raises(ValueError, match="projection or lens"):
        stage([triangle()], good, materials(1), projection=N.PROJ_ORTHOGRAPHIC, ortho_size=(0.0, 1.0))
    with pytest.raises(ValueError, match="alpha cutoff"):
        stage([triangle()], [Placement(0, IDENTITY, 0, alpha_cutoff=1.5)], materials(1))


def test_placements_carry_their_shadow_flags():
    """stage.rs:971-1065, the shadow flag (content slots are not staged
    here): casts_shadow false, shadow_only and one-sided placements set the
    instance surface words the tracer reads."""
    staged = stage([triangle()], [Placement(0, casts_shadow=False), Placement(0, shadow_only=True),
                                  Placement(0, two_sided=False), Placement(0)], materials(1))
    words = staged.scene.instances["word"].tolist()
    assert words == [2, 1, 1 << 23, 0]


def test_staged_detail_is_the_default_with_the_stage_lens():
    staged = stage([triangle()], [Placement(0)], materials(1), aperture=0.1, focus_distance=3.0,
                   lens_shift=(0.25, -0.5))
    d = staged.detail
    assert (d.bounces.total, d.bounces.diffuse, d.bounces.glossy, d.bounces.transmission) == (16, 4, 8, 12)
    assert d.aperture == pytest.approx(0.1) and d.focus_distance == 3.0 and tuple(d.lens_shift) == (0.25, -0.5)


def test_staged_scene_lays_out_like_the_hand_built_one():
    """The staged scene's two-level layout equals that of the same instances
    built by hand (the same mesh triangles, ranges and words)."""
    import ctypes as C

    from pfx_amd.scenes import icosphere

    ico = icosphere(1).astype(np.float32)  # (n, 3, 3)
    pos = ico.reshape(-1, 3)
    idx = np.arange(len(pos), dtype=np.uint32)
    moved = IDENTITY.copy()
    moved[3] = [3.0, 0.5, 0.0, 1.0]
    staged = stage([StageMesh(pos, pos, idx)], [Placement(0, IDENTITY, 0), Placement(0, moved, 1)], materials(2))
    tris = np.zeros(2 * len(ico), N.TRIANGLE_DETAIL)
    tris["vertices"] = np.concatenate([ico, ico]).reshape(-1, 9)
    tris["normals"] = tris["vertices"]
    tris["material"] = [0] * len(ico) + [1] * len(ico)
    np.testing.assert_array_equal(staged.scene.mesh_triangles, tris)
    assert staged.scene.meshes.tolist() == [(0, len(ico)), (len(ico), len(ico))]

    def levels(scene):
        sc, keep = scene.to_c()
        nc, oc = C.c_uint32(), C.c_uint32()
        err = C.create_string_buffer(256)
        assert N.lib().pfxh_layout_levels(C.byref(sc), None, C.byref(nc), None, C.byref(oc), err, 256) == 0
        nodes = np.zeros(nc.value, N.BVH_NODE)
        order = np.zeros(oc.value, np.uint32)
        assert N.lib().pfxh_layout_levels(C.byref(sc), nodes.ctypes.data, C.byref(nc), order.ctypes.data,
                                          C.byref(oc), err, 256) == 0
        return nodes

    from pfx_amd.scene import Scene

    hand = Scene(mesh_triangles=tris, meshes=np.array([(0, len(ico)), (len(ico), len(ico))], N.MESH),
                 instances=staged.scene.instances.copy(), materials=materials(2))
    assert levels(staged.scene).tobytes() == levels(hand).tobytes()
