"""Pins of the oracle's restated layout functions, from the reference's own
known answers and invariants (crates/trace/src/bvh.rs, lights.rs, tree.rs,
adaptive.rs, sky.rs tests), plus the product's host layout matching it byte
for byte.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import _native as N
from pfx_amd import scenes
from pfx_amd.scene import disc_light, point_light, rect_light, spot_light


def lcg17():
    """bvh.rs:331-335: an LCG (Numerical Recipes constants) seeded 17."""
    state = [17]

    def random():
        state[0] = (state[0] * 1664525 + 1013904223) % 2**32
        return np.float32((state[0] >> 8) / float(1 << 24))

    return random


def random_triangles(random, count=120):
    tris = np.zeros(count, N.TRIANGLE)
    for i in range(count):
        x = random() * np.float32(8.0) - np.float32(4.0)
        y = random() * np.float32(8.0) - np.float32(4.0)
        z = random() * np.float32(8.0) - np.float32(4.0)
        bx = x + random() + np.float32(0.1)
        cy = y + random() + np.float32(0.1)
        tris[i]["vertices"] = [x, y, z, bx, y, z, x, cy, z]
    return tris


def bvh_of(tris):
    nodes = np.zeros(2 * len(tris), N.BVH_NODE)
    order = np.zeros(len(tris), np.uint32)
    n = O.lib().or_bvh(tris.ctypes.data, len(tris), nodes.ctypes.data, order.ctypes.data)
    return nodes[:n], order


def test_bvh_matches_brute_force_and_is_stable():
    """bvh.rs:330-381: 120 random triangles, 2000 random rays."""
    random = lcg17()
    tris = random_triangles(random)
    nodes, order = bvh_of(tris)
    nodes2, order2 = bvh_of(tris)
    assert nodes.tobytes() == nodes2.tobytes() and np.array_equal(order, order2)
    lib = O.lib()
    for _ in range(2000):
        o = np.float32([random() * np.float32(12) - np.float32(6) for _ in range(3)])
        d = np.float32([random() * np.float32(2) - np.float32(1) for _ in range(3)])
        d = (d * (np.float32(1) / np.sqrt(np.float32(d @ d)))).astype(np.float32)
        fast = lib.or_bvh_intersect(tris.ctypes.data, len(tris), nodes.ctypes.data, len(nodes), order.ctypes.data,
                                    o.ctypes.data, d.ctypes.data)
        best, best_t = 0xFFFFFFFF, np.inf
        for i in range(len(tris)):
            t = lib.or_triangle_hit(tris[i]["vertices"].ctypes.data, o.ctypes.data, d.ctypes.data, 1e30)
            if t >= 0 and t < best_t:
                best, best_t = i, t
        assert fast == best


def test_bvh_structure_invariants():
    random = lcg17()
    tris = random_triangles(random, 500)
    nodes, order = bvh_of(tris)
    assert sorted(order.tolist()) == list(range(500))
    covered = np.zeros(500, int)
    v = tris["vertices"].reshape(-1, 3, 3)
    for k, n in enumerate(nodes):
        if n["count"] == 0:
            for c in (k + 1, n["first"]):  # children boxes lie inside the parent's
                assert np.all(nodes[c]["lo"][:3] >= n["lo"][:3]) and np.all(nodes[c]["hi"][:3] <= n["hi"][:3])
        else:
            assert n["count"] <= 16
            for j in order[n["first"]:n["first"] + n["count"]]:
                covered[j] += 1
                assert np.all(v[j].min(0) >= n["lo"][:3]) and np.all(v[j].max(0) <= n["hi"][:3])
    assert np.all(covered == 1)


def _product_bvh(tris):
    nodes = np.zeros(2 * len(tris), N.BVH_NODE)
    order = np.zeros(len(tris), np.uint32)
    count = C.c_uint32()
    err = C.create_string_buffer(256)
    st = N.lib().pfxh_layout_bvh(tris.ctypes.data, len(tris), nodes.ctypes.data, C.byref(count), order.ctypes.data,
                                 err, 256)
    assert st == 0, err.value
    return nodes[: count.value], order


def test_product_bvh_is_byte_identical():
    random = lcg17()
    tris = random_triangles(random, 3000)
    a = _product_bvh(tris)
    b = bvh_of(tris)
    assert a[0].tobytes() == b[0].tobytes() and np.array_equal(a[1], b[1])


def _levels(scene, product):
    sc, keep = scene.to_c()
    nc, oc = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    call = N.lib().pfxh_layout_levels if product else O.lib().or_levels
    extra = () if product else (None,)
    assert call(C.byref(sc), None, C.byref(nc), None, C.byref(oc), *extra, err, 512) == 0, err.value
    nodes = np.zeros(nc.value, N.BVH_NODE)
    order = np.zeros(oc.value, np.uint32)
    assert call(C.byref(sc), nodes.ctypes.data, C.byref(nc), order.ctypes.data, C.byref(oc), *extra, err, 512) == 0
    return nodes, order


@pytest.mark.parametrize("which", ["box", "terrain"])
def test_product_two_level_layout_is_byte_identical(which):
    if which == "box":
        scene = scenes.box_spheres(width=32, height=32)
    else:
        scene = scenes.terrain(seed=4, grid=(48, 24), size=(24.0, 12.0), meshes=3, mesh_tris=(100, 400),
                               instances=10, lights=8, width=32, height=18)
    a, b = _levels(scene, True), _levels(scene, False)
    assert a[0].tobytes() == b[0].tobytes()
    assert np.array_equal(a[1], b[1])
    if which == "terrain":
        # levels.rs: top level is 3*groups-1 rows and holds one leaf per group
        groups = 1 + 10
        top = a[0][: 3 * groups - 1]
        leaves = top[: 2 * groups - 1][top[: 2 * groups - 1]["pad"][:, 0] == 1]
        assert len(leaves) == groups


def test_emitters_are_weighted_by_power():
    """lights.rs:465-507: two unit-half triangles with emission 1, two of a
    size-2 quad with emission 0.5."""

    def quad(size, height):
        return [
            [0, height, 0, size, height, 0, size, height, size],
            [0, height, 0, size, height, size, 0, height, size],
        ]

    rows = []
    for v in quad(1.0, 2.0):
        rows.append(v + [1.0, 1.0, 1.0])
    for v in quad(2.0, 2.0):
        rows.append(v + [0.5, 0.5, 0.5])
    tris = np.array(rows, np.float32)
    entries = np.zeros((5, 16), np.float32)
    cdf = np.zeros(4, np.float32)
    n = O.lib().or_emitters(tris.ctypes.data, 4, 1, 0, entries.ctypes.data, cdf.ctypes.data)
    assert n == 5
    assert entries[0, 0] == 4.0 and entries[0, 1] == 1.0
    assert abs(entries[0, 2] - 3.0) < 1e-5
    want = [1 / 6, 1 / 6, 1 / 3, 1 / 3]
    assert np.allclose(entries[1:, 15], want, atol=1e-6)
    assert np.allclose(cdf, np.cumsum(want), atol=1e-6)
    assert cdf[3] == 1.0
    assert entries[3, 7] == 2.0  # record
    assert abs(entries[3, 11] - 2.0) < 1e-6  # area


def test_empty_emitter_list_keeps_its_header():
    """lights.rs:509-515."""
    entries = np.zeros((1, 16), np.float32)
    cdf = np.zeros(1, np.float32)
    assert O.lib().or_emitters(None, 0, 2, 0, entries.ctypes.data, cdf.ctypes.data) == 1
    assert entries[0, 0] == 0.0 and entries[0, 1] == 2.0


def _pack(light):
    out = np.zeros(16, np.float32)
    O.lib().or_light_pack(light.ctypes.data, out.ctypes.data)
    return out.reshape(4, 4)


def test_lights_validate_and_pack():
    """lights.rs:578-628."""
    lib = O.lib()
    point = point_light((0.0, 2.0, 0.0), (1.0, 0.5, 0.25), 4.0, 10.0)
    assert lib.or_light_valid(point.ctypes.data)
    packed = _pack(point)
    assert list(packed[1]) == [4.0, 2.0, 1.0, 10.0]
    assert packed[3, 3] == 1.0
    want = 4.0 * (1.0 - 0.2**4) ** 2 / 4.0
    assert abs(lib.or_light_falloff(point.ctypes.data, 2.0) - want) < 1e-6
    assert lib.or_light_falloff(point.ctypes.data, 10.0) == 0.0
    spot = spot_light((0.0, 2.0, 0.0), (1.0, 0.5, 0.25), 4.0, 10.0, (0.0, -1.0, 0.0), 20.0, 30.0)
    assert lib.or_light_valid(spot.ctypes.data)
    down = np.float32([0.0, -1.0, 0.0])
    assert lib.or_light_cone(spot.ctypes.data, down.ctypes.data) == 1.0
    side = np.float32([1.0, -1.0, 0.0])
    assert lib.or_light_cone(spot.ctypes.data, side.ctypes.data) == 0.0
    edge = np.float32([np.tan(np.radians(25.0)), -1.0, 0.0])
    e = lib.or_light_cone(spot.ctypes.data, edge.ctypes.data)
    assert 0.0 < e < 1.0
    wide = spot_light((0.0, 2.0, 0.0), (1.0, 0.5, 0.25), 4.0, 10.0, (0.0, -1.0, 0.0), 40.0, 30.0)
    assert not lib.or_light_valid(wide.ctypes.data)
    # the rect light of the same test: area 0.5, radiance 2 / 0.5 = 4 per channel
    rect = rect_light((0.0, 2.0, 0.0), (0.5, 0.0, 0.0), (0.0, 0.0, 0.25), (1.0, 1.0, 1.0), 2.0)
    assert lib.or_light_valid(rect.ctypes.data)
    packed = _pack(rect)
    assert list(packed[1, :3]) == [4.0, 4.0, 4.0]
    skew = rect.copy()
    skew["half_v"] = (0.5, 0.0, 0.25)
    assert not lib.or_light_valid(skew.ctypes.data)


def test_area_lights_pack_and_validate():
    """lights.rs:101-152, 207-230 for rect and disc lights: kind in a.w,
    power over area in b.xyz with b.w = f32::MAX, half_u or the unit normal
    with the two-sidedness in c, half_v or the radius with the shadow flag
    in d; the validity rules of both shapes; product rows == oracle rows."""
    lib = O.lib()
    rect = rect_light((1.0, 2.0, 3.0), (0.3, 0.0, 0.0), (0.0, 0.0, 0.2), (0.8, 0.9, 1.0), 2.0, two_sided=True)
    packed = _pack(rect)
    area = np.float32(4.0) * np.float32(0.3 * 0.2)
    assert packed[0, 3] == 2.0 and list(packed[0, :3]) == [1.0, 2.0, 3.0]
    np.testing.assert_allclose(packed[1, :3], np.float32([0.8, 0.9, 1.0]) * np.float32(2.0) / area, rtol=1e-6)
    assert packed[1, 3] == np.finfo(np.float32).max
    assert list(packed[2]) == [np.float32(0.3), 0.0, 0.0, 1.0] and list(packed[3]) == [0.0, 0.0, np.float32(0.2), 1.0]
    disc = disc_light((0.0, 1.0, 0.0), (0.0, -2.0, 0.0), 0.25, (1.0, 1.0, 1.0), np.pi * 0.0625, shadow=False)
    packed = _pack(disc)
    assert packed[0, 3] == 3.0
    np.testing.assert_allclose(packed[1, :3], [1.0, 1.0, 1.0], rtol=1e-6)  # power pi r^2 over area pi r^2
    assert list(packed[2]) == [0.0, -1.0, 0.0, 0.0] and list(packed[3]) == [0.25, 0.0, 0.0, 0.0]
    for bad in (dict(disc_radius=0.0), dict(direction=(0.0, 0.0, 0.0)), dict(disc_radius=np.nan)):
        d = disc.copy()
        for k, v in bad.items():
            d[k] = v
        assert not lib.or_light_valid(d.ctypes.data)
    flat = rect.copy()
    flat["half_u"] = (0.0, 0.0, 0.0)
    assert not lib.or_light_valid(flat.ctypes.data)
    # area lights never go to the tree: with five point lights and two
    # area lights the rows are [area lights][treed in tree order][nodes]
    lights = np.concatenate([point_light((k, 3.0, 0.0), (1, 1, 1), 5.0, 20.0) for k in range(3)] + [rect, disc] +
                            [point_light((k, 3.0, 2.0), (1, 1, 1), 5.0, 20.0) for k in range(2)])
    rows = np.zeros((3 * len(lights), 16), np.float32)
    looped, tree = C.c_uint32(), C.c_uint32()
    n = lib.or_lights(lights.ctypes.data, len(lights), rows.ctypes.data, C.byref(looped), C.byref(tree))
    assert looped.value == 2 and tree.value == 5
    assert rows[0, 3] == 2.0 and rows[1, 3] == 3.0
    prod = np.zeros_like(rows)
    pn, pl, pt = C.c_uint32(), C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(256)
    assert N.lib().pfxh_layout_lights(lights.ctypes.data, len(lights), prod.ctypes.data, C.byref(pn), C.byref(pl),
                                      C.byref(pt), err, 256) == 0, err.value
    assert pn.value == n and pl.value == 2 and pt.value == 5
    assert prod[:n].tobytes() == rows[:n].tobytes()


def street(count):
    """tree.rs:401-418."""
    rows = []
    for k in range(count):
        side = -6.0 if k % 2 == 0 else 6.0
        if k % 3 == 0:
            rows.append(spot_light((side, 5.0, k * 4.0), (1.0, 0.8, 0.6), 40.0, 18.0, (0.0, -1.0, 0.0), 30.0, 50.0))
        else:
            rows.append(point_light((side, 5.0, k * 4.0), (1.0, 0.8, 0.6), 40.0, 18.0))
    return np.concatenate(rows)


def lay(lights):
    rows = np.zeros((3 * len(lights) + 1, 16), np.float32)
    looped, tree = C.c_uint32(), C.c_uint32()
    n = O.lib().or_lights(lights.ctypes.data, len(lights), rows.ctypes.data, C.byref(looped), C.byref(tree))
    return rows[:n], looped.value, tree.value


def product_lay(lights):
    rows = np.zeros((3 * len(lights) + 1, 16), np.float32)
    count, looped, tree = C.c_uint32(), C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(256)
    st = N.lib().pfxh_layout_lights(lights.ctypes.data, len(lights), rows.ctypes.data, C.byref(count),
                                    C.byref(looped), C.byref(tree), err, 256)
    assert st == 0, err.value
    return rows[: count.value], looped.value, tree.value


@pytest.mark.parametrize("count", [5, 32, 128, 1024])
def test_every_light_sits_in_exactly_one_leaf(count):
    """tree.rs:435-448 (our rows index lights by their row, so coverage is
    checked through the packed positions)."""
    lights = street(count)
    rows, looped, tree = lay(lights)
    assert looped == 0 and tree == count
    nodes = rows[count:]
    assert len(nodes) < 2 * count
    seen = []

    def walk(at):
        d = rows[at, 12:16]
        n, link = int(d[1]), int(d[2])
        if n == 0:
            walk(at + 1)
            walk(link)
        else:
            assert n <= 8
            seen.extend(range(link, link + n))

    walk(count)
    assert sorted(seen) == list(range(count))
    # the treed rows are the lights, each exactly once
    pos = {tuple(np.round(p, 4)) for p in lights["position"]}
    assert {tuple(np.round(r[:3], 4)) for r in rows[:count]} == pos


def test_a_node_bounds_its_lights_energy_and_reach():
    """tree.rs:451-466."""
    lights = street(128)
    rows, _, _ = lay(lights)
    root = rows[128]
    energy = float(sum(sum(l["colour"]) * l["intensity"] / 3.0 for l in lights))
    assert abs(root[3] - energy) <= energy * 1e-5
    assert root[7] == 18.0
    for p in lights["position"]:
        assert np.all(root[:3] <= p) and np.all(p <= root[4:7])


def test_few_lights_keep_their_order_and_their_own_slots():
    """tree.rs:468-476."""
    rows, looped, tree = lay(street(4))
    assert (looped, tree) == (4, 0) and len(rows) == 4
    _, looped, tree = lay(street(5))
    assert (looped, tree) == (0, 5)


def scattered(count, seed, grid=None):
    """Point and spot lights at random positions, as the terrain scene
    scatters them; with `grid`, coordinates snap to a few values so many
    lights tie on the split axis."""
    rng = np.random.default_rng(seed)
    rows = []
    for k in range(count):
        pos = rng.uniform(-50.0, 50.0, 3)
        if grid:
            pos = np.round(pos / (100.0 / grid)) * (100.0 / grid)
        intensity = float(np.exp(rng.uniform(0.0, np.log(1000.0))))
        if k % 2 == 0:
            rows.append(point_light(tuple(pos), (1.0, 0.9, 0.8), intensity, 30.0))
        else:
            d = np.array([rng.uniform(-0.5, 0.5), -1.0, rng.uniform(-0.5, 0.5)])
            outer = float(rng.uniform(15.0, 60.0))
            rows.append(spot_light(tuple(pos), (1.0, 0.9, 0.8), intensity, 30.0, tuple(d / np.linalg.norm(d)),
                                   0.8 * outer, outer))
    return np.concatenate(rows)


def test_product_light_tree_is_byte_identical():
    """The split order (Rust's select_nth_unstable_by, restated on both
    sides) included: random positions exercise its partition rounds,
    snapped ones its runs of equal keys."""
    cases = [street(c) for c in (3, 5, 128)]
    cases += [scattered(c, 30 + c) for c in (9, 17, 40, 128, 1000)]
    cases += [scattered(c, 40 + c, grid=3) for c in (64, 128, 700)]
    for lights in cases:
        a, b = product_lay(lights), lay(lights)
        assert a[1:] == b[1:] and a[0].tobytes() == b[0].tobytes()


def test_adaptive_rounds_known_answers():
    """adaptive.rs:391-407."""
    a = N.AdaptiveC(0.01, 16, 100, 1.5)
    out = np.zeros(32, np.uint32)
    n = O.lib().or_rounds(C.byref(a), out.ctypes.data, 32)
    assert out[:n].tolist() == [8, 16, 24, 36, 54, 81, 100]
    a = N.AdaptiveC(0.01, 16, 64, 1.5)
    n = O.lib().or_rounds(C.byref(a), out.ctypes.data, 32)
    assert out[:n].tolist() == [8, 16, 24, 36, 54, 64]


def test_adaptive_error_known_answers():
    """adaptive.rs:447-459 (the half-buffer error at a half split)."""
    even = np.float32([0.5, 0.4, 0.3, 1.0])
    odd = np.float32([0.52, 0.38, 0.33, 1.0])
    all_ = ((even + odd) * np.float32(0.5)).astype(np.float32)
    f = O.lib().or_adaptive_error
    got = f(even.ctypes.data, 64.0, all_.ctypes.data, 128.0)
    diff = float(np.abs(even[:3] - odd[:3]).sum()) * 0.5
    want = diff / (1e-4 + np.sqrt(all_[:3].sum()))
    assert abs(got - want) < 1e-6
    assert f(even.ctypes.data, 0.0, all_.ctypes.data, 8.0) == np.float32(1e9)
    twos = np.full(4, 0.2, np.float32)
    assert f(twos.ctypes.data, 10.0, twos.ctypes.data, 30.0) == 0.0


def test_sky_cdf_sums_to_one():
    """sky.rs:506-524 (on an HDR lat-long sky) and the black-sky uniform case."""
    rng = np.random.default_rng(1)
    tex = rng.uniform(0.0, 4.0, (16, 32, 4)).astype(np.float32)
    rows = np.zeros(128, np.float32)
    cols = np.zeros(256 * 128, np.float32)
    O.lib().or_sky_cdf(tex.ctypes.data, 32, 16, 256, 128, rows.ctypes.data, cols.ctypes.data)
    mass = 0.0
    for r in range(128):
        rm = rows[r] - (rows[r - 1] if r else 0.0)
        cm = np.diff(np.concatenate([[0.0], cols[r * 256:(r + 1) * 256]]))
        assert rm >= -1e-5 and np.all(cm >= -1e-5)
        mass += rm * cm.sum()
    assert abs(mass - 1.0) < 1e-4
    assert abs(rows[-1] - 1.0) < 1e-6
    O.lib().or_sky_cdf(None, 0, 0, 256, 128, rows.ctypes.data, cols.ctypes.data)
    assert np.allclose(rows, (np.arange(128) + 1) / 128.0)
    assert np.allclose(cols[:256], (np.arange(256) + 1) / 256.0)


@pytest.mark.parametrize("seed,count", [(1, 40), (2, 200), (3, 1024)])
def test_leaf_order_does_not_depend_on_the_selection_pivots(seed, count):
    """tree.rs:263-293 recurses to one emitter per leaf, so with distinct
    centroids each split's selection property (pinned in test_select.py)
    already fixes which emitters lie on each side; the pivots Rust's
    select_nth_unstable_by picks only reorder the emitters inside a side,
    which the recursion sorts again. Building the tree with a stable sort as
    the selection must therefore give the same leaf order: the treed rows
    byte-identical. The node rows sum the same measures in another order,
    so they agree to rounding."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-50.0, 50.0, (count, 3)).astype(np.float32)
    col = rng.uniform(0.1, 1.0, (count, 3)).astype(np.float32)
    lights = np.concatenate([point_light(tuple(p), tuple(c), float(rng.uniform(1.0, 40.0)), 30.0)
                             for p, c in zip(pos, col)])
    rows, looped, tree = lay(lights)
    try:
        O.lib().or_set_tree_split(1)
        by_sort, looped2, tree2 = lay(lights)
    finally:
        O.lib().or_set_tree_split(0)
    assert (looped2, tree2) == (looped, tree) and tree == count and len(by_sort) == len(rows)
    leaves = slice(looped, looped + tree)
    assert rows[leaves].tobytes() == by_sort[leaves].tobytes()
    nodes = slice(looped + tree, None)
    np.testing.assert_allclose(by_sort[nodes], rows[nodes], rtol=1e-5, atol=1e-6)
    # and the default build is the product's, byte for byte
    assert product_lay(lights)[0].tobytes() == rows.tobytes()
