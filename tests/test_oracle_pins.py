"""Pins of the oracle that the reference's own tests and the mathematics fix:

* the two-level layout against the reference's flattened-triangle test
  (crates/trace/src/levels.rs:591-681), on the product's and the oracle's
  layouts;
* the light-tree pick (lights.wgsl:276-386): the pdf it returns for a light
  is the probability that it picks that light. Enumerated over a fine grid
  of the uniform, each light's share of the grid equals its pdf, and the
  picked lights' pdfs plus the share of uniforms that pick nothing sum to 1
  (a descent can end in a leaf none of whose lights reaches the point: the
  node bounds are conservative, and the reference then returns pdf 0); a
  chi-squared test on random uniforms agrees.
"""
import ctypes as C

import numpy as np
import pytest

import levels_case as L
from oracle import Oracle
from oracle import oracle as O
from pfx_amd import Bounces, Detail, scenes
from pfx_amd import _native as N


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


def _bvh_order(tris):
    flat = np.zeros(len(tris), N.TRIANGLE)
    flat["vertices"] = tris.reshape(len(tris), 9)
    nodes = np.zeros(2 * len(tris), N.BVH_NODE)
    order = np.zeros(len(tris), np.uint32)
    count = C.c_uint32()
    err = C.create_string_buffer(256)
    assert N.lib().pfxh_layout_bvh(flat.ctypes.data, len(flat), nodes.ctypes.data, C.byref(count), order.ctypes.data,
                                   err, 256) == 0
    return order


def test_two_levels_find_the_hits_of_the_flattened_triangles():
    """levels.rs:591-681: 3 bottom-level builds (the bit-equal copy of the
    shared mesh reuses its build), 130 slots, and 3,000 random rays walked
    through the two-level buffer hit what brute force over the placed
    triangles hits, within 1e-4 * max(distance, 1)."""
    scene, placed, origins, directions = L.case()
    nodes, order = _levels(scene, True)
    onodes, oorder = _levels(scene, False)
    assert nodes.tobytes() == onodes.tobytes() and np.array_equal(order, oorder)
    world = scene.world_details["vertices"].reshape(-1, 3, 3)
    mt = scene.mesh_triangles["vertices"].reshape(-1, 3, 3)
    meshes = [mt[m["first"]:m["first"] + m["count"]] for m in scene.meshes[:2]]
    slots, builds, nslots = L.slot_triangles(nodes, order, world, meshes, _bvh_order)
    assert builds == 3
    assert nslots == 30 + 60 + 40
    hits = 0
    for o, d in zip(origins, directions):
        brute = L.triangle_hits(placed, o, d).min()
        walked = L.walk(nodes, slots, o, d)
        if np.isfinite(brute):
            assert walked is not None, (o, d, brute)
            a = walked[1]
            assert abs(a - brute) <= 1e-4 * max(brute, 1.0), (a, brute)
            hits += 1
        else:
            assert walked is None, (o, d, walked)
    assert hits >= 50  # 74 of the 3,000 rays hit a triangle


def test_two_levels_a_distinct_copy_is_built_again():
    """The dedup is by geometry: moving one vertex of the copy makes it a
    fourth bottom-level build (levels.rs:284-303)."""
    scene, _, _, _ = L.case()
    scene.mesh_triangles = scene.mesh_triangles.copy()
    scene.mesh_triangles[100]["vertices"][0] += np.float32(0.25)
    nodes, order = _levels(scene, True)
    world = scene.world_details["vertices"].reshape(-1, 3, 3)
    mt = scene.mesh_triangles["vertices"].reshape(-1, 3, 3)
    groups = L._top_groups(nodes)
    assert len({int(g["first"]) for g in groups}) == 4
    assert len({int(g["count"]) for g in groups}) == 4


def _tree_scene(lights=128, seed=4):
    scene = scenes.terrain(seed=seed, grid=(32, 16), size=(24.0, 12.0), meshes=2, mesh_tris=(100, 200), instances=4,
                           lights=lights, width=16, height=16)
    return Oracle(scene, Detail(bounces=Bounces.deep(2)), 16, 16, threads=8)


POINTS = [((3.0, 1.0, 2.0), (0.0, 1.0, 0.0)), ((-6.0, 0.5, -3.0), (0.2, 0.9, 0.1)), ((8.0, 3.0, 4.0), (-0.6, 0.8, 0.0)),
          ((0.0, 12.0, 0.0), (0.0, -1.0, 0.0)), ((-10.0, 0.0, 5.0), (0.0, 0.0, 1.0))]


@pytest.mark.parametrize("k", range(len(POINTS)))
def test_tree_pick_pdf_is_the_pick_probability(k):
    """Enumeration over 2^20 evenly spaced uniforms: each light's pdf is the
    same wherever it is picked, each light's share of the grid is its pdf,
    and the pdfs of the picked lights plus the failed share sum to 1."""
    ref = _tree_scene()
    rows, looped, tree = ref.light_meta()
    assert tree == 128 and looped == 0
    p, n = POINTS[k]
    n = np.asarray(n, np.float32) / np.float32(np.linalg.norm(n))
    count = 1 << 20
    u = ((np.arange(count, dtype=np.float64) + 0.5) / count).astype(np.float32)
    picks, pdfs = ref.tree_picks(p, n, u)
    ok = pdfs > 0
    assert ok.mean() > 0.5  # the point sees lights
    picked = np.unique(picks[ok])
    assert picked.min() >= looped and picked.max() < looped + tree
    total = 0.0
    for light in picked:
        mine = pdfs[ok & (picks == light)]
        assert np.all(mine == mine[0]), light  # one pdf per light
        share = len(mine) / count
        assert abs(share - float(mine[0])) <= 2e-5 + 1e-3 * float(mine[0]), (light, share, float(mine[0]))
        total += float(mine[0])
    failed = 1.0 - ok.mean()
    assert abs(total + failed - 1.0) <= 1e-4, (total, failed)


def test_tree_pick_frequencies_chi_squared():
    """Random uniforms at one point: picked-light counts against the pdfs the
    picks return (chi-squared, bins with at least 5 expected)."""
    from scipy.stats import chi2

    ref = _tree_scene()
    p, n = POINTS[0]
    rng = np.random.default_rng(5)
    u = rng.random(400_000).astype(np.float32)
    picks, pdfs = ref.tree_picks(p, np.asarray(n, np.float32), u)
    lights = np.unique(picks[pdfs > 0])
    pdf_of = {int(l): float(pdfs[picks == l][0]) for l in lights}
    # one more bin: the uniforms that pick nothing, expected 1 - sum of pdfs
    probs = np.array([pdf_of[int(l)] for l in lights])
    expected = np.append(probs, max(0.0, 1.0 - probs.sum())) * len(u)
    observed = np.append([(picks[pdfs > 0] == l).sum() for l in lights], (pdfs <= 0).sum())
    keep = expected >= 5
    rest_e, rest_o = expected[~keep].sum(), observed[~keep].sum()
    e = np.append(expected[keep], rest_e) if rest_e > 0 else expected[keep]
    o = np.append(observed[keep], rest_o) if rest_e > 0 else observed[keep]
    stat = float(((o - e) ** 2 / e).sum())
    assert chi2.sf(stat, len(e) - 1) > 1e-3, (stat, len(e))
