"""Batches of uneven stills against single traces: every still of a batch
(own scene, size, sample count and seed) reads back byte-identical colour,
albedo and normal images to its own Trace sampled with the same seed --
with one wave or many (units split across waves), with the captured graph or
direct launches, across repeated renders, and after the autotuner."""
import numpy as np
import pytest

from pfx_amd import Bounces, Detail, Trace, scenes
from pfx_amd.batch import Batch, Still, Tuning

pytestmark = pytest.mark.gpu

DETAIL = Detail(bounces=Bounces.deep(4))


def mix():
    return [
        Still(scenes.sphere_terrain(11, tris=20000, lights=3, width=96, height=64), 96, 64, 4, 7),
        Still(scenes.sphere_terrain(12, tris=60000, lights=17, width=128, height=128), 128, 128, 2, 99),
        Still(scenes.box_spheres(seed=3, spheres=8, width=64, height=64), 64, 64, 8, 5),
        Still(scenes.terrain(seed=4, grid=(64, 32), meshes=2, mesh_tris=(200, 800), instances=8, lights=12,
                             width=80, height=48), 80, 48, 3, 1234),
        Still(scenes.sphere_terrain(13, tris=5000, lights=1, width=32, height=200), 32, 200, 5, 3),
    ]


def single(still, detail, passes=1):
    with Trace.new_detailed(still.scene, detail, still.width, still.height) as t:
        for _ in range(passes):
            t.sample(still.spp, still.seed)
        return t.readback()


def same(a, b):
    for name in ("color", "albedo", "normal"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.tobytes() == y.tobytes(), (name, float(np.abs(x - y).max()))


@pytest.mark.parametrize("wave_paths,graph", [(None, True), (16384, True), (16384, False)])
def test_batch_equals_single_traces(wave_paths, graph):
    print(f"batch wave_paths={wave_paths} graph={graph}", flush=True)
    stills = mix()
    with Batch(stills, DETAIL) as b:
        t = b.tuning
        if wave_paths:
            t.wave_paths = wave_paths
        t.graph = graph
        b.tuning = t
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
        s = b.stats()
    assert s["paths"] == sum(x.width * x.height * x.spp for x in stills)
    assert s["closest_rays"] >= s["paths"]
    for still, out in zip(stills, outs):
        same(out, single(still, DETAIL))


def test_batch_renders_accumulate_like_repeated_samples():
    print("batch accumulate", flush=True)
    stills = mix()[:3]
    with Batch(stills, DETAIL) as b:
        t = b.tuning
        t.wave_paths = 20000
        b.tuning = t
        b.render()
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
    for still, out in zip(stills, outs):
        same(out, single(still, DETAIL, passes=2))


def test_autotuner_picks_a_valid_shape_and_results_hold():
    print("batch tune", flush=True)
    stills = mix()
    with Batch(stills, DETAIL) as b:
        t = b.tune()
        assert isinstance(t, Tuning)
        assert max(x.width * x.height for x in stills) <= t.wave_paths
        assert 1 <= t.refill <= 64 and 1 <= t.trace_blocks_per_cu <= 16
        assert all(b.trace(i).samples() == 0 for i in range(len(stills)))
        b.render()
        outs = [b.trace(i).readback() for i in range(len(stills))]
        with pytest.raises(ValueError):
            b.tuning = Tuning(wave_paths=16, refill=16, trace_blocks_per_cu=16, shade_blocks_per_cu=8)
    for still, out in zip(stills, outs):
        same(out, single(still, DETAIL))


def test_lent_trace_samples_through_the_batch_wave():
    print("batch lent trace", flush=True)
    stills = mix()[:2]
    with Batch(stills, DETAIL) as b:
        lent = b.trace(1)
        lent.sample(stills[1].spp, stills[1].seed)
        out = lent.readback()
        lent.close()  # does not free the batch's still
        again = b.trace(1).readback()
    same(out, single(stills[1], DETAIL))
    same(again, out)
