"""The learned denoiser's training path at its own shapes: `train.generate()`
renders seeded terrain frames on the device (4-spp inputs with albedo,
normal and probe depth; targets from an independent sample sequence), cuts
256x256 crops into the device-resident training set, and a batch-16 BF16
gradient is taken over them. Checked:

* the batch gradient is the mean of the sixteen single-crop gradients (the
  loss is a mean over the batch);
* two of the crops at their full 256x256 size, under two of the eight
  symmetries, against the FP64 oracle (loss, every layer's gradient);
* a few Adam steps on the tracer-made set lower the loss.

The targets use 64 spp instead of the training runs' 1,024, to keep the
test short; the shapes, the data path and the kernels are the same.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import learn as Lm
from pfx_amd import train
from test_gpu_learn import layer_slices, params_for_test, symmetry

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_tracer_made_crops_batch16_gradient_against_the_fp64_oracle():
    p = params_for_test(31)
    with Lm.Learner(params=p) as learner:
        recorded = []
        add = learner.add

        def keep(color, albedo, normal, depth, target, deviation=None):
            recorded.append([np.array(a, np.float32) for a in (color, albedo, normal, depth, target, deviation)])
            add(color, albedo, normal, depth, target, deviation)

        learner.add = keep
        print("generating 16 crops of 256^2 from terrain renders", flush=True)
        n = train.generate(learner, crops=16, side=256, input_spp=4, target_spp=64, seed=3, frame=(512, 256),
                           views_per_scene=8)
        assert n == 16 and learner.crops == 16 and len(recorded) == 16
        for c, a, nr, d, t, v in recorded:
            assert c.shape == (256, 256, 4) and d.shape == (256, 256) and v.shape == (256, 256)
            assert np.all(np.isfinite(c)) and np.all(np.isfinite(t)) and np.all(v >= 0.0)
        picks = [k << 3 | (k % 8) for k in range(16)]
        loss16 = learner.gradient(picks)
        g16 = learner.grads()
        singles, losses = [], []
        for pk in picks:
            losses.append(learner.gradient([pk]))
            singles.append(learner.grads())
        mean = np.mean(singles, axis=0)
        print(f"  batch loss {loss16:.6f}, mean of singles {np.mean(losses):.6f}", flush=True)
        assert loss16 == pytest.approx(float(np.mean(losses)), rel=1e-4)
        assert np.linalg.norm(g16 - mean) <= 1e-3 * np.linalg.norm(mean)
        for k in (3, 10):
            sym = picks[k] & 7
            data = [symmetry(a, sym)[None] for a in recorded[k]]
            print(f"  FP64 oracle on crop {k} (symmetry {sym})", flush=True)
            ref_loss, ref_g = O.learn_grad(p.astype(np.float64), *data)
            g = singles[k]
            assert losses[k] == pytest.approx(ref_loss, rel=1e-2)
            worst = 0.0
            for j, (lo, hi) in enumerate(layer_slices()):
                gap = np.linalg.norm(g[lo:hi] - ref_g[lo:hi]) / max(np.linalg.norm(ref_g[lo:hi]), 1e-30)
                worst = max(worst, gap)
                assert gap <= 0.05, (j, gap)
            total = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
            print(f"    worst layer {worst:.4f}, whole gradient {total:.4f}", flush=True)
            assert total <= 0.02
        first = learner.train(1, batch=16, lr=5e-4, seed=7)
        later = learner.train(60, batch=16, lr=5e-4, seed=8)
        print(f"  loss {float(first[0]):.4f} -> {float(np.mean(later[-10:])):.4f}", flush=True)
        assert float(np.mean(later[-10:])) < float(first[0])
