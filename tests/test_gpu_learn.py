"""The learned denoiser's HIP kernels against the FP64 oracle: loss and
gradient of every layer (FP32 operands: 1e-4 of each layer's gradient norm,
measured 4.7e-5 and 6.9e-5, where an FP32 pre-activation within rounding of
zero flips its ReLU against the FP64 oracle at a few pixels (2.6e-7 in layer
0 with the noise channel zeroed); BF16 operands, the training mode: 0.12 per
layer and 0.03 over the whole gradient, measured 0.075 and 0.014, the deep layers
of a 32-pixel crop seeing only 16 to 64 pixels, where single ReLU flips from
rounding show; the 256-pixel crops of test_gpu_learn_c4 hold 0.05 and
0.02, measured 0.029 and 0.017), the batch gather under all eight
symmetries of the square crop, inference on a frame that is not a multiple of
16, the first Adam step in closed form (p - lr * g / (|g| + eps)), and
training that lowers the loss."""
import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import learn as Lm

pytestmark = pytest.mark.gpu


def crops(rng, b, h, w):
    color = rng.exponential(0.5, (b, h, w, 4)).astype(np.float32)
    albedo = rng.uniform(0, 1, (b, h, w, 4)).astype(np.float32)
    normal = rng.uniform(-1, 1, (b, h, w, 4)).astype(np.float32)
    depth = rng.uniform(0, 20, (b, h, w)).astype(np.float32)
    target = rng.exponential(0.5, (b, h, w, 4)).astype(np.float32)
    deviation = rng.exponential(0.1, (b, h, w)).astype(np.float32)
    return color, albedo, normal, depth, target, deviation


def symmetry(a, sym):
    """The crop as the gather kernel reads it under `sym` (bit 0 mirrors x,
    bit 1 mirrors y, bit 2 transposes)."""
    side = a.shape[0]
    oy, ox = np.indices((side, side))
    sx = side - 1 - ox if sym & 1 else ox
    sy = side - 1 - oy if sym & 2 else oy
    if sym & 4:
        sx, sy = sy, sx
    return a[sy, sx]


def params_for_test(seed):
    rng = np.random.default_rng(seed)
    p = Lm.init_params(seed)
    p += rng.normal(0, 0.01, len(p)).astype(np.float32) * (np.abs(p) == 0)  # nonzero biases
    return p


def layer_slices():
    out, off = [], 0
    for l in O.learn_layers():
        n = l[3] * l[2] * 9 + l[3]
        out.append((off, off + n))
        off += n
    return out


def filled(learner, data, side):
    learner.reserve(side, len(data[0]))
    for i in range(len(data[0])):
        learner.add(*(a[i] for a in data))
    return learner


@pytest.mark.parametrize("fp32,tol,total,loss", [(True, 1e-4, 1e-4, "l1_log"), (False, 0.12, 0.03, "l1_log"),
                                                  (True, 1e-4, 1e-4, "rel_l2")])
def test_gradient_matches_fp64_oracle(fp32, tol, total, loss):
    print(f"learn gradient fp32={fp32} loss={loss}", flush=True)
    rng = np.random.default_rng(21)
    data = crops(rng, 2, 32, 32)
    p = params_for_test(21)
    picks = [0 << 3 | 0, 1 << 3 | 5]
    with filled(Lm.Learner(fp32=fp32, params=p), data, 32) as L:
        L.set_loss(loss)
        loss_value = L.gradient(picks)
        g = L.grads()
    ref_data = [np.stack([a[0], symmetry(a[1], 5)]) for a in data]
    ref_loss, ref_g = O.learn_grad(p.astype(np.float64), *ref_data, loss=loss)
    loss = loss_value
    assert loss == pytest.approx(ref_loss, rel=1e-5 if fp32 else 1e-2)
    worst = 0.0
    for k, (lo, hi) in enumerate(layer_slices()):
        gap = np.linalg.norm(g[lo:hi] - ref_g[lo:hi])
        size = np.linalg.norm(ref_g[lo:hi])
        worst = max(worst, gap / size)
        assert gap <= tol * size, (k, gap, size)
    whole = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    print(f"  worst layer {worst:.4g}, whole gradient {whole:.4g}", flush=True)
    assert whole <= total


def test_gather_symmetries_match_the_oracle():
    print("learn gather symmetries", flush=True)
    rng = np.random.default_rng(22)
    data = crops(rng, 1, 32, 32)
    p = params_for_test(22)
    with filled(Lm.Learner(fp32=True, params=p), data, 32) as L:
        for sym in range(8):
            loss = L.gradient([sym])
            ref = [symmetry(a[0], sym)[None] for a in data]
            ref_loss, _ = O.learn_grad(p.astype(np.float64), *ref, want_grad=False)
            assert loss == pytest.approx(ref_loss, rel=1e-5), sym


@pytest.mark.parametrize("fp32,tol", [(True, 1e-4), (False, 5e-2)])
def test_inference_matches_fp64_oracle(fp32, tol):
    print(f"learn inference fp32={fp32}", flush=True)
    rng = np.random.default_rng(23)
    h, w = 24, 40
    color, albedo, normal, depth, _, deviation = (a[0] for a in crops(rng, 1, h, w))
    p = params_for_test(23)
    with Lm.Learner(fp32=fp32, params=p) as L:
        out = L.denoise(color, albedo, normal, depth, deviation)
    ref = O.learn_forward(p.astype(np.float64), color, albedo, normal, depth, deviation)
    gap = np.abs(np.log1p(out[..., :3]) - np.log1p(ref[..., :3]))
    assert float(gap.max()) <= tol * max(1.0, float(np.log1p(ref[..., :3]).max()))
    np.testing.assert_array_equal(out[..., 3], color[..., 3])


def test_mirror_passes_average_the_mirrored_denoises():
    """pfxh_learner_set_passes: 2 and 4 passes give the mean, in linear
    radiance, of one-pass denoises of the frame mirrored in x, y and both,
    each mirrored back (the batch runs every frame with a one-pass frame's
    arithmetic; the sum is taken in frame order, then scaled). The frame is
    not a multiple of 16, so the padding lies outside the mirrored image."""
    print("learn mirror passes", flush=True)
    rng = np.random.default_rng(29)
    h, w = 40, 56
    color, albedo, normal, depth, _, deviation = (a[0] for a in crops(rng, 1, h, w))

    def mirror(a, b):
        a = a[:, ::-1] if b & 1 else a
        return np.ascontiguousarray(a[::-1] if b & 2 else a)

    with Lm.Learner(params=params_for_test(29)) as L:
        singles = [mirror(L.denoise(*(mirror(x, b) for x in (color, albedo, normal, depth, deviation))), b)
                   for b in range(4)]
        for passes in (2, 4):
            L.set_passes(passes)
            got = L.denoise(color, albedo, normal, depth, deviation)
            want = singles[0][..., :3].copy()
            for k in range(1, passes):
                want += singles[k][..., :3]
            want *= np.float32(1.0 / passes)
            np.testing.assert_allclose(got[..., :3], want, rtol=1e-6, atol=1e-7)
            np.testing.assert_array_equal(got[..., 3], color[..., 3])
        L.set_passes(1)
        np.testing.assert_array_equal(L.denoise(color, albedo, normal, depth, deviation), singles[0])
        with pytest.raises(ValueError, match="passes must be 1, 2 or 4"):
            L.set_passes(3)


def splitmix_picks(seed, count, crops_n):
    mask = (1 << 64) - 1
    x, out = seed, []
    for _ in range(count):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        out.append(((z >> 3) % crops_n) << 3 | (z & 7))
    return out


def test_first_adam_step_is_the_closed_form():
    print("learn adam step", flush=True)
    rng = np.random.default_rng(24)
    data = crops(rng, 3, 32, 32)
    p0 = params_for_test(24)
    lr, seed, batch = 1e-3, 99, 2
    with filled(Lm.Learner(fp32=True, params=p0), data, 32) as L:
        L.gradient(splitmix_picks(seed, batch, 3))
        g = L.grads()
        losses = L.train(1, batch=batch, lr=lr, seed=seed)
        p1 = L.params()
    assert np.isfinite(losses).all()
    want = p0 - lr * g / (np.abs(g) + 1e-8)
    live = np.abs(g) > 1e-5
    assert live.sum() > 1000
    np.testing.assert_allclose(p1[live], want[live], rtol=0, atol=2e-3 * lr)
    np.testing.assert_array_equal(p1[g == 0], p0[g == 0])


def test_training_lowers_the_loss():
    print("learn training", flush=True)
    rng = np.random.default_rng(25)
    side, n = 32, 8
    yy, xx = np.indices((side, side)) / side
    targets, noisy = [], []
    for i in range(n):
        base = np.stack([np.sin(3 * xx + i) + 1.2, np.cos(2 * yy + i) + 1.2, xx * yy + 0.2, np.ones_like(xx)], -1)
        targets.append(base.astype(np.float32))
        noisy.append((base * rng.exponential(1.0, base.shape)).astype(np.float32))
    albedo = np.full((side, side, 4), 0.5, np.float32)
    normal = np.zeros((side, side, 4), np.float32)
    with Lm.Learner(params=Lm.init_params(25)) as L:
        L.reserve(side, n)
        for i in range(n):
            L.add(noisy[i], albedo, normal, None, targets[i])
        losses = L.train(150, batch=4, lr=1e-3, seed=7)
    assert np.isfinite(losses).all()
    first, last = float(losses[:10].mean()), float(losses[-10:].mean())
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.5 * first


def test_checked_denoise_falls_back_per_tile():
    """Tiles whose removed variance exceeds the threshold take the repo
    denoiser's output; the others keep the learned output, byte for byte."""
    print("learn checked fallback", flush=True)
    from pfx_amd import Bounces, Detail, Trace, scenes
    from pfx_amd import denoise as D

    scene = scenes.terrain(seed=4, grid=(64, 32), meshes=2, mesh_tris=(200, 800), instances=8, lights=12,
                           width=120, height=72)
    with Trace.new_detailed(scene, Detail(bounces=Bounces.deep(4)), 120, 72) as t, \
            Lm.Learner(weights=Lm.WEIGHTS) as L, D.Denoiser() as d:
        t.sample(4, 3)
        noisy = t.readback_color()
        learned = t.learned_denoise(L)
        repo = t.denoise(d)
        keep, n_keep = t.learned_denoise_checked(L, d, float("inf"), 32)
        swap, n_swap = t.learned_denoise_checked(L, d, -1.0, 32)
        # a threshold between the tiles' values
        r = []
        for ty in range(0, 72, 32):
            for tx in range(0, 120, 32):
                a = np.log1p(np.maximum(noisy[ty:ty + 32, tx:tx + 32, :3], 0))
                b = np.log1p(np.maximum(learned[ty:ty + 32, tx:tx + 32, :3], 0))
                r.append(float(np.mean((a - b) ** 2)))
        thr = float(np.median(r))
        mixed, n_mixed = t.learned_denoise_checked(L, d, thr, 32)
    assert n_keep == 0 and keep.tobytes() == learned.tobytes()
    assert n_swap == len(r) and swap.tobytes() == repo.tobytes()
    k = 0
    for ty in range(0, 72, 32):
        for tx in range(0, 120, 32):
            src = repo if r[k] > thr * (1 + 1e-4) else (learned if r[k] < thr * (1 - 1e-4) else None)
            if src is not None:
                assert mixed[ty:ty + 32, tx:tx + 32].tobytes() == src[ty:ty + 32, tx:tx + 32].tobytes(), k
            k += 1
    assert 0 < n_mixed < len(r)


def test_learned_denoise_follows_the_view():
    """The depth guide is kept between denoiser calls and refreshed when the
    view changes: after set_camera the result equals a fresh trace's."""
    print("learn view change", flush=True)
    from pfx_amd import Bounces, Camera, Detail, Trace, scenes

    scene = scenes.terrain(seed=4, grid=(64, 32), meshes=2, mesh_tris=(200, 800), instances=8, lights=12,
                           width=96, height=64)
    other = Camera.look((10.0, 25.0, -30.0), (0.0, 0.0, 5.0), 50.0, 96 / 64)
    detail = Detail(bounces=Bounces.deep(4))
    with Lm.Learner(weights=Lm.WEIGHTS) as L:
        with Trace.new_detailed(scene, detail, 96, 64) as t:
            t.sample(4, 5)
            first = t.learned_denoise(L)
            again = t.learned_denoise(L)
            t.set_camera(other)
            t.sample(4, 5)
            moved = t.learned_denoise(L)
        with Trace.new_detailed(scene, detail, 96, 64) as fresh:
            fresh.set_camera(other)
            fresh.sample(4, 5)
            want = fresh.learned_denoise(L)
    assert first.tobytes() == again.tobytes()
    assert moved.tobytes() == want.tobytes()
    assert moved.tobytes() != first.tobytes()
