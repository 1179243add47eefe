"""The learned denoiser on the host: the FP64 oracle's gradient is pinned by
central finite differences in every layer, its forward pass by networks
whose output is known in closed form (a pass-through of the colour along the
full-resolution skip, a one-pixel shift by a corner tap, a constant), and the
library's layer table and parameter count agree with the oracle's. The
reference has no trained denoiser, so there are no reference vectors: these
checks are what the mathematics fixes."""
import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import learn as Lm


def crops(rng, b, h, w):
    color = rng.exponential(0.5, (b, h, w, 4)).astype(np.float32)
    albedo = rng.uniform(0, 1, (b, h, w, 4)).astype(np.float32)
    normal = rng.uniform(-1, 1, (b, h, w, 4)).astype(np.float32)
    depth = rng.uniform(0, 20, (b, h, w)).astype(np.float32)
    target = rng.exponential(0.5, (b, h, w, 4)).astype(np.float32)
    deviation = rng.exponential(0.1, (b, h, w)).astype(np.float32)
    return color, albedo, normal, depth, target, deviation


def layer_slices():
    out, off = [], 0
    for l in O.learn_layers():
        ci, co = l[2], l[3]
        out.append((off, off + co * ci * 9, off + co * ci * 9 + co))
        off += co * ci * 9 + co
    return out


def test_library_and_oracle_define_the_same_network(native_lib):
    lib_layers = Lm.layers()
    orc = O.learn_layers()
    assert len(lib_layers) == len(orc) == 16
    for a, b in zip(lib_layers, orc):
        assert (a["in0"], a["in1"], a["in_channels"], a["out_channels"], a["level"], int(a["upsample"]),
                int(a["pool"]), int(a["relu"])) == b
    assert Lm.param_count() == O.learn_param_count() == 1239651
    # 5 levels, 32 -> 128 channels
    assert {l["level"] for l in lib_layers} == {0, 1, 2, 3, 4}
    assert max(l["out_channels"] for l in lib_layers) == 128 and lib_layers[0]["out_channels"] == 32


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = Lm.init_params(5).astype(np.float64)
    params += rng.normal(0, 0.01, len(params))  # nonzero biases too
    data = crops(rng, 1, 16, 16)
    loss, g = O.learn_grad(params, *data)
    assert loss > 0
    eps = 1e-6
    checked = 0
    for w0, w1, b1 in layer_slices():
        for lo, hi in ((w0, w1), (w1, b1)):
            idx = np.arange(lo, hi)
            live = idx[np.abs(g[idx]) > 1e-7]
            for i in rng.choice(live, min(3, len(live)), replace=False) if len(live) else []:
                p1, p2 = params.copy(), params.copy()
                p1[i] += eps
                p2[i] -= eps
                fd = (O.learn_grad(p1, *data, want_grad=False)[0] - O.learn_grad(p2, *data, want_grad=False)[0]) / (2 * eps)
                # central differences at eps 1e-6 carry about 1e-9 of rounding
                # (the loss sums 768 terms): relative 1e-5, at least 1e-8
                assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3), (i, fd, g[i])
                checked += 1
    assert checked >= 60


def identity_params(shift=(0, 0)):
    """Carry log(1 + colour) through the full-resolution skip into the
    output: dec_conv1a reads the input's colour channels with one tap,
    dec_conv1b and dec_conv0 copy them with the centre tap."""
    layers = O.learn_layers()
    p = np.zeros(O.learn_param_count())
    sl = layer_slices()

    def w(k, o, c, ky, kx, v):
        ci = layers[k][2]
        p[sl[k][0] + ((o * ci + c) * 3 + ky) * 3 + kx] = v

    c0 = layers[12][3]  # dec_conv1a's first source: dec_conv2b's 64 channels
    ky, kx = 1 - shift[0], 1 - shift[1]
    for ch in range(3):
        w(13, ch, c0 + ch, ky, kx, 1.0)
        w(14, ch, ch, 1, 1, 1.0)
        w(15, ch, ch, 1, 1, 1.0)
    return p


def test_oracle_forward_reduces_to_known_networks():
    rng = np.random.default_rng(4)
    h, w = 24, 40  # not multiples of 16: the frame is padded with zeros
    color, albedo, normal, depth, _, deviation = (a[0] for a in crops(rng, 1, h, w))
    out = O.learn_forward(identity_params(), color, albedo, normal, depth, deviation)
    np.testing.assert_allclose(out[..., :3], color[..., :3], rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(out[..., 3], color[..., 3])
    # tap (0, 0) reads the pixel above and to the left
    out = O.learn_forward(identity_params((1, 1)), color, albedo, normal, depth)
    np.testing.assert_allclose(out[1:, 1:, :3], color[:-1, :-1, :3], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(out[0, :, :3], 0.0, atol=1e-12)
    # only the last bias: a constant
    p = np.zeros(O.learn_param_count())
    p[layer_slices()[15][1]:] = [0.5, 1.0, 2.0]
    out = O.learn_forward(p, color, albedo, normal, depth)
    np.testing.assert_allclose(out[..., :3], np.broadcast_to(np.expm1([0.5, 1.0, 2.0]), (h, w, 3)), rtol=1e-6)


def test_oracle_loss_is_mean_l1_of_log_radiance():
    rng = np.random.default_rng(5)
    data = list(crops(rng, 2, 16, 16))
    p = np.zeros(O.learn_param_count())
    p[layer_slices()[15][1]:] = [0.25, 0.5, 0.75]
    loss, _ = O.learn_grad(p, *data, want_grad=False)
    want = np.mean(np.abs(np.array([0.25, 0.5, 0.75]) - np.log1p(data[4][..., :3].astype(np.float64))))
    assert loss == pytest.approx(want, rel=1e-12)


def test_oracle_relative_l2_loss_and_its_gradient_in_closed_form():
    """The relative-L2 loss (mode 'rel_l2'): with only the last layer's
    bias set, y is that bias at every pixel, x = expm1(y), and the loss is
    the mean of (x - t)^2 / (x^2 + 0.01); its gradient with the denominator
    held fixed is, for the last bias of channel k, the mean over pixels of
    2 (x - t) / (x^2 + 0.01) (x + 1) / 3 per channel. A wrong sign, a missing
    (x + 1) or a denominator taken from the target fails one of these; the
    other layers' backward pass is the one the finite differences pin."""
    rng = np.random.default_rng(6)
    data = list(crops(rng, 2, 16, 16))
    p = np.zeros(O.learn_param_count())
    bias = np.array([0.25, 1.5, -0.5])
    p[layer_slices()[15][1]:] = bias
    loss, g = O.learn_grad(p, *data, loss="rel_l2")
    t = np.maximum(data[4][..., :3].astype(np.float64), 0.0)
    x = np.expm1(bias)
    den = x * x + 0.01
    want = np.mean((x - t) ** 2 / den)
    assert loss == pytest.approx(want, rel=1e-12)
    want_g = np.mean(2.0 * (x - t) / den * (x + 1.0), axis=(0, 1, 2)) / 3.0
    np.testing.assert_allclose(g[layer_slices()[15][1]:], want_g, rtol=1e-10)
    # the L1 mode is untouched by the switch
    loss1, _ = O.learn_grad(p, *data, want_grad=False)
    assert loss1 == pytest.approx(np.mean(np.abs(bias - np.log1p(t))), rel=1e-12)


def test_weights_file_round_trip(tmp_path, native_lib):
    p = Lm.init_params(7)
    assert p.dtype == np.float32 and len(p) == Lm.param_count()
    path = tmp_path / "w.pfxl"
    Lm.save_params(path, p)
    np.testing.assert_array_equal(Lm.load_params(path), p)
    path.write_bytes(path.read_bytes()[:-4])
    with pytest.raises(ValueError):
        Lm.load_params(path)
