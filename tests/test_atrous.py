"""The tracer's own a-trous denoiser (crates/trace/src/denoise.rs:1-62) in
the oracle: the reference's test design (denoise.rs:72-97: a material edge
survives, grain shrinks, the result is deterministic) and an independent
float64 numpy restatement of the filter on small random images."""
import numpy as np
import pytest

from oracle import oracle as O


def edge_case():
    width, height = 32, 16
    color = np.zeros((height, width, 4), np.float32)
    albedo = np.zeros_like(color)
    normal = np.zeros_like(color)
    for y in range(height):
        for x in range(width):
            side = 0.2 if x < width // 2 else 0.8
            grain = 0.08 if (x + y) % 2 == 0 else -0.08
            color[y, x] = (side + grain, side + grain, side + grain, 1.0)
            albedo[y, x] = (side, side, side, 1.0)
            normal[y, x] = (0.0, 0.0, 1.0, 1.0)
    return color, albedo, normal


def test_reference_design_edge_kept_grain_reduced_deterministic():
    color, albedo, normal = edge_case()
    out = O.atrous(color, albedo, normal, 64)
    h, w = color.shape[:2]
    left, right = out[h // 2, w // 2 - 1, 0], out[h // 2, w // 2, 0]
    assert right - left > 0.5
    y, x = h // 2, w // 4
    assert abs(out[y, x, 0] - 0.2) < abs(color[y, x, 0] - 0.2)
    assert out.tobytes() == O.atrous(color, albedo, normal, 64).tobytes()
    np.testing.assert_array_equal(out[..., 3], color[..., 3])
    with pytest.raises(ValueError):
        O.atrous(color, albedo, normal, 0)


def numpy_filter(color, albedo, normal, samples):
    """denoise.rs:1-62 restated in float64 with numpy's exp."""
    src = color.astype(np.float64).copy()
    al, nm = albedo.astype(np.float64), normal.astype(np.float64)
    h, w = color.shape[:2]
    k = [0.375, 0.25, 0.0625]
    lum = lambda c: c[..., 0] * 0.2126 + c[..., 1] * 0.7152 + c[..., 2] * 0.0722
    for step in (1, 2, 4):
        dst = src.copy()
        for y in range(h):
            for x in range(w):
                luma = lum(src[y, x])
                tol = 2.8 / np.sqrt(samples) * (luma + 0.05)
                s, ws = np.zeros(3), 0.0
                for dy in range(-2, 3):
                    for dx in range(-2, 3):
                        xx, yy = x + dx * step, y + dy * step
                        if not (0 <= xx < w and 0 <= yy < h):
                            continue
                        bg = np.abs(al[y, x, :3] - al[yy, xx, :3]).sum()
                        ng = np.abs(nm[y, x, :3] - nm[yy, xx, :3]).sum()
                        lg = abs(lum(src[yy, xx]) - luma)
                        wt = k[abs(dx)] * k[abs(dy)] * np.exp(-bg * 8 - ng * 4) * np.exp(-lg / max(tol, 1e-5))
                        s += src[yy, xx, :3] * wt
                        ws += wt
                dst[y, x, :3] = s / max(ws, 1e-8)
        src = dst
    return src


@pytest.mark.parametrize("seed,spp", [(1, 4), (2, 64)])
def test_oracle_matches_a_float64_restatement(seed, spp):
    rng = np.random.default_rng(seed)
    h, w = 9, 13
    color = rng.exponential(0.5, (h, w, 4)).astype(np.float32)
    albedo = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    normal = rng.uniform(-1, 1, (h, w, 4)).astype(np.float32)
    # a shared edge so the guides matter
    albedo[:, w // 2:] *= 0.3
    ref = numpy_filter(color, albedo, normal, spp)
    out = O.atrous(color, albedo, normal, spp)
    np.testing.assert_allclose(out[..., :3], ref[..., :3], rtol=2e-5, atol=1e-7)
    np.testing.assert_array_equal(out[..., 3], color[..., 3])
