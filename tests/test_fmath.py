"""Pins of the shared single-precision math (pfx_amd/csrc/fmath.h).

The kernels and the oracle share these functions so they round identically;
here they are checked against numpy's libm (float64 reference rounded to
float32) over the argument ranges the tracer passes them.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

SIN, COS, ACOS, ATAN2, EXP2, LOG2, POW, ASIN, COS_ACOS_GAP = range(9)


def run(which, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    yp = None
    if y is not None:
        y = np.ascontiguousarray(y, np.float32)
        yp = y.ctypes.data
    O.lib().or_fm_many(which, x.ctypes.data, yp, out.ctypes.data, len(x))
    return out


def ulp_err(got, want64):
    want = want64.astype(np.float32)
    spacing = np.spacing(np.abs(want).astype(np.float32))
    spacing = np.where(spacing == 0, np.float32(1.4e-45), spacing)
    return np.abs(got.astype(np.float64) - want64) / spacing


rng = np.random.default_rng(5)


@pytest.mark.parametrize(
    "which,fn,lo,hi,bound",
    [
        (SIN, np.sin, -2 * np.pi, 4 * np.pi, 2.0),
        (COS, np.cos, -2 * np.pi, 4 * np.pi, 2.0),
        (ACOS, np.arccos, -1.0, 1.0, 3.0),
        (ASIN, np.arcsin, -1.0, 1.0, 3.0),
        (EXP2, np.exp2, -60.0, 60.0, 2.0),
        (LOG2, np.log2, 1e-30, 1e30, 3.0),
    ],
)
def test_unary_within_ulps(which, fn, lo, hi, bound):
    if which == LOG2:
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), 20000)).astype(np.float32)
    else:
        x = rng.uniform(lo, hi, 20000).astype(np.float32)
    got = run(which, x)
    want = fn(x.astype(np.float64))
    if which in (SIN, COS):
        # absolute error near zeros of sin/cos is what matters
        assert np.max(np.abs(got - want)) < 4e-7
    else:
        assert np.percentile(ulp_err(got, want), 100) <= bound


def test_atan2_all_quadrants():
    y = rng.uniform(-10, 10, 20000).astype(np.float32)
    x = rng.uniform(-10, 10, 20000).astype(np.float32)
    got = run(ATAN2, y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - want)) < 5e-7
    assert run(ATAN2, np.float32([1.0]), np.float32([0.0]))[0] == np.float32(np.pi / 2)


def test_cos_of_an_acos_gap():
    """cos(max(0, acos a - acos b)) through its identity, against numpy in
    float64 over the light tree's arguments (a a clamped dot product, b the
    cosine of a light's half angle, in (0, 1]), near the fold a = b and at the
    ends of the range."""
    a = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(0.9, 1.0, 5000), [-1.0, 1.0, 0.0]])
    b = np.concatenate([rng.uniform(0, 1, 20000), rng.uniform(0.9, 1.0, 5000), [1.0, 1.0, 0.5]])
    near = rng.uniform(0.0, 1.0, 2000)
    a = np.concatenate([a, near * (1.0 - 1e-6)]).astype(np.float32)
    b = np.concatenate([b, near]).astype(np.float32)
    got = run(COS_ACOS_GAP, a, b)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    want = np.cos(np.maximum(0.0, np.arccos(a64) - np.arccos(b64)))
    assert np.max(np.abs(got - want)) < 2.5e-7
    assert np.all(got <= 1.0) and np.all(got[a >= b] == 1.0)


def test_pow_shading_range():
    x = rng.uniform(0.0, 1.0, 20000).astype(np.float32)
    for e in (4.0, 5.0):
        got = run(POW, x, np.full_like(x, e))
        want = x.astype(np.float64) ** e
        # exp2(y log2 x) in float32, as GPU pow is built: ~|y log2 x| ulp relative
        assert np.max(np.abs(got - want) / np.maximum(want, 1e-30)) < 1e-5
    assert run(POW, np.float32([0.0]), np.float32([5.0]))[0] == 0.0
    assert run(POW, np.float32([1.0]), np.float32([5.0]))[0] == 1.0


def test_pcg_is_the_rxs_m_xs_hash():
    """PCG-RXS-M-XS 32 (O'Neill 2014, the output permutation the reference's
    trace.wgsl:168-172 applies to one LCG step): checked against a bignum
    evaluation on random and edge inputs."""

    def pcg(v):
        s = (v * 747796405 + 2891336453) % 2**32
        w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) % 2**32
        return (w >> 22) ^ w

    vals = [0, 1, 2**31, 2**32 - 1] + list(rng.integers(0, 2**32, 2000, dtype=np.uint64))
    for v in vals:
        assert O.lib().or_pcg(int(v)) == pcg(int(v))


def test_random_stream_is_uniform():
    # rnd = (pcg(state) >> 8) / 2^24: mean, variance and a 64-bin chi-square
    state = 12345
    out = np.empty(200000)
    lib = O.lib()
    for i in range(len(out)):
        state = lib.or_pcg(state)
        out[i] = (state >> 8) / 16777216.0
    assert abs(out.mean() - 0.5) < 0.003
    assert abs(out.var() - 1 / 12) < 0.002
    hist = np.histogram(out, 64, (0, 1))[0]
    chi2 = ((hist - len(out) / 64) ** 2 / (len(out) / 64)).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5


def test_offset_ray_known_answers():
    """crates/trace/src/math.rs:80-92."""
    lib = O.lib()
    assert lib.or_offset_axis(0.0, 0.0) == 0.0
    assert lib.or_offset_axis(0.0, 1.0) == np.float32(1.0 / 65536.0)
    assert lib.or_offset_axis(1.0, 1.0) == np.float32(1.0 + 1.0 / 32768.0)
    assert lib.or_offset_axis(-1.0, -1.0) == np.float32(-1.0 - 1.0 / 32768.0)


def test_spawn_point_half_millimetre_gap():
    """crates/trace/src/math.rs:94-105."""
    out = np.zeros(3, np.float32)
    p = np.float32([0, 0, 0])
    d = np.float32([0, 1, 0])
    O.lib().or_spawn_point(p.ctypes.data, d.ctypes.data, np.float32([0, 2, 0]).ctypes.data, 2.0, out.ctypes.data)
    assert 0.0 < out[1] < 0.0005 * 0.5
    p = np.float32([0, -0.05, 0])
    O.lib().or_spawn_point(p.ctypes.data, d.ctypes.data, np.float32([0, 2600, 0]).ctypes.data, 2600.0,
                           out.ctypes.data)
    assert out[1] + np.float32(0.05) > 0.001


def test_pow_of_a_rounding_negative_base_is_finite():
    """1 - cos with cos a hair above 1 (a normalised vector's z rounding past
    1) is a tiny negative base: pow takes it as 0 instead of NaN, for the
    integer and non-integer exponents shading uses."""
    for y in (4.0, 5.0, 2.5):
        x = np.float32([-1.1920929e-07, -0.0, 0.0])
        assert np.array_equal(run(POW, x, np.full_like(x, y)), np.zeros(3, np.float32))
