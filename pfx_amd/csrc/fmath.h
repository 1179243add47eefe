// Deterministic single-precision math shared by the HIP kernels and the CPU
// checker. Every function here is written with plain IEEE +, -, *, / and
// sqrtf only (no fused multiply-add: all code is compiled with
// -ffp-contract=off), so a CPU build and a gfx950 build give the same bits.
//
// The polynomials and range reductions follow S. L. Moshier's Cephes single
// precision library (sinf/cosf, asinf/acosf, atanf, exp2f, log2f), a
// published method; accuracy is pinned against numpy's libm in
// tests/test_fmath.py (a few ulp over the ranges the tracer uses).
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define PFX_HD __host__ __device__ inline
#else
#define PFX_HD inline
#endif

namespace pfx {

constexpr float PI = 3.141592653589793f;
constexpr float TAU = 6.283185307179586f;

PFX_HD uint32_t f2u(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}
PFX_HD float u2f(uint32_t x) {
    union { float f; uint32_t u; } v;
    v.u = x;
    return v.f;
}

PFX_HD float fm_sin_poly(float r) {
    float z = r * r;
    return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
}
PFX_HD float fm_cos_poly(float r) {
    float z = r * r;
    return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
           - 0.5f * z + 1.0f;
}

// Reduce x by multiples of pi/2 (three-part Cody-Waite split) and return the
// quadrant in *q.
PFX_HD float fm_reduce(float x, int* q) {
    float j = floorf(x * 0.63661977236758134f + 0.5f);
    *q = (int)j;
    return ((x - j * 1.5703125f) - j * 4.837512969970703125e-4f) - j * 7.54978995489188216e-8f;
}

PFX_HD float fm_sin(float x) {
    int q;
    float r = fm_reduce(x, &q);
    float s = fm_sin_poly(r), c = fm_cos_poly(r);
    switch (q & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}
PFX_HD float fm_cos(float x) {
    int q;
    float r = fm_reduce(x, &q);
    float s = fm_sin_poly(r), c = fm_cos_poly(r);
    switch (q & 3) {
        case 0: return c;
        case 1: return -s;
        case 2: return -c;
        default: return s;
    }
}

// asin on |x| <= 0.5
PFX_HD float fm_asin_core(float x) {
    float z = x * x;
    return ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z
            + 1.6666752422e-1f) * z * x + x;
}
PFX_HD float fm_acos(float x) {
    if (x < -0.5f) return PI - 2.0f * fm_asin_core(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * fm_asin_core(sqrtf(0.5f * (1.0f - x)));
    return 1.5707963267948966f - fm_asin_core(x);
}
// cos(max(0, acos(a) - acos(b))) for a, b in [-1, 1], without transcendental
// calls: acos falls as its argument grows, so the gap is 0 (and the cosine 1)
// when a >= b; otherwise cos(x - y) = cos x cos y + sin x sin y with the
// sines of angles in [0, pi] the non-negative square roots (1 - v^2 taken
// as (1 - v)(1 + v), whose factors are exact near v = +-1).
PFX_HD float fm_cos_acos_gap(float a, float b) {
    if (a >= b) return 1.0f;
    float sa = (1.0f - a) * (1.0f + a), sb = (1.0f - b) * (1.0f + b);
    float r = a * b + sqrtf(sa > 0.0f ? sa : 0.0f) * sqrtf(sb > 0.0f ? sb : 0.0f);
    return r < 1.0f ? r : 1.0f;
}
PFX_HD float fm_asin(float x) {
    if (x > 0.5f) return 1.5707963267948966f - 2.0f * fm_asin_core(sqrtf(0.5f * (1.0f - x)));
    if (x < -0.5f) return -(1.5707963267948966f - 2.0f * fm_asin_core(sqrtf(0.5f * (1.0f + x))));
    return fm_asin_core(x);
}

PFX_HD float fm_atan(float x) {
    float sign = 1.0f;
    if (x < 0.0f) {
        sign = -1.0f;
        x = -x;
    }
    float y;
    if (x > 2.414213562373095f) {
        y = 1.5707963267948966f;
        x = -1.0f / x;
    } else if (x > 0.4142135623730950f) {
        y = 0.7853981633974483f;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x
             + x);
    return sign * y;
}
PFX_HD float fm_atan2(float y, float x) {
    if (x == 0.0f) {
        if (y > 0.0f) return 1.5707963267948966f;
        if (y < 0.0f) return -1.5707963267948966f;
        return 0.0f;
    }
    float w = fm_atan(y / x);
    if (x < 0.0f) w = (y >= 0.0f) ? w + PI : w - PI;
    return w;
}

PFX_HD float fm_ldexp(float x, int e) {
    // e in [-252, 254]: two steps keep every intermediate scale normal
    int e1 = e / 2, e2 = e - e1;
    return x * u2f((uint32_t)(e1 + 127) << 23) * u2f((uint32_t)(e2 + 127) << 23);
}

PFX_HD float fm_exp2(float x) {
    if (!(x == x)) return x;
    if (x > 128.0f) return u2f(0x7f800000u);
    if (x < -150.0f) return 0.0f;
    float i = floorf(x + 0.5f);
    float f = x - i;
    float px = f * (((((1.535336188319500e-4f * f + 1.339887440266574e-3f) * f + 9.618437357674640e-3f) * f
                      + 5.550332471162809e-2f) * f + 2.402264791363012e-1f) * f + 6.931472028550421e-1f);
    return fm_ldexp(1.0f + px, (int)i);
}

PFX_HD float fm_log2(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? u2f(0xff800000u) : u2f(0x7fc00000u);
    if (x == u2f(0x7f800000u)) return x;
    int e = 0;
    uint32_t b = f2u(x);
    if ((b & 0x7f800000u) == 0) {  // subnormal: scale into the normal range
        x = x * 8388608.0f;
        b = f2u(x);
        e = -23;
    }
    e += (int)((b >> 23) & 0xffu) - 126;
    float m = u2f((b & 0x807fffffu) | 0x3f000000u);  // [0.5, 1)
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = m + m - 1.0f;
    } else {
        m = m - 1.0f;
    }
    float z = m * m;
    float y = ((((((((7.0376836292e-2f * m - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m - 1.2420140846e-1f) * m
                   + 1.4249322787e-1f) * m - 1.6668057665e-1f) * m + 2.0000714765e-1f) * m - 2.4999993993e-1f) * m
               + 3.3333331174e-1f) * m * z;
    y = y + -0.5f * z;
    float r = y * 0.44269504088896340736f;
    r = r + m * 0.44269504088896340736f;
    r = r + y;
    r = r + m;
    r = r + (float)e;
    return r;
}

// pow for the non-negative bases the shading code passes (x^y = 2^(y log2 x))
// x^y for x >= 0. WGSL's pow leaves a negative base undefined; shading
// reaches one only through rounding (1 - cos with cos a hair above 1), and
// such a base is taken as 0 so the result stays finite.
PFX_HD float fm_pow(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x < 0.0f) x = 0.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : u2f(0x7f800000u);
    if (x == 1.0f) return 1.0f;
    return fm_exp2(y * fm_log2(x));
}

// min / max with IEEE minNum/maxNum NaN handling and a fixed answer for
// equal operands (the first one), so +0 and -0 come out the same on every
// compiler. Rust's f32::min/max leave that case open.
PFX_HD float minf(float a, float b) {
    if (b != b) return a;
    if (a != a) return b;
    return b < a ? b : a;
}
PFX_HD float maxf(float a, float b) {
    if (b != b) return a;
    if (a != a) return b;
    return b > a ? b : a;
}
PFX_HD float fm_clamp(float x, float lo, float hi) { return minf(maxf(x, lo), hi); }

}  // namespace pfx
