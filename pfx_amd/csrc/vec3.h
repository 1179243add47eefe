// Three-component float vectors with a fixed evaluation order, shared by the
// HIP kernels and the CPU checker so both round identically. dot() sums
// left to right ((x + y) + z), the order WGSL's dot() is written in;
// normalize() multiplies by the IEEE reciprocal of the IEEE square root.
#pragma once
#include "fmath.h"

namespace pfx {

struct v3 {
    float x, y, z;
};

PFX_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
PFX_HD v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PFX_HD v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PFX_HD v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
PFX_HD v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
PFX_HD v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PFX_HD v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
PFX_HD v3 operator/(v3 a, v3 b) { return v3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PFX_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PFX_HD v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PFX_HD float length(v3 a) { return sqrtf(dot(a, a)); }
// WGSL normalize: v / |v| (a zero vector gives non-finite values, as WGSL does)
PFX_HD v3 normalize(v3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
PFX_HD float maxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
PFX_HD v3 vabs(v3 a) { return v3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
PFX_HD v3 vmin(v3 a, v3 b) { return v3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
PFX_HD v3 vmax(v3 a, v3 b) { return v3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
PFX_HD v3 vclamp(v3 a, float lo, float hi) {
    return v3{fm_clamp(a.x, lo, hi), fm_clamp(a.y, lo, hi), fm_clamp(a.z, lo, hi)};
}
PFX_HD bool all_zero(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
PFX_HD float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Column-major 3x3 (WGSL mat3x3f): columns t, b, n.
struct m3 {
    v3 c0, c1, c2;
};
// M * v
PFX_HD v3 mul(m3 m, v3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }
// transpose(M) * v
PFX_HD v3 tmul(m3 m, v3 v) { return v3{dot(m.c0, v), dot(m.c1, v), dot(m.c2, v)}; }

// Orthonormal frame around n (trace.wgsl `basis`, crates/trace/src/trace.wgsl:1270-1273)
PFX_HD m3 basis(v3 n) {
    v3 helper = fabsf(n.y) > 0.9f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
    v3 t = normalize(cross(helper, n));
    return m3{t, cross(n, t), n};
}

// Integer-ULP ray origin offset (Wachter and Binder, Ray Tracing Gems ch. 6),
// crates/trace/src/trace.wgsl:1257-1265, Rust twin crates/trace/src/math.rs:42-62
PFX_HD float offset_axis(float position, float normal) {
    if (fabsf(position) < 1.0f / 32.0f) return position + normal / 65536.0f;
    int32_t bits = (int32_t)f2u(position);
    float scaled = normal * 256.0f;
    int32_t shift = (int32_t)scaled;  // truncation toward zero, |scaled| <= 256
    uint32_t moved = (uint32_t)bits + (uint32_t)(position < 0.0f ? -shift : shift);
    return u2f(moved);
}
PFX_HD v3 offset_ray(v3 p, v3 n) {
    return v3{offset_axis(p.x, n.x), offset_axis(p.y, n.y), offset_axis(p.z, n.z)};
}
// crates/trace/src/trace.wgsl:1266-1269
PFX_HD v3 spawn_point(v3 p, v3 dir, v3 origin, float t) {
    float err = 4e-7f * (fmaxf(fabsf(origin.x), fmaxf(fabsf(origin.y), fabsf(origin.z))) + t);
    return offset_ray(p, dir) + dir * err;
}

// trace.wgsl `pcg` (crates/trace/src/trace.wgsl:168-172): PCG-RXS-M-XS hash
PFX_HD uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
// trace.wgsl `random` (:423-426)
PFX_HD float rnd(uint32_t* state) {
    *state = pcg(*state);
    return (float)(*state >> 8u) / 16777216.0f;
}

}  // namespace pfx
