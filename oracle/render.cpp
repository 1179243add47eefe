// Oracle (test infrastructure only): the per-pixel megakernel of the
// reference, restated on the CPU. One thread per pixel walks the whole path,
// as `main`/`radiance` do in crates/trace/src/trace.wgsl:2153-2579, with the
// light sampling of crates/trace/src/lights.wgsl and the BRDF of
// crates/materials/src/wgsl/brdf.wgsl. Features the still configs never
// enable are compiled out exactly as the reference's pipeline constants do
// (gpu.rs:486-529): no textures, lobes, wet, flakes, thin film, alpha, fog,
// sprites, motion, shapes or transmission.
#include <math.h>
#include <string.h>

#include <atomic>
#include <utility>
#include <cmath>
#include <string>
#include <vector>

#include "oracle.h"
#include "../pfx_amd/csrc/sobol.h"

namespace oracle {

using pfx::dot;
using pfx::fm_clamp;
using pfx::fm_cos;
using pfx::fm_pow;
using pfx::fm_sin;
using pfx::m3;
using pfx::mk;
using pfx::normalize;

static const float PI = 3.141592653589793f;  // brdf.wgsl:1
static const uint32_t SENTINEL = 0x76543210u;  // trace.wgsl:447
static const uint32_t INSTANCE_LEAF = 1u;
static const uint32_t ONE_SIDED = 8388608u;
static const uint32_t NO_MATERIAL = 0xffffffffu;

struct Mat {
    v3 base;
    float roughness, metalness, specular, clearcoat, clearcoat_roughness, sheen, fresnel_power;
    v3 emission;
};

struct Hit {
    float t;
    v3 n, geometric;
    uint32_t material, surface, record;
};

enum { ACC_COMP = 0, ACC_PLAIN = 1, ACC_F64 = 2 };

struct Scene {
    Laid laid;
    std::vector<Mat> materials;
    EmitterList emitters;
    LaidLights lights;
    Env env;
    SkyCdf sky;
    bool lit = false;
    bool area_lights = false;  // a looped row is a shadowed rect or disc light (light_hits can see it)
    std::vector<float> sobol;  // Sobol patterns (empty: PCG)
    uint32_t sobol_len = 0;
    // Frame (gpu.rs:2295-2316)
    v3 origin, forward, right, up;
    v3 sun_dir, sun_color;
    float sun_intensity = 0.0f, sun_w = 1.0f;
    float lens[4] = {0, 0, 0, 0};
    uint32_t window[4] = {0, 0, 0, 0};
    uint32_t projection = 0;            // 0 perspective, 1 orthographic, 2 equirect
    float half_view[2] = {0.0f, 0.0f};  // orthographic half extents
    uint32_t width = 0, height = 0;
    pfxh_bounces bounces{16, 4, 8, 12};
    uint32_t accumulate = ACC_COMP;
    bool watertight = false;  // PFXH_TRI_WATERTIGHT
    uint32_t samples = 0;
    // accumulation: 12 sums per pixel (color.xyzw, albedo, normal)
    std::vector<float> sum, comp;
    std::vector<double> sum64;
    std::vector<float> m2;  // per pixel: sum over samples of ((r + g + b) / 3)^2
    std::vector<float> snaps;  // adaptive Snap{mean, info}: 8 floats per pixel
    std::atomic<uint64_t> closest{0}, shadow{0}, paths{0};
};

// ---- traversal (trace.wgsl:449-468, 681-694, 754-804, 999-1168) -------------
static bool box_hit(const Node& node, v3 o, v3 d, float limit) {
    float near_ = 0.0f, far_ = limit;
    for (int axis = 0; axis < 3; ++axis) {
        float da = pfx::comp(d, axis), oa = pfx::comp(o, axis);
        if (fabsf(da) < 1e-12f) {
            if (oa < node.lo[axis] || oa > node.hi[axis]) return false;
        } else {
            float a = (node.lo[axis] - oa) / da;
            float b = (node.hi[axis] - oa) / da;
            near_ = fmaxf(near_, fminf(a, b));
            far_ = fminf(far_, fmaxf(a, b));
            if (near_ > far_) return false;
        }
    }
    return true;
}
static float dot4(const float* m, v3 p, float w) { return m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3] * w; }
static v3 to_object(const Node& m, v3 p, float w) {
    float r2[4] = {pfx::u2f(m.first), pfx::u2f(m.count), pfx::u2f(m.pad[0]), pfx::u2f(m.pad[1])};
    return mk(dot4(m.lo, p, w), dot4(m.hi, p, w), dot4(r2, p, w));
}
static v3 t3(const float* p) { return mk(p[0], p[1], p[2]); }

static void tri_fill(const Scene& sc, uint32_t i, uint32_t record, float t, float u, float v, Hit& hit) {
    const Tri& tri = sc.laid.tris[i];
    v3 e1 = t3(tri.b) - t3(tri.a), e2 = t3(tri.c) - t3(tri.a);
    hit.t = t;
    hit.surface = i;
    hit.record = record;
    v3 geometric = normalize(pfx::cross(e1, e2));
    hit.geometric = geometric;
    float w = (1.0f - u) - v;
    v3 shaded = t3(tri.na) * w + t3(tri.nb) * u + t3(tri.nc) * v;
    if (dot(shaded, shaded) > 1e-12f)
        hit.n = normalize(shaded);
    else
        hit.n = record > 0 ? mk(0.0f, 0.0f, 0.0f) : geometric;
    uint32_t info_y = sc.laid.surfaces[record].info[1];
    hit.material = info_y == NO_MATERIAL ? (uint32_t)tri.a[3] : info_y;
}
// Watertight ray/triangle test (Woop, Benthin and Wald, JCGT 2(1), 2013),
// FP32 without the paper's FP64 recheck of an edge function equal to 0
// (then counted inside). The ray frame: kz = axis of the largest |d|
// component (first of equals), kx = kz + 1, ky = kx + 1 (mod 3), swapped
// when d[kz] < 0; shear sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz =
// 1 / d[kz]. Each vertex minus the origin is sheared to (x, y) and the edge
// functions U, V, W (the barycentric weights of a, b, c times det) must
// share a sign. Candidate t = (U az + V bz + W cz) / det with az = sz (a -
// o)[kz]; u = V / det (weight of b), v = W / det (weight of c), each as a
// product with 1 / det. Returns false when the ray misses the triangle.
static bool watertight_test(const Tri& tri, v3 o, v3 d, float& t, float& u, float& v, float& det) {
    float dx = fabsf(d.x), dy = fabsf(d.y), dz = fabsf(d.z);
    int kz = (dx >= dy && dx >= dz) ? 0 : (dy >= dz ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (pfx::comp(d, kz) < 0.0f) std::swap(kx, ky);
    float sx = pfx::comp(d, kx) / pfx::comp(d, kz);
    float sy = pfx::comp(d, ky) / pfx::comp(d, kz);
    float sz = 1.0f / pfx::comp(d, kz);
    v3 A = t3(tri.a) - o, B = t3(tri.b) - o, C = t3(tri.c) - o;
    float ax = pfx::comp(A, kx) - sx * pfx::comp(A, kz), ay = pfx::comp(A, ky) - sy * pfx::comp(A, kz);
    float bx = pfx::comp(B, kx) - sx * pfx::comp(B, kz), by = pfx::comp(B, ky) - sy * pfx::comp(B, kz);
    float cx = pfx::comp(C, kx) - sx * pfx::comp(C, kz), cy = pfx::comp(C, ky) - sy * pfx::comp(C, kz);
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (sz * pfx::comp(A, kz)) + V * (sz * pfx::comp(B, kz)) + W * (sz * pfx::comp(C, kz));
    float rcp = 1.0f / det;
    t = T * rcp;
    u = V * rcp;
    v = W * rcp;
    return true;
}
static void tri_hit(const Scene& sc, uint32_t i, uint32_t record, uint32_t word, float inv_det, v3 o, v3 d,
                    bool shadow_ray, Hit& hit) {
    const Tri& tri = sc.laid.tris[i];
    if ((word & 1u) != 0u && !shadow_ray) return;
    if ((word & 2u) != 0u && shadow_ray) return;
    if (sc.watertight) {
        float t, u, v, det;
        if (!watertight_test(tri, o, d, t, u, v, det)) return;
        if (!shadow_ray && (word & ONE_SIDED) != 0u && det * inv_det < 0.0f) return;
        if (t > 0.0001f && t < hit.t) tri_fill(sc, i, record, t, u, v, hit);
        return;
    }
    v3 a = t3(tri.a);
    v3 e1 = t3(tri.b) - a, e2 = t3(tri.c) - a;
    v3 p = pfx::cross(d, e2);
    float det = dot(e1, p);
    float seen = det * inv_det;
    if (fabsf(seen) < 1e-9f) return;
    if (!shadow_ray && (word & ONE_SIDED) != 0u && seen < 0.0f) return;
    float inv = 1.0f / det;
    v3 s = o - a;
    float u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return;
    v3 q = pfx::cross(s, e1);
    float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return;
    float t = dot(e2, q) * inv;
    if (t > 0.0001f && t < hit.t) tri_fill(sc, i, record, t, u, v, hit);
}
static Hit world_hit(const Scene& sc, Hit h) {
    if (h.record == 0) return h;
    const Surface& s = sc.laid.surfaces[h.record];
    float mirror = pfx::u2f(s.info[3]) < 0.0f ? -1.0f : 1.0f;
    v3 i0 = t3(s.inverse), i1 = t3(s.inverse + 4), i2 = t3(s.inverse + 8);
    v3 g = h.geometric;
    h.geometric = normalize(i0 * g.x + i1 * g.y + i2 * g.z) * mirror;
    v3 shaded = i0 * h.n.x + i1 * h.n.y + i2 * h.n.z;
    h.n = pfx::all_zero(h.n) ? h.geometric : normalize(shaded);
    return h;
}

struct Counter {
    uint64_t nodes = 0, prims = 0;
};

static Hit scene_hit(const Scene& sc, v3 o, v3 d, float limit, bool shadow_ray, Counter* ctr = nullptr) {
    Hit hit;
    hit.t = limit;
    hit.n = mk(0, 0, 0);
    hit.geometric = mk(0, 0, 0);
    hit.material = 0;
    hit.surface = 0;
    hit.record = 0;
    if (sc.laid.slots > 0) {
        uint32_t stack[64];
        uint32_t top = 1;
        stack[0] = 0;
        v3 ro = o, rd = d;
        uint32_t base = 0, record = 0, word = 0;
        float inv_det = 1.0f;
        while (top > 0) {
            top -= 1;
            uint32_t index = stack[top];
            if (index == SENTINEL) {
                ro = o;
                rd = d;
                base = 0;
                record = 0;
                word = 0;
                inv_det = 1.0f;
                continue;
            }
            const Node& node = sc.laid.nodes[index];
            if (ctr) ctr->nodes++;
            if (!box_hit(node, ro, rd, hit.t)) continue;
            if (node.pad[0] == INSTANCE_LEAF) {
                if (top + 2 <= 64) {
                    stack[top] = SENTINEL;
                    stack[top + 1] = node.first;
                    top += 2;
                    base = node.count;
                    record = pfx::f2u(node.lo[3]);
                    const uint32_t* info = sc.laid.surfaces[record].info;
                    word = info[0];
                    inv_det = pfx::u2f(info[3]);
                    if (node.pad[1] > 0) {
                        const Node& m = sc.laid.nodes[node.pad[1]];
                        ro = to_object(m, o, 1.0f);
                        rd = to_object(m, d, 0.0f);
                    }
                }
            } else if (node.count == 0) {
                if (top + 2 <= 64) {
                    stack[top] = node.first;
                    stack[top + 1] = index + 1;
                    top += 2;
                }
            } else {
                for (uint32_t j = 0; j < node.count; ++j) {
                    if (ctr) ctr->prims++;
                    tri_hit(sc, base + node.first + j, record, word, inv_det, ro, rd, shadow_ray, hit);
                }
            }
        }
    }
    return world_hit(sc, hit);
}

static bool tri_blocks(const Scene& sc, uint32_t i, uint32_t word, float inv_det, v3 o, v3 d, float limit) {
    const Tri& tri = sc.laid.tris[i];
    if ((word & 2u) != 0u) return false;
    if (sc.watertight) {
        float t, u, v, det;
        if (!watertight_test(tri, o, d, t, u, v, det)) return false;
        return !(t <= 0.0001f || t >= limit);
    }
    v3 a = t3(tri.a);
    v3 e1 = t3(tri.b) - a, e2 = t3(tri.c) - a;
    v3 p = pfx::cross(d, e2);
    float det = dot(e1, p);
    if (fabsf(det * inv_det) < 1e-9f) return false;
    float inv = 1.0f / det;
    v3 s = o - a;
    float u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    v3 q = pfx::cross(s, e1);
    float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(e2, q) * inv;
    if (t <= 0.0001f || t >= limit) return false;
    return true;
}

static bool occluded(const Scene& sc, v3 o, v3 d, float limit, Counter* ctr = nullptr) {
    if (sc.laid.slots == 0) return false;
    uint32_t stack[64];
    uint32_t top = 1;
    stack[0] = 0;
    v3 ro = o, rd = d;
    uint32_t base = 0, word = 0;
    float inv_det = 1.0f;
    while (top > 0) {
        top -= 1;
        uint32_t index = stack[top];
        if (index == SENTINEL) {
            ro = o;
            rd = d;
            base = 0;
            word = 0;
            inv_det = 1.0f;
            continue;
        }
        const Node& node = sc.laid.nodes[index];
        if (ctr) ctr->nodes++;
        if (!box_hit(node, ro, rd, limit)) continue;
        if (node.pad[0] == INSTANCE_LEAF) {
            if (top + 2 <= 64) {
                stack[top] = SENTINEL;
                stack[top + 1] = node.first;
                top += 2;
                base = node.count;
                uint32_t record = pfx::f2u(node.lo[3]);
                const uint32_t* info = sc.laid.surfaces[record].info;
                word = info[0];
                inv_det = pfx::u2f(info[3]);
                if (node.pad[1] > 0) {
                    const Node& m = sc.laid.nodes[node.pad[1]];
                    ro = to_object(m, o, 1.0f);
                    rd = to_object(m, d, 0.0f);
                }
            }
        } else if (node.count == 0) {
            if (top + 2 <= 64) {
                stack[top] = node.first;
                stack[top + 1] = index + 1;
                top += 2;
            }
        } else {
            for (uint32_t j = 0; j < node.count; ++j) {
                if (ctr) ctr->prims++;
                if (tri_blocks(sc, base + node.first + j, word, inv_det, ro, rd, limit)) return true;
            }
        }
    }
    return false;
}

// ---- BRDF (crates/materials/src/wgsl/brdf.wgsl) ------------------------------
// A path's random numbers (trace.wgsl:423-446): PCG draws in order, or with
// the Sobol sampler the components of one dimension's point at a time
// (sobol.wgsl:44-55; the lookup shared with the product in sobol.h, the
// patterns built by this file's own generate_4d).
struct Rng {
    uint32_t state = 0;         // PCG state, or the Sobol pixel seed (rng.x)
    uint32_t index = 0;         // Sobol sample index (rng.y)
    uint32_t base = 0;          // rng.z
    const float* table = nullptr;
    uint32_t length = 0;
    pfx::Sobol4 u{};
    void at(uint32_t dim) {
        if (table) u = pfx::sobol_point(table, length, state, index, base + dim);
    }
    float x() { return table ? u.x : pfx::rnd(&state); }
    float y() { return table ? u.y : pfx::rnd(&state); }
    float z() { return table ? u.z : pfx::rnd(&state); }
    float w() { return table ? u.w : pfx::rnd(&state); }
};
static uint32_t light_dim(uint32_t k) { return k == 0u ? 1u : (k < 13u ? 3u + k : k << 16); }  // :442-446

static v3 mix3(v3 a, v3 b, float t) { return a * (1.0f - t) + b * t; }
static v3 schlick(v3 f0, float cos_theta, float power) {  // brdf.wgsl:41-44
    float k = fm_pow(1.0f - fm_clamp(cos_theta, 0.0f, 1.0f), power);
    return mix3(f0, mk(1, 1, 1), k);
}
static float ggx_d(float nh, float a2) {  // :88-92
    float n = fmaxf(nh, 0.0f);
    float k = n * n * (a2 - 1.0f) + 1.0f;
    return a2 / (PI * k * k);
}
static float smith_g1(float ndot, float a2) {  // :94-97
    float n = fmaxf(ndot, 0.0f);
    return 2.0f * n / (n + sqrtf(a2 + (1.0f - a2) * n * n));
}
static float alpha_rough(float r) {  // :110-113
    float rough = fm_clamp(r, 0.02f, 1.0f);
    return rough * rough;
}
static float fresnel_exponent(const Mat& s) { return s.fresnel_power <= 0.0f ? 5.0f : s.fresnel_power; }
static v3 reflection_filmed(const Mat& s, float ndh_cos) {  // :206-228 (no transmission, no film)
    float power = fresnel_exponent(s);
    v3 f0 = mix3(mk(s.specular, s.specular, s.specular), s.base, fm_clamp(s.metalness, 0.0f, 1.0f));
    if (f0.x <= 0.0f && f0.y <= 0.0f && f0.z <= 0.0f) return mk(0, 0, 0);
    return schlick(f0, ndh_cos, power);
}
static v3 ggx_spec(float ndv, float ndl, float ndh, float a2, v3 fresnel) {  // :230-232
    return fresnel * ggx_d(ndh, a2) * (smith_g1(ndv, a2) * smith_g1(ndl, a2)) / (4.0f * ndv * ndl);
}
static v3 eval(const Mat& s, v3 wo, v3 wi) {  // :234-284 (eval_filmed without film)
    float ndv = wo.z;
    if (ndv <= 1e-5f) return mk(0, 0, 0);
    float ndl = wi.z;
    if (fabsf(ndl) <= 1e-8f) return mk(0, 0, 0);
    float metal = fm_clamp(s.metalness, 0.0f, 1.0f);
    float sheen = fm_clamp(s.sheen, 0.0f, 1.0f);
    float sheen_cover = sheen * fm_pow(1.0f - ndv, 4.0f);
    float view = fm_clamp(ndv, 0.0f, 1.0f);
    if (ndl < 0.0f) return mk(0, 0, 0);  // back lobe weight is zero without transmission/subsurface
    v3 h = normalize(wo + wi);
    float ndh = fmaxf(h.z, 0.0f);
    float vdh = fmaxf(dot(wo, h), 0.0f);
    v3 fresnel = reflection_filmed(s, vdh);
    float a = alpha_rough(s.roughness);
    v3 spec = ggx_spec(ndv, ndl, ndh, a * a, fresnel);
    v3 transmitted = mk(1, 1, 1) - fresnel;
    float body = (1.0f - metal) * 1.0f * 1.0f;
    float diffuse_w = body * (1.0f - sheen_cover);
    float sheen_w = body * sheen_cover * fm_pow(1.0f - ndh, 5.0f);
    v3 colour = spec + s.base * transmitted * (diffuse_w + sheen_w) / PI;
    float coat = fm_clamp(s.clearcoat, 0.0f, 1.0f);
    if (coat > 0.0f) {
        float coat_a = alpha_rough(s.clearcoat_roughness);
        v3 coat_f = schlick(mk(0.04f, 0.04f, 0.04f), vdh, fresnel_exponent(s));
        v3 coat_spec = ggx_spec(ndv, ndl, ndh, coat_a * coat_a, coat_f);
        float fc = schlick(mk(0.04f, 0.04f, 0.04f), view, fresnel_exponent(s)).x;
        colour = coat_spec * coat + colour * (1.0f - coat * fc);
    }
    return colour;
}
static float coat_weight(const Mat& s, v3 wo) {  // trace.wgsl:1749, 1777
    return fm_clamp(s.clearcoat * (0.12f + 0.88f * (0.04f + 0.96f * fm_pow(1.0f - fmaxf(wo.z, 0.0f), 5.0f))), 0.0f, 0.9f);
}
static float bsdf_pdf(const Mat& s, v3 wo, v3 wi) {  // trace.wgsl:1774-1788
    if (wi.z <= 0.0f) return 0.0f;
    float coat = coat_weight(s, wo);
    float spec = (1.0f - coat) * fm_clamp(s.metalness + s.specular, 0.0f, 0.9f);
    v3 h = normalize(wo + wi);
    float a = alpha_rough(s.roughness);
    float ca = alpha_rough(s.clearcoat_roughness);
    float denom = fmaxf(4.0f * fabsf(dot(wi, h)), 1e-8f);
    float hz = fmaxf(h.z, 0.0f);
    float base_pdf = ggx_d(hz, a * a) * hz / denom;
    float coat_pdf = ggx_d(hz, ca * ca) * hz / denom;
    return 1.0f * (coat * coat_pdf + spec * base_pdf + (1.0f - coat - spec) * wi.z / PI);
}
struct BsdfSample {
    v3 dir;
    float pdf;
    uint32_t lobe;
};
static v3 reflect(v3 e1, v3 e2) { return e1 - e2 * (2.0f * dot(e2, e1)); }
static BsdfSample sample_bsdf(const Mat& s, v3 wo, Rng* state) {  // trace.wgsl:1716-1773
    float transmission = 0.0f;
    state->at(3u);
    if (state->w() < transmission) { /* never: no transmission */ }
    float cw = coat_weight(s, wo);
    float sw = (1.0f - cw) * fm_clamp(s.metalness + s.specular, 0.0f, 0.9f);
    float pick = state->z();
    if (pick < cw + sw) {
        bool coat = pick < cw;
        float alpha = alpha_rough(coat ? s.clearcoat_roughness : s.roughness);
        float u1 = state->x();
        float u2 = state->y();
        float phi = 2.0f * PI * u1;
        float cos_theta = sqrtf((1.0f - u2) / (1.0f + (alpha * alpha - 1.0f) * u2));
        float st = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
        v3 h = mk(fm_cos(phi) * st, fm_sin(phi) * st, cos_theta);
        v3 wi = reflect(-wo, h);
        return BsdfSample{wi, bsdf_pdf(s, wo, wi), 1u};
    }
    float u1 = state->x();
    float u2 = state->y();
    float phi = 2.0f * PI * u1;
    float r = sqrtf(u2);
    v3 wi = mk(fm_cos(phi) * r, fm_sin(phi) * r, sqrtf(1.0f - u2));
    return BsdfSample{wi, bsdf_pdf(s, wo, wi), 0u};
}

// ---- lights (crates/trace/src/lights.wgsl) -----------------------------------
struct v4 {
    v3 xyz;
    float w;
};
struct Candidate {
    v3 wi, scale;
    float pdf, roughness;
    v3 direction;
    float limit;
};
static Candidate candidate_none() { return Candidate{mk(0, 0, 1), mk(0, 0, 0), 0.0f, 0.0f, mk(0, 0, 1), -1.0f}; }
static float mis_power(float chosen, float other) {  // :7-11
    if (chosen <= 0.0f) return 0.0f;
    float ratio = other / chosen;
    return 1.0f / (1.0f + ratio * ratio);
}
static uint32_t find_emitter(const Scene& sc, uint32_t count, float u) {  // :12-21
    uint32_t low = 0, high = count;
    while (low < high) {
        uint32_t mid = (low + high) / 2;
        if (sc.emitters.cdf[mid] <= u)
            low = mid + 1;
        else
            high = mid;
    }
    return low < count - 1 ? low : count - 1;
}
static float emitter_pdf(const GEmitter& e, v3 x, v3 apex, v3 normal) {  // :31-37 (triangles)
    v3 to = x - apex;
    float distance2 = dot(to, to);
    float cosine = fabsf(dot(normal, to)) / sqrtf(fmaxf(distance2, 1e-30f));
    return e.e[3] / fmaxf(e.c[3], 1e-30f) * distance2 / fmaxf(cosine, 1e-7f);
}
static float emitter_weight(const Scene& sc, const Hit& hit, v3 p, v3 previous_point, float previous_pdf) {  // :38-44
    const Mat& m = sc.materials[hit.material];
    if (previous_pdf <= 0.0f || (m.emission.x <= 0.0f && m.emission.y <= 0.0f && m.emission.z <= 0.0f)) return 1.0f;
    uint32_t index = (uint32_t)sc.laid.tris[hit.surface].c[3];
    if (index == 0) return 1.0f;
    return mis_power(previous_pdf, emitter_pdf(sc.emitters.entries[index], p, previous_point, hit.geometric));
}
static v4 shadow_ray_to(v3 origin, v3 aim, bool shadow) {  // lights.wgsl:95-102
    if (!shadow) return v4{mk(0, 0, 1), 0.0f};
    v3 ray = aim - origin;
    float distance = pfx::length(ray);
    float limit = distance * 0.9999f - 0.0002f;
    if (limit <= 0.0f) return v4{mk(0, 0, 1), 0.0f};
    return v4{ray / distance, limit};
}
static Candidate emitter_candidate(const Scene& sc, const m3& T, v3 p, v3 origin, Rng* state) {  // :103-144
    uint32_t count = (uint32_t)sc.emitters.entries[0].a[0];
    const GEmitter& e = sc.emitters.entries[find_emitter(sc, count, state->z()) + 1];
    float r1 = sqrtf(state->x());
    float r2 = state->y();
    v3 a = t3(e.a), b = t3(e.b), c = t3(e.c);
    v3 x = a * (1.0f - r1) + b * (r1 * (1.0f - r2)) + c * (r1 * r2);
    v3 to = x - p;
    float distance2 = dot(to, to);
    v3 direction = to / sqrtf(fmaxf(distance2, 1e-30f));
    v3 local = pfx::tmul(T, direction);
    if (local.z <= 0.0f) return candidate_none();
    v3 normal = normalize(pfx::cross(b - a, c - a));
    if (fabsf(dot(normal, p - a)) <= 1e-5f) return candidate_none();
    float cosine = fabsf(dot(normal, direction));
    if (cosine <= 1e-7f) return candidate_none();
    v4 ray = shadow_ray_to(origin, x, true);
    return Candidate{local, t3(e.e), e.e[3] / e.c[3] * distance2 / cosine, 0.0f, ray.xyz, ray.w};
}
// lights.wgsl:57-63: (unit normal, area) of a rect or disc light row
static v4 area_frame(const GLight& light) {
    if ((uint32_t)light.a[3] == 3u) return v4{t3(light.c), PI * light.d[0] * light.d[0]};
    v3 spanned = pfx::cross(t3(light.c), t3(light.d)) * 4.0f;
    float area = pfx::length(spanned);
    return v4{spanned / area, area};
}
// lights.wgsl:64-91: radiance of the shadowed area lights among the looped
// rows that the ray (o, d) meets before `limit`, MIS-weighted against light
// sampling when the ray was BSDF-sampled with pdf previous_pdf
static v3 light_hits(const Scene& sc, v3 o, v3 d, float limit, v3 previous_point, float previous_pdf) {
    uint32_t count = sc.lights.looped;
    v3 total = mk(0, 0, 0);
    for (uint32_t i = 0; i < count; ++i) {
        const GLight& light = sc.lights.rows[i];
        uint32_t kind = (uint32_t)light.a[3];
        if ((kind != 2u && kind != 3u) || light.d[3] < 0.5f) continue;
        v4 plane = area_frame(light);
        float facing = dot(d, plane.xyz);
        if (fabsf(facing) < 1e-9f || (light.c[3] < 0.5f && facing >= 0.0f)) continue;
        float t = dot(t3(light.a) - o, plane.xyz) / facing;
        if (t <= 0.0001f || t >= limit) continue;
        v3 x = o + d * t;
        v3 q = x - t3(light.a);
        if (kind == 3u) {
            if (dot(q, q) > light.d[0] * light.d[0]) continue;
        } else if (fabsf(dot(q, t3(light.c))) > dot(t3(light.c), t3(light.c)) ||
                   fabsf(dot(q, t3(light.d))) > dot(t3(light.d), t3(light.d))) {
            continue;
        }
        float weight = 1.0f;
        if (previous_pdf > 0.0f) {
            v3 to = x - previous_point;
            float distance2 = dot(to, to);
            float cosine = fabsf(dot(plane.xyz, to)) / sqrtf(fmaxf(distance2, 1e-30f));
            weight = mis_power(previous_pdf, distance2 / (plane.w * fmaxf(cosine, 1e-7f)));
        }
        total = total + t3(light.b) * weight;
    }
    return total;
}
static Candidate light_candidate(const GLight& light, const m3& T, v3 p, v3 origin, v3 geometric,
                                 Rng* state) {  // :145-200
    bool shadow = light.d[3] > 0.5f;
    uint32_t kind = (uint32_t)light.a[3];
    if (kind == 2u || kind == 3u) {  // :149-170: a point on the rect or disc, area pdf in solid angle
        float u1 = state->x();
        float u2 = state->y();
        v4 plane = area_frame(light);
        v3 x = t3(light.a) + t3(light.c) * (2.0f * u1 - 1.0f) + t3(light.d) * (2.0f * u2 - 1.0f);
        if (kind == 3u) {
            m3 disc = pfx::basis(t3(light.c));
            float r = light.d[0] * sqrtf(u1);
            float phi = 2.0f * PI * u2;
            x = t3(light.a) + disc.c0 * (r * fm_cos(phi)) + disc.c1 * (r * fm_sin(phi));
        }
        v3 to = x - p;
        float distance2 = dot(to, to);
        v3 direction = to / sqrtf(fmaxf(distance2, 1e-30f));
        v3 local = pfx::tmul(T, direction);
        if (local.z <= 0.0f) return candidate_none();
        float facing = -dot(plane.xyz, direction);
        float cosine = light.c[3] > 0.5f ? fabsf(facing) : facing;
        if (cosine <= 1e-7f) return candidate_none();
        float pdf = distance2 / (plane.w * cosine);
        v4 ray = shadow_ray_to(origin, x, shadow);
        if (!shadow) return Candidate{local, t3(light.b) / pdf, 0.0f, 0.0f, ray.xyz, ray.w};
        return Candidate{local, t3(light.b), pdf, 0.0f, ray.xyz, ray.w};
    }
    v3 to = t3(light.a) - p;
    float distance = pfx::length(to);
    float range = light.b[3];
    if (distance >= range || distance <= 1e-6f) return candidate_none();
    v3 l = to / distance;
    v3 local = pfx::tmul(T, l);
    if (local.z <= 0.0f || dot(geometric, l) <= 0.0f) return candidate_none();
    float cone = 1.0f;
    if ((uint32_t)light.a[3] == 1u) {
        float along = fm_clamp(dot(-l, t3(light.c)) * light.d[0] + light.d[1], 0.0f, 1.0f);
        cone = along * along;
        if (cone <= 0.0f) return candidate_none();
    }
    float radius = light.c[3];
    v3 aim = t3(light.a);
    if (shadow && radius > 0.0f) {
        m3 disc = pfx::basis(l);
        float r = radius * sqrtf(state->x());
        float phi = 2.0f * PI * state->y();
        aim = aim + (disc.c0 * (r * fm_cos(phi)) + disc.c1 * (r * fm_sin(phi)));
    }
    float ratio = distance / range;
    float window = fm_clamp(1.0f - ratio * ratio * ratio * ratio, 0.0f, 1.0f);
    float reach = fmaxf(distance, radius);
    v4 ray = shadow_ray_to(origin, aim, shadow);
    float roughness = sqrtf(fm_clamp(radius / (2.0f * distance), 0.0f, 1.0f));
    return Candidate{local, t3(light.b) * (window * window * cone / (reach * reach)), 0.0f, roughness, ray.xyz, ray.w};
}
static float sin_from_cos(float c) { return sqrtf(fmaxf(1.0f - c * c, 0.0f)); }
struct F2 {
    float x, y;
};
static F2 tree_importance(v3 n, v3 to, float cos_u, v3 axis, float theta_o, float theta_e, float far_,
                          float near_, float energy) {  // :204-237
    float sin_u = sin_from_cos(cos_u);
    float cos_i = dot(to, n);
    float sin_i = sin_from_cos(cos_i);
    float cos_min_i = cos_i >= cos_u ? 1.0f : cos_i * cos_u + sin_i * sin_u;
    if (cos_min_i < 0.0f) return F2{0.0f, 0.0f};
    float cos_max_i = fmaxf(cos_i * cos_u - sin_i * sin_u, 0.0f);
    float cos_t = 1.0f, sin_t = 0.0f;
    v3 nt = -to;
    if (axis.x != nt.x || axis.y != nt.y || axis.z != nt.z) {
        cos_t = dot(axis, nt);
        sin_t = sin_from_cos(cos_t);
    }
    float cos_t_minus_u = cos_t * cos_u + sin_t * sin_u;
    float cos_o = fm_cos(theta_o);
    float sin_o = fm_sin(theta_o);
    float cos_min_out;
    if (cos_t >= cos_u || cos_t_minus_u >= cos_o) {
        cos_min_out = 1.0f;
    } else if (theta_o + theta_e > PI || cos_t_minus_u > fm_cos(theta_o + theta_e)) {
        cos_min_out = cos_t_minus_u * cos_o + sin_from_cos(cos_t_minus_u) * sin_o;
    } else {
        return F2{0.0f, 0.0f};
    }
    float most = fabsf(cos_min_i * energy * cos_min_out / (near_ * near_));
    float cos_t_plus_u = cos_t * cos_u - sin_t * sin_u;
    if (theta_e - theta_o < 0.0f || cos_t < 0.0f || cos_u < 0.0f || cos_t_plus_u < fm_cos(theta_e - theta_o))
        return F2{most, 0.0f};
    float cos_max_out = cos_t_plus_u * cos_o - sin_from_cos(cos_t_plus_u) * sin_o;
    return F2{most, fabsf(cos_max_i * energy * cos_max_out / (far_ * far_))};
}
static F2 node_importance(const GLight& node, v3 p, v3 n) {  // :238-250
    v3 lo = t3(node.a), hi = t3(node.b);
    v3 centroid = (lo + hi) * 0.5f;
    v3 extent = hi - centroid;
    if (dot(n, centroid - p) + dot(pfx::vabs(n), pfx::vabs(extent)) <= 0.0f) return F2{0.0f, 0.0f};
    if (pfx::length(p - pfx::vmin(pfx::vmax(p, lo), hi)) >= node.b[3]) return F2{0.0f, 0.0f};
    v3 to = centroid - p;
    float reach = pfx::length(to);
    float radius2 = dot(extent, extent);
    float cos_u = -1.0f;
    if (reach * reach > radius2) cos_u = sqrtf(fmaxf(1.0f - radius2 / (reach * reach), 0.0f));
    float distance = fmaxf(fmaxf(0.5f * sqrtf(radius2), reach), 1e-6f);
    return tree_importance(n, to / fmaxf(reach, 1e-30f), cos_u, t3(node.c), node.c[3], node.d[0], distance, distance,
                           node.a[3]);
}
static float spot_cone(const GLight& light, float cosine) {  // :251-254
    float along = fm_clamp(cosine * light.d[0] + light.d[1], 0.0f, 1.0f);
    return along * along;
}
static F2 emitter_importance(const GLight& light, v3 p, v3 n) {  // :255-271
    v3 to = t3(light.a) - p;
    float near_ = pfx::length(to);
    if (near_ >= light.b[3] || near_ <= 1e-6f) return F2{0.0f, 0.0f};
    v3 toward = to / near_;
    float radius = light.c[3];
    float far_ = sqrtf(radius * radius + near_ * near_);
    float cos_u = near_ / far_;
    float energy = (light.b[0] + light.b[1] + light.b[2]) / 3.0f;
    if ((uint32_t)light.a[3] == 1u) {
        float theta_e = pfx::fm_acos(fm_clamp(-light.d[1] / light.d[0], -1.0f, 1.0f));
        // cos(max(0, acos(cone . -toward) - acos(cos_u))) through the identity in
        // fmath.h (the same value in real arithmetic, no transcendental calls)
        float cos_min_out = pfx::fm_cos_acos_gap(fm_clamp(dot(t3(light.c), -toward), -1.0f, 1.0f), cos_u);
        energy *= spot_cone(light, cos_min_out);
        return tree_importance(n, toward, cos_u, t3(light.c), 0.0f, theta_e, far_, near_, energy);
    }
    return tree_importance(n, toward, cos_u, -toward, 0.0f, PI / 2.0f, far_, near_, energy);
}
static F2 child_importance(const Scene& sc, const GLight& node, v3 p, v3 n) {  // :272-275
    if ((uint32_t)node.d[1] == 1u) return emitter_importance(sc.lights.rows[(uint32_t)node.d[2]], p, n);
    return node_importance(node, p, n);
}
static uint32_t popcount(uint32_t x) {
    uint32_t c = 0;
    while (x) {
        c += x & 1u;
        x >>= 1;
    }
    return c;
}
static F2 tree_leaf(const Scene& sc, uint32_t first, uint32_t count, v3 p, v3 n, float start) {  // :276-349
    float rand = start;
    bool sample_max = rand > 0.5f;
    if (count > 1u) rand = rand * 2.0f - (sample_max ? 1.0f : 0.0f);
    int selected = -1;
    float chosen = 0.0f, chosen_other = 0.0f, total = 0.0f, total_other = 0.0f;
    uint32_t has = 0;
    for (uint32_t i = 0; i < count; ++i) {
        int index = (int)(first + i);
        F2 importance = emitter_importance(sc.lights.rows[first + i], p, n);
        float weight = sample_max ? importance.x : importance.y;
        float other = sample_max ? importance.y : importance.x;
        if (weight > 0.0f) {
            total += weight;
            if (selected < 0) {
                selected = index;
                chosen = weight;
            } else {
                float thresh = weight / total;
                if (rand <= thresh) {
                    selected = index;
                    chosen = weight;
                    rand = rand / thresh;
                } else {
                    rand = (rand - thresh) / (1.0f - thresh);
                }
                rand = fm_clamp(rand, 0.0f, 1.0f);
            }
        }
        if (selected == index) chosen_other = other;
        total_other += other;
        has |= (importance.x > 0.0f ? 1u : 0u) << i;
    }
    if (has == 0u) return F2{0.0f, 0.0f};
    float chosen_max = sample_max ? chosen : chosen_other;
    float total_max = sample_max ? total : total_other;
    float chosen_min = sample_max ? chosen_other : chosen;
    float total_min = sample_max ? total_other : total;
    if (total_min == 0.0f) {
        if (sample_max) {
            chosen_min = 1.0f;
            total_min = (float)popcount(has);
        } else {
            selected = -1;
            float valid = 0.0f;
            uint32_t bits = has;
            for (uint32_t i = 0; i < count; ++i) {
                if ((bits & 1u) != 0u) {
                    valid += 1.0f;
                    if (selected < 0) {
                        selected = (int)(first + i);
                    } else {
                        float thresh = 1.0f / valid;
                        if (rand <= thresh) {
                            selected = (int)(first + i);
                            rand = rand / thresh;
                        } else {
                            rand = (rand - thresh) / (1.0f - thresh);
                        }
                        rand = fm_clamp(rand, 0.0f, 1.0f);
                    }
                }
                bits >>= 1u;
            }
            chosen_min = 1.0f;
            total_min = valid;
            chosen_max = emitter_importance(sc.lights.rows[(uint32_t)selected], p, n).x;
        }
    }
    return F2{(float)selected, 0.5f * (chosen_max / total_max + chosen_min / total_min)};
}
F2 tree_pick(const Scene& sc, v3 p, v3 n, float start) {  // :350-386
    uint32_t at = (uint32_t)(sc.emitters.entries[0].a[1] + sc.emitters.entries[0].a[3]);
    float rand = start;
    float pdf = 1.0f;
    for (;;) {
        const GLight& node = sc.lights.rows[at];
        uint32_t count = (uint32_t)node.d[1];
        if (count > 0u) {
            F2 leaf = tree_leaf(sc, (uint32_t)node.d[2], count, p, n, rand);
            return F2{leaf.x, leaf.y * pdf};
        }
        uint32_t left = at + 1u;
        uint32_t right = (uint32_t)node.d[2];
        F2 l = child_importance(sc, sc.lights.rows[left], p, n);
        F2 r = child_importance(sc, sc.lights.rows[right], p, n);
        float total_max = l.x + r.x;
        if (total_max == 0.0f) return F2{0.0f, 0.0f};
        float total_min = l.y + r.y;
        float probability_min = 0.5f * ((l.x > 0.0f ? 1.0f : 0.0f) + (r.x == 0.0f ? 1.0f : 0.0f));
        if (total_min > 0.0f) probability_min = l.y / total_min;
        float left_probability = 0.5f * (l.x / total_max + probability_min);
        float right_weight = 1.0f - left_probability;
        at = left;
        if (right_weight > 0.0f) {
            float thresh = right_weight / (left_probability + right_weight);
            if (rand <= thresh) {
                at = right;
                rand = rand / thresh;
            } else {
                rand = (rand - thresh) / (1.0f - thresh);
            }
            rand = fm_clamp(rand, 0.0f, 1.0f);
        }
        pdf *= at == left ? left_probability : 1.0f - left_probability;
    }
}
static Candidate tree_light(const Scene& sc, const m3& T, v3 p, v3 origin, v3 geometric, Rng* state) {  // :387-393
    F2 picked = tree_pick(sc, p, T.c2, state->z());
    if (!(picked.y > 0.0f)) return candidate_none();
    Candidate c = light_candidate(sc.lights.rows[(uint32_t)picked.x], T, p, origin, geometric, state);
    c.scale = c.scale / picked.y;
    return c;
}
static uint32_t light_slots(const Scene& sc) {  // :400-402
    if (!sc.lit) return 0u;
    return 1u + (uint32_t)sc.emitters.entries[0].a[1] + (sc.emitters.entries[0].a[3] > 0.5f ? 1u : 0u);
}
static Candidate direct_light(const Scene& sc, uint32_t slot, const m3& T, v3 p, v3 origin, v3 geometric,
                              Rng* state) {  // :403-409
    uint32_t looped = (uint32_t)sc.emitters.entries[0].a[1];
    if (slot > looped) return tree_light(sc, T, p, origin, geometric, state);
    if (slot > 0u) return light_candidate(sc.lights.rows[slot - 1u], T, p, origin, geometric, state);
    if (sc.emitters.entries[0].a[0] < 0.5f) return candidate_none();
    return emitter_candidate(sc, T, p, origin, state);
}

// ---- sky sampling (crates/trace/src/sky.rs:371-440, SKY_SAMPLING_WGSL) -------
static uint32_t sky_find(const float* cdf, uint32_t n, float needle) {
    uint32_t low = 0, high = n;
    while (low < high) {
        uint32_t mid = (low + high) / 2u;
        if (cdf[mid] <= needle)
            low = mid + 1u;
        else
            high = mid;
    }
    return low < n - 1u ? low : n - 1u;
}
struct SkySample {
    v3 direction;
    float pdf;
};
static SkySample sky_sample(const Scene& sc, float rx, float ry) {
    const SkyCdf& k = sc.sky;
    float u = fm_clamp(rx, 0.0f, 0.99999994f);
    float v = fm_clamp(ry, 0.0f, 0.99999994f);
    uint32_t row = sky_find(k.rows.data(), k.height, v);
    float row_before = row > 0u ? k.rows[row - 1u] : 0.0f;
    float fy = (v - row_before) / fmaxf(k.rows[row] - row_before, 0.000001f);
    float needle = fmaxf(u - 0.0f, 0.0f) / fmaxf(1.0f - 0.0f, 0.000001f);
    const float* cols = k.columns.data() + (size_t)row * k.width;
    uint32_t col = sky_find(cols, k.width, needle);
    float col_before = col > 0u ? cols[col - 1u] : 0.0f;
    float fx = (needle - col_before) / fmaxf(cols[col] - col_before, 0.000001f);
    float theta = 3.14159265f * ((float)row + fy) / (float)k.height;
    float phi = 6.28318531f * ((float)col + fx) / (float)k.width;
    v3 direction = mk(fm_sin(theta) * fm_cos(phi), fm_cos(theta), fm_sin(theta) * fm_sin(phi));
    float area = 6.28318531f / (float)k.width * 3.14159265f / (float)k.height * fmaxf(fm_sin(theta), 0.00001f);
    float mass = (k.rows[row] - row_before) * (cols[col] - col_before);
    return SkySample{direction, (1.0f - 0.0f) * mass / area + 0.0f};
}
static float wgsl_fmod(float x, float y) { return x - y * truncf(x / y); }
static float sky_pdf(const Scene& sc, v3 direction) {
    const SkyCdf& k = sc.sky;
    v3 d = normalize(direction);
    float theta = pfx::fm_acos(fm_clamp(d.y, -1.0f, 1.0f));
    float phi = wgsl_fmod(pfx::fm_atan2(d.z, d.x) + 6.28318531f, 6.28318531f);
    float fr = theta / 3.14159265f * (float)k.height, fc = phi / 6.28318531f * (float)k.width;
    uint32_t row = fr >= (float)k.height ? k.height - 1u : (uint32_t)fmaxf(fr, 0.0f);
    uint32_t col = fc >= (float)k.width ? k.width - 1u : (uint32_t)fmaxf(fc, 0.0f);
    if (row > k.height - 1u) row = k.height - 1u;
    if (col > k.width - 1u) col = k.width - 1u;
    float row_before = row > 0u ? k.rows[row - 1u] : 0.0f;
    const float* cols = k.columns.data() + (size_t)row * k.width;
    float col_before = col > 0u ? cols[col - 1u] : 0.0f;
    float mass = (k.rows[row] - row_before) * (cols[col] - col_before);
    float area = 6.28318531f / (float)k.width * 3.14159265f / (float)k.height * fmaxf(fm_sin(theta), 0.00001f);
    return (1.0f - 0.0f) * mass / area + 0.0f;
}

// ---- radiance (crates/trace/src/trace.wgsl:2153-2502) ------------------------
struct Result {
    v3 color;
    float albedo[4];
    float normal[4];
};
static Result radiance(const Scene& sc, v3 o0, v3 d0, Rng* state, uint64_t* closest, uint64_t* shadows) {
    Result res;
    memset(&res, 0, sizeof res);
    v3 o = o0, d = d0;
    v3 throughput = mk(1, 1, 1);
    v3 result = mk(0, 0, 0);
    float previous_pdf = 0.0f;
    uint32_t lobes[3] = {0, 0, 0};
    bool last = false;
    const pfxh_bounces& B = sc.bounces;
    uint32_t slots = light_slots(sc);
    for (uint32_t bounce = 0; bounce <= B.total + 1u; ++bounce) {
        state->base = 16u * (bounce + 1u);  // trace.wgsl:2170
        Hit hit = scene_hit(sc, o, d, 1e20f, false);
        *closest += 1;
        v3 origin_point = pfx::spawn_point(o, -d, o, 0.0f);
        if (sc.area_lights) result = result + throughput * light_hits(sc, o, d, hit.t, origin_point, previous_pdf);
        if (hit.t >= 1e19f) {
            float pdf = sky_pdf(sc, d);
            float weight = (bounce > 0u && previous_pdf > 0.0f) ? previous_pdf / (previous_pdf + pdf) : 1.0f;
            v3 environment = env_radiance(sc.env, d);  // sky_seen with PLAIN_LOOK is the radiance itself
            result = result + throughput * environment * weight;
            break;
        }
        v3 p = o + d * hit.t;
        v3 n = dot(hit.n, d) > 0.0f ? -hit.n : hit.n;
        const Mat& s0 = sc.materials[hit.material];
        Mat s = s0;
        v3 geometric = dot(hit.geometric, d) > 0.0f ? -hit.geometric : hit.geometric;
        if (dot(n, geometric) < 0.05f) n = normalize(n + geometric * 0.1f);
        s.roughness = fmaxf(s.roughness, 0.05f);
        s.clearcoat_roughness = fmaxf(s.clearcoat_roughness, 0.05f);
        if (bounce == 0u) {
            res.albedo[0] = s.base.x;
            res.albedo[1] = s.base.y;
            res.albedo[2] = s.base.z;
            res.albedo[3] = 1.0f;
            res.normal[0] = n.x;
            res.normal[1] = n.y;
            res.normal[2] = n.z;
            res.normal[3] = 1.0f;
        }
        if (sc.lit) s.emission = s.emission * emitter_weight(sc, hit, p, origin_point, previous_pdf);
        result = result + throughput * s.emission;
        if (last) break;
        m3 T = pfx::basis(n);
        v3 wo = pfx::tmul(T, -d);
        v3 origin = pfx::spawn_point(p, n, o, hit.t);
        for (uint32_t k = 0; k < 2u + slots; ++k) {
            state->at(light_dim(k));
            v3 local = mk(0, 0, 1), direction = mk(0, 0, 1), strength = mk(0, 0, 0);
            float limit = 1e19f, pdf = 0.0f;
            Mat lit = s;
            if (k == 0u) {
                if (sc.sun_intensity <= 0.0f) continue;
                direction = normalize(sc.sun_dir);
                if (sc.sun_w < 1.0f) {
                    float height = 1.0f - state->x() * (1.0f - sc.sun_w);
                    float radius = sqrtf(fmaxf(0.0f, 1.0f - height * height));
                    float phi = 2.0f * PI * state->y();
                    m3 orient = pfx::basis(direction);
                    direction = normalize(pfx::mul(orient, mk(radius * fm_cos(phi), radius * fm_sin(phi), height)));
                }
                local = pfx::tmul(T, direction);
                if (local.z <= 0.0f) continue;
                strength = sc.sun_color * (sc.sun_intensity * 1.0f);
            } else if (k == 1u) {
                float rx = state->x();
                float ry = state->y();
                SkySample light = sky_sample(sc, rx, ry);
                direction = light.direction;
                local = pfx::tmul(T, direction);
                if (local.z <= 0.0f) continue;
                strength = env_radiance(sc.env, direction) * 1.0f;
                pdf = light.pdf;
                if (!(pdf > 0.0f)) continue;
            } else {
                Candidate picked = direct_light(sc, k - 2u, T, p, origin, geometric, state);
                if (picked.limit < 0.0f) continue;
                local = picked.wi;
                direction = picked.direction;
                strength = picked.scale;
                pdf = picked.pdf;
                limit = picked.limit;
                lit.roughness = fmaxf(s.roughness, picked.roughness);
            }
            if (pdf > 0.0f) {
                float bsdf = bsdf_pdf(s, wo, local);
                if (k == 1u) {
                    strength = strength / (pdf + bsdf);
                } else {
                    float ratio = bsdf / pdf;
                    strength = strength / (pdf * (1.0f + ratio * ratio));
                }
            }
            v3 add = throughput * eval(lit, wo, local) * (pdf > 0.0f ? 1.0f : 1.0f) * local.z * strength;
            if (limit > 0.0f) {
                if (add.x == 0.0f && add.y == 0.0f && add.z == 0.0f) continue;  // a zero term leaves result unchanged
                *shadows += 1;
                if (occluded(sc, origin, direction, limit)) continue;
            }
            result = result + add;
        }
        BsdfSample sample = sample_bsdf(s, wo, state);
        v3 wi = sample.dir;
        lobes[sample.lobe] += 1u;
        last = bounce >= B.total || lobes[0] > B.diffuse || lobes[1] > B.glossy || lobes[2] > B.transmission;
        if (wi.z <= 0.0f || sample.pdf <= 0.0f) break;
        throughput = throughput * (eval(s, wo, wi) * 1.0f * wi.z / sample.pdf * 1.0f);
        d = normalize(pfx::mul(T, wi));
        previous_pdf = sample.pdf;
        o = pfx::spawn_point(p, d, o, hit.t);
        if (bounce >= 3u) {
            float keep = fm_clamp(fmaxf(throughput.x, fmaxf(throughput.y, throughput.z)), 0.05f, 0.95f);
            state->at(0u);
            if (state->x() > keep) break;
            throughput = throughput / keep;
        }
    }
    res.color = result;
    return res;
}

// ---- main: camera rays and accumulation (trace.wgsl:2503-2579) ----------------
static void camera_ray(const Scene& sc, uint32_t px, uint32_t py, float jx, float jy, v3* o, v3* d, Rng* state,
                       bool centre) {
    float ndcx = ((float)px + jx) / (float)sc.window[0] * 2.0f - 1.0f;
    float ndcy = ((float)py + jy) / (float)sc.window[1] * 2.0f - 1.0f;
    v3 direction = normalize(sc.forward + sc.right * (ndcx + sc.lens[0]) + sc.up * (-ndcy + sc.lens[1]));
    v3 origin = sc.origin;
    if (sc.projection == 1u) {
        // trace.wgsl:2546-2548: orthographic, rays along forward from the view rectangle
        direction = normalize(sc.forward);
        origin = origin + (normalize(sc.right) * ((ndcx + sc.lens[0]) * sc.half_view[0]) +
                           normalize(sc.up) * ((-ndcy + sc.lens[1]) * sc.half_view[1]));
    } else if (sc.projection == 2u) {
        // trace.wgsl:2549-2553: equirect, direction from longitude and latitude
        float phi = ((ndcx * 0.5f) * 2.0f) * PI;
        float theta = (ndcy * 0.5f + 0.5f) * PI;
        v3 local = mk(fm_sin(theta) * fm_sin(phi), fm_cos(theta), -fm_sin(theta) * fm_cos(phi));
        direction = normalize(normalize(sc.right) * local.x + normalize(sc.up) * local.y - normalize(sc.forward) * local.z);
    } else if (!centre && sc.lens[2] > 0.0f) {
        v3 focus = origin + direction * (sc.lens[3] / dot(direction, normalize(sc.forward)));
        state->at(1u);
        float radius = sqrtf(state->x()) * sc.lens[2];
        float phi = 2.0f * PI * state->y();
        origin = origin + (normalize(sc.right) * (radius * fm_cos(phi)) + normalize(sc.up) * (radius * fm_sin(phi)));
        direction = normalize(focus - origin);
    }
    *o = origin;
    *d = direction;
}

static inline void neumaier(float& s, float& c, float x) {
    float t = s + x;
    if (fabsf(s) >= fabsf(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

static void sample_pixel(Scene& sc, uint32_t x, uint32_t y, uint32_t count, uint32_t seed, uint64_t* closest,
                         uint64_t* shadows) {
    size_t index = (size_t)y * sc.width + x;
    float* S = &sc.sum[index * 12];
    float* C = &sc.comp[index * 12];
    double* D = &sc.sum64[index * 12];
    uint32_t acc = sc.samples;
    if (acc == 0) {
        for (int k = 0; k < 12; ++k) {
            S[k] = 0.0f;
            C[k] = 0.0f;
            D[k] = 0.0;
        }
        sc.m2[index] = 0.0f;
    }
    if (S[3] < 0.0f) return;
    uint32_t px = x + sc.window[2], py = y + sc.window[3];
    uint32_t stream = py * sc.window[0] + px;
    for (uint32_t sample = 0; sample < count; ++sample) {
        Rng state;
        state.state = pfx::pcg(stream * 1973u + pfx::pcg(seed * 9277u + (acc + sample) * 7919u + 26699u));
        if (!sc.sobol.empty()) {  // trace.wgsl:2522
            state.state = pfx::hash_iqnt2d(px, py) ^ seed;
            state.index = acc + sample;
            state.table = sc.sobol.data();
            state.length = sc.sobol_len;
        }
        state.at(0u);
        float jx = state.x();
        float jy = state.y();
        v3 o, d;
        camera_ray(sc, px, py, jx, jy, &o, &d, &state, false);
        Result r = radiance(sc, o, d, &state, closest, shadows);
        float l = (r.color.x + r.color.y + r.color.z) / 3.0f;
        sc.m2[index] = sc.m2[index] + l * l;
        float vals[12] = {r.color.x, r.color.y, r.color.z, 0.0f, r.albedo[0], r.albedo[1], r.albedo[2], r.albedo[3],
                          r.normal[0], r.normal[1], r.normal[2], r.normal[3]};
        for (int k = 0; k < 12; ++k) {
            if (k == 3) continue;
            if (sc.accumulate == ACC_PLAIN) {
                S[k] = S[k] + vals[k];
            } else if (sc.accumulate == ACC_F64) {
                D[k] = D[k] + (double)vals[k];
            } else {
                neumaier(S[k], C[k], vals[k]);
            }
        }
    }
    S[3] = (float)(acc + count);
    D[3] = (double)(acc + count);
}

static float total_of(const Scene& sc, size_t index, int k) {
    if (sc.accumulate == ACC_F64) return (float)sc.sum64[index * 12 + k];
    if (sc.accumulate == ACC_PLAIN) return sc.sum[index * 12 + k];
    return sc.sum[index * 12 + k] + sc.comp[index * 12 + k];
}

static void develop(const Scene& sc, float* color, float* albedo, float* normal) {
    for (size_t i = 0; i < (size_t)sc.width * sc.height; ++i) {
        float w = sc.sum[i * 12 + 3];
        float count = fmaxf(fabsf(w), 1.0f);
        if (sc.accumulate == ACC_F64) {
            double c64 = (double)count;
            if (color) {
                for (int k = 0; k < 3; ++k) color[i * 4 + k] = (float)(sc.sum64[i * 12 + k] / c64);
                color[i * 4 + 3] = 1.0f;
            }
            if (albedo)
                for (int k = 0; k < 4; ++k) albedo[i * 4 + k] = (float)(sc.sum64[i * 12 + 4 + k] / c64);
            v3 n = mk((float)(sc.sum64[i * 12 + 8] / c64), (float)(sc.sum64[i * 12 + 9] / c64),
                      (float)(sc.sum64[i * 12 + 10] / c64));
            v3 nn = pfx::length(n) > 1e-6f ? normalize(n) : mk(0, 0, 0);
            if (normal) {
                normal[i * 4 + 0] = nn.x;
                normal[i * 4 + 1] = nn.y;
                normal[i * 4 + 2] = nn.z;
                normal[i * 4 + 3] = (float)(sc.sum64[i * 12 + 11] / c64);
            }
            continue;
        }
        if (color) {
            for (int k = 0; k < 3; ++k) color[i * 4 + k] = total_of(sc, i, k) / count;
            color[i * 4 + 3] = 1.0f;
        }
        if (albedo)
            for (int k = 0; k < 4; ++k) albedo[i * 4 + k] = total_of(sc, i, 4 + k) / count;
        v3 n = mk(total_of(sc, i, 8) / count, total_of(sc, i, 9) / count, total_of(sc, i, 10) / count);
        v3 nn = pfx::length(n) > 1e-6f ? normalize(n) : mk(0, 0, 0);
        if (normal) {
            normal[i * 4 + 0] = nn.x;
            normal[i * 4 + 1] = nn.y;
            normal[i * 4 + 2] = nn.z;
            normal[i * 4 + 3] = total_of(sc, i, 11) / count;
        }
    }
}

static void sample_rows(Scene& sc, uint32_t count, uint32_t seed, uint32_t row0, uint32_t row1) {
    if (row1 == 0 || row1 > sc.height) row1 = sc.height;
    uint64_t closest = 0, shadows = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : closest, shadows)
    for (int y = (int)row0; y < (int)row1; ++y) {
        uint64_t c = 0, s = 0;
        for (uint32_t x = 0; x < sc.width; ++x) sample_pixel(sc, x, (uint32_t)y, count, seed, &c, &s);
        closest += c;
        shadows += s;
    }
    sc.closest += closest;
    sc.shadow += shadows;
    sc.paths += (uint64_t)(row1 - row0) * sc.width * count;
    sc.samples += count;
}

// adaptive.rs:375-385 (the same formula as converge.wgsl:19-27)
float adaptive_error(const float prior[4], float prior_count, const float now[4], float count) {
    if (prior_count <= 0.0f || count <= prior_count) return 1e9f;
    float scale = sqrtf(1.0f / count) / sqrtf(1.0f / prior_count + 1.0f / (count - prior_count));
    float gap[4];
    for (int k = 0; k < 4; ++k) {
        float rest = (now[k] * count - prior[k] * prior_count) / (count - prior_count);
        gap[k] = fabsf(prior[k] - rest) * scale;
    }
    return (gap[0] + gap[1] + gap[2]) / (1e-4f + sqrtf(fmaxf(now[0] + now[1] + now[2], 0.0f))) + gap[3];
}

// converge.wgsl:9-30 measure, :32-51 mark
static uint64_t converge(Scene& sc, const pfxh_adaptive& a, bool mark) {
    size_t pixels = (size_t)sc.width * sc.height;
    for (size_t i = 0; i < pixels; ++i) {
        float count = sc.sum[i * 12 + 3];
        float* snap = &sc.snaps[i * 8];
        if (count <= 0.0f) {
            snap[5] = 0.0f;
            continue;
        }
        float now[4] = {total_of(sc, i, 0) / count, total_of(sc, i, 1) / count, total_of(sc, i, 2) / count,
                        total_of(sc, i, 7) / count};
        float prior = snap[4];
        float error = adaptive_error(snap, prior, now, count);
        for (int k = 0; k < 4; ++k) snap[k] = now[k];
        snap[4] = count;
        snap[5] = error;
        snap[6] = 0.0f;
        snap[7] = 0.0f;
    }
    if (!mark) return pixels;
    uint64_t running = 0;
    int W = (int)sc.width, H = (int)sc.height;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            float count = sc.sum[i * 12 + 3];
            if (count <= 0.0f) continue;
            float worst = 0.0f;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = std::min(std::max(x + dx, 0), W - 1), yy = std::min(std::max(y + dy, 0), H - 1);
                    worst = fmaxf(worst, sc.snaps[((size_t)yy * W + xx) * 8 + 5]);
                }
            if (count >= (float)a.min_samples && worst < a.threshold) {
                sc.sum[i * 12 + 3] = -count;
            } else {
                running += 1;
            }
        }
    return running;
}

// ---- Sobol patterns (sobol.rs:40-150) ----------------------------------------
static const uint32_t XOR_TABLE[4][32] = {  // sobol.rs:41-70
    {0},
    {0x00000000, 0x00000001, 0x00000001, 0x00000007, 0x00000001, 0x00000013, 0x00000015, 0x0000007f,
     0x00000001, 0x00000103, 0x00000105, 0x0000070f, 0x00000111, 0x00001333, 0x00001555, 0x00007fff,
     0x00000001, 0x00010003, 0x00010005, 0x0007000f, 0x00010011, 0x00130033, 0x00150055, 0x007f00ff,
     0x00010101, 0x01030303, 0x01050505, 0x070f0f0f, 0x01111111, 0x13333333, 0x15555555, 0x7fffffff},
    {0x00000000, 0x00000001, 0x00000003, 0x00000001, 0x00000005, 0x0000001f, 0x0000002b, 0x0000003d,
     0x00000011, 0x00000133, 0x00000377, 0x00000199, 0x00000445, 0x00001ccf, 0x00002ddb, 0x0000366d,
     0x00000101, 0x00010303, 0x00030707, 0x00010909, 0x00051515, 0x001f3f3f, 0x002b6b6b, 0x003dbdbd,
     0x00101011, 0x01303033, 0x03707077, 0x01909099, 0x04515145, 0x1cf3f3cf, 0x2db6b6db, 0x36dbdb6d},
    {0x00000000, 0x00000001, 0x00000000, 0x00000003, 0x0000000d, 0x0000000c, 0x00000005, 0x0000004f,
     0x00000014, 0x000000e7, 0x00000329, 0x0000039c, 0x00000011, 0x00001033, 0x00000044, 0x000030bb,
     0x0000d1cd, 0x0000c2ec, 0x00005415, 0x0004fc3f, 0x00015054, 0x000e5c97, 0x0032e5b9, 0x0039725c,
     0x00000101, 0x01000303, 0x00000404, 0x03000b0b, 0x0d001d1d, 0x0c002c2c, 0x05004545, 0x4f00cfcf},
};
static uint32_t hash_hp_uint(uint32_t i) {  // sobol.rs:72-79
    i ^= i >> 16;
    i *= 0x21f0aaadu;
    i ^= i >> 15;
    i *= 0xd35a2d97u;
    i ^= i >> 15;
    return i ^ 0xe6fe3bebu;
}
static void generate_4d(float* points, uint32_t size, uint32_t seed) {  // sobol.rs:106-135
    uint32_t s = 0x44605a73u;
    s ^= s << 19;
    uint32_t rng = hash_hp_uint(seed ^ s);
    const float inv = 1.0f / 4294967808.0f;
    auto next = [&]() {
        float value = (float)hash_hp_uint(rng) * inv;
        rng += 1u;
        return value;
    };
    for (int a = 0; a < 4; ++a) points[a] = next();
    uint32_t log_n = 0, n = 1;
    while (n < size) {
        float strata = (float)(n * 2u);
        for (uint32_t i = 0; i < n && n + i < size; ++i)
            for (int axis = 0; axis < 4; ++axis) {
                uint32_t occupied = (uint32_t)(points[4 * (i ^ XOR_TABLE[axis][log_n]) + axis] * strata);
                points[4 * (n + i) + axis] = ((float)(occupied ^ 1u) + next()) / strata;
            }
        log_n += 1;
        n *= 2u;
    }
}
extern "C" {
void or_sobol_pattern(float* points, uint32_t size, uint32_t seed) { generate_4d(points, size, seed); }
void or_sobol_point(const float* table, uint32_t length, uint32_t seed, uint32_t index, uint32_t dim, float* out) {
    pfx::Sobol4 r = pfx::sobol_point(table, length, seed, index, dim);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}
}

static std::string prepare(Scene& sc, const pfxh_scene& in, const pfxh_detail& detail, uint32_t w, uint32_t h) {
    if (in.material_count == 0) return "invalid trace scene detail";
    for (uint32_t i = 0; i < in.light_count; ++i)
        if (!light_valid(in.lights[i])) return "invalid trace light";
    std::string e = lay_scene(in, sc.laid);
    if (!e.empty()) return e;
    for (uint32_t m = 0; m < in.material_count; ++m) {
        const pfxh_material& pm = in.materials[m];
        Mat mm;
        mm.base = t3(pm.base);
        mm.roughness = pm.roughness;
        mm.metalness = pm.metalness;
        mm.specular = pm.specular;
        mm.clearcoat = pm.clearcoat;
        mm.clearcoat_roughness = pm.clearcoat_roughness;
        mm.sheen = pm.sheen;
        mm.fresnel_power = pm.fresnel_power;
        mm.emission = t3(pm.emission);
        sc.materials.push_back(mm);
    }
    for (auto& t : sc.laid.tris)
        if ((uint32_t)t.a[3] >= in.material_count) return "invalid trace geometry";
    // gpu.rs:866-889: emissive world triangles become emitters, marked in c.w
    std::vector<EmitterCandidate> cands;
    uint32_t world_slots = in.triangle_count + in.world_detail_count;
    for (uint32_t slot = 0; slot < world_slots; ++slot) {
        Tri& t = sc.laid.tris[slot];
        const pfxh_material& pm = in.materials[(uint32_t)t.a[3]];
        float verts[9] = {t.a[0], t.a[1], t.a[2], t.b[0], t.b[1], t.b[2], t.c[0], t.c[1], t.c[2]};
        v3 e1 = t3(t.b) - t3(t.a), e2 = t3(t.c) - t3(t.a);
        v3 cr = pfx::cross(e1, e2);
        float area = 0.5f * sqrtf(dot(cr, cr));
        if (detail.no_emitter_sampling == 0u && emits(pm.emission) && area > 1e-12f) {  // gpu.rs:781
            EmitterCandidate c;
            memcpy(c.vertices, verts, sizeof verts);
            c.record = 0;
            memcpy(c.emission, pm.emission, 12);
            cands.push_back(c);
            t.c[3] = (float)cands.size();
        }
    }
    for (size_t slot = world_slots; slot < sc.laid.tris.size(); ++slot)
        if (detail.no_emitter_sampling == 0u && emits(in.materials[(uint32_t)sc.laid.tris[slot].a[3]].emission))
            return "emissive instanced meshes are not supported";
    sc.lights = tree_lay(in.lights, in.light_count);
    sc.emitters = build_emitters(cands, sc.lights.looped, sc.lights.tree);
    sc.lit = !cands.empty() || in.light_count > 0;
    for (uint32_t i = 0; i < sc.lights.looped; ++i) {
        uint32_t kind = (uint32_t)sc.lights.rows[i].a[3];
        if ((kind == 2u || kind == 3u) && sc.lights.rows[i].d[3] >= 0.5f) sc.area_lights = true;
    }
    sc.env.width = in.sky.width;
    sc.env.height = in.sky.height;
    if (in.sky.width > 0 && in.sky.height > 0)
        sc.env.texels.assign(in.sky.texels, in.sky.texels + (size_t)in.sky.width * in.sky.height * 4);
    sc.sky = sky_build(sc.env, 256, 128);
    sc.origin = t3(in.camera.origin);
    sc.forward = t3(in.camera.forward);
    sc.right = t3(in.camera.right);
    sc.up = t3(in.camera.up);
    sc.sun_dir = t3(in.sun.direction);
    sc.sun_color = t3(in.sun.color);
    sc.sun_intensity = in.sun.intensity;
    sc.sun_w = fm_cos(detail.sun_radius_deg * (3.14159265358979f / 180.0f));
    sc.lens[0] = detail.lens_shift[0];
    sc.lens[1] = detail.lens_shift[1];
    sc.lens[2] = detail.aperture * 0.5f;
    sc.lens[3] = detail.focus_distance;
    sc.projection = detail.projection;
    sc.half_view[0] = detail.projection == 1u ? detail.ortho_size[0] * 0.5f : 0.0f;  // gpu.rs:210-217
    sc.half_view[1] = detail.projection == 1u ? detail.ortho_size[1] * 0.5f : 0.0f;
    sc.width = w;
    sc.height = h;
    sc.window[0] = w;
    sc.window[1] = h;
    sc.bounces = detail.bounces;
    sc.accumulate = detail.accumulate;
    sc.watertight = detail.triangle_test == PFXH_TRI_WATERTIGHT;
    if (detail.sampler == PFXH_SAMPLER_SOBOL) {
        sc.sobol_len = pfx::sobol_length(detail.sobol_samples);
        sc.sobol.assign((size_t)sc.sobol_len * pfx::SOBOL_PATTERNS * 4, 0.0f);
        for (uint32_t p = 0; p < pfx::SOBOL_PATTERNS; ++p)
            generate_4d(&sc.sobol[4 * (size_t)p * sc.sobol_len], sc.sobol_len, p);
    }
    sc.sum.assign((size_t)w * h * 12, 0.0f);
    sc.comp.assign((size_t)w * h * 12, 0.0f);
    sc.sum64.assign((size_t)w * h * 12, 0.0);
    sc.m2.assign((size_t)w * h, 0.0f);
    sc.snaps.assign((size_t)w * h * 8, 0.0f);
    return "";
}

static void put_err(char* err, size_t n, const std::string& s) {
    if (!err || n == 0) return;
    size_t k = std::min(n - 1, s.size());
    memcpy(err, s.data(), k);
    err[k] = 0;
}

}  // namespace oracle

using namespace oracle;

extern "C" {

void* or_scene_new(const pfxh_scene* scene, const pfxh_detail* detail, uint32_t w, uint32_t h, char* err,
                   size_t err_len) {
    Scene* sc = new Scene();
    std::string e = prepare(*sc, *scene, *detail, w, h);
    if (!e.empty()) {
        put_err(err, err_len, e);
        delete sc;
        return nullptr;
    }
    return sc;
}
void or_scene_free(void* h) { delete (Scene*)h; }
void or_set_camera(void* h, const pfxh_camera* c) {
    Scene* sc = (Scene*)h;
    sc->origin = t3(c->origin);
    sc->forward = t3(c->forward);
    sc->right = t3(c->right);
    sc->up = t3(c->up);
    sc->samples = 0;
}
// Trace::set_sun_radius (gpu.rs:2903-2910), Trace::set_environment (:2810-2815)
void or_set_sun_radius(void* h, float radius_deg) {
    Scene* sc = (Scene*)h;
    sc->sun_w = fm_cos(radius_deg * (3.14159265358979f / 180.0f));
    sc->samples = 0;
}
void or_set_environment(void* h, const pfxh_environment* sky) {
    Scene* sc = (Scene*)h;
    sc->env.width = sky->width;
    sc->env.height = sky->height;
    sc->env.texels.clear();
    if (sky->width > 0 && sky->height > 0)
        sc->env.texels.assign(sky->texels, sky->texels + (size_t)sky->width * sky->height * 4);
    sc->sky = sky_build(sc->env, 256, 128);
    sc->samples = 0;
}
int or_set_window(void* h, uint32_t fw, uint32_t fh, uint32_t x, uint32_t y) {
    Scene* sc = (Scene*)h;
    if ((uint64_t)x + sc->width > fw || (uint64_t)y + sc->height > fh) return 1;
    sc->window[0] = fw;
    sc->window[1] = fh;
    sc->window[2] = x;
    sc->window[3] = y;
    sc->samples = 0;
    return 0;
}
void or_sample(void* h, uint32_t count, uint32_t seed, uint32_t row0, uint32_t row1) {
    sample_rows(*(Scene*)h, count, seed, row0, row1);
}
uint32_t or_samples(void* h) { return ((Scene*)h)->samples; }
void or_readback(void* h, float* color, float* albedo, float* normal) { develop(*(Scene*)h, color, albedo, normal); }
// the standard error of each pixel's mean luminance, sqrt(max(m2 / n - l^2,
// 0) / n) with l the mean of the developed colour's channels; the product's
// noise channel for the learned denoiser (no reference counterpart)
void or_deviation(void* h, float* out) {
    const Scene& sc = *(const Scene*)h;
    for (size_t i = 0; i < (size_t)sc.width * sc.height; ++i) {
        float n = fabsf(sc.sum[i * 12 + 3]);
        float se = 0.0f;
        if (n > 0.0f) {
            float l = (total_of(sc, i, 0) / n + total_of(sc, i, 1) / n + total_of(sc, i, 2) / n) / 3.0f;
            se = sqrtf(fmaxf(sc.m2[i] / n - l * l, 0.0f) / n);
        }
        out[i] = se;
    }
}
void or_counters(void* h, uint64_t* out) {
    Scene* sc = (Scene*)h;
    out[0] = sc->closest;
    out[1] = sc->shadow;
    out[2] = sc->paths;
}
// probe.wgsl:3-39 plus the caller's primitive index
void or_probe(void* h, pfxh_probe_hit* out) {
    Scene* sc = (Scene*)h;
    for (uint32_t y = 0; y < sc->height; ++y)
        for (uint32_t x = 0; x < sc->width; ++x) {
            v3 o, d;
            camera_ray(*sc, x + sc->window[2], y + sc->window[3], 0.5f, 0.5f, &o, &d, nullptr, true);
            Hit hit = scene_hit(*sc, o, d, 1e20f, false);
            pfxh_probe_hit& r = out[(size_t)y * sc->width + x];
            memset(&r, 0, sizeof r);
            r.primitive = 0xffffffffu;
            if (hit.t >= 1e19f) continue;
            v3 n = dot(hit.n, d) > 0.0f ? -hit.n : hit.n;
            r.depth = sc->projection == 2u ? hit.t : hit.t * dot(d, normalize(sc->forward));  // probe.wgsl:33-34
            r.distance = hit.t;
            r.material_id = hit.material + 1u;
            r.primitive = sc->laid.slot_prim[hit.surface];
            r.instance = sc->laid.instance_of_record[hit.record];
            r.normal[0] = n.x;
            r.normal[1] = n.y;
            r.normal[2] = n.z;
        }
}
// adaptive.rs:136-324 without checkpoints
int or_sample_adaptive(void* h, const pfxh_adaptive* a, uint32_t seed, pfxh_converged* out) {
    Scene* sc = (Scene*)h;
    if (!(std::isfinite(a->threshold) && a->threshold > 0.0f && a->min_samples >= 2 &&
          a->max_samples >= a->min_samples && std::isfinite(a->growth) && a->growth > 1.0f))
        return 1;
    sc->samples = 0;
    std::fill(sc->snaps.begin(), sc->snaps.end(), 0.0f);
    std::vector<uint32_t> rounds = adaptive_rounds(*a);
    memset(out, 0, sizeof *out);
    out->pixels = (uint64_t)sc->width * sc->height;
    for (size_t step = 0; step < rounds.size(); ++step) {
        uint32_t target = rounds[step];
        uint32_t count = target > sc->samples ? target - sc->samples : 0;
        sample_rows(*sc, count, seed, 0, 0);
        uint64_t running = converge(*sc, *a, step > 0);
        if (out->round_count < PFXH_MAX_ROUNDS) {
            out->samples[out->round_count] = target;
            out->active[out->round_count] = running;
            out->round_count += 1;
        }
        if (running == 0) break;
    }
    return 0;
}
float or_adaptive_error(const float* prior, float prior_count, const float* now, float count) {
    return adaptive_error(prior, prior_count, now, count);
}
void or_tree_pick(void* h, const float* p, const float* n, float u, float* out) {
    F2 r = tree_pick(*(Scene*)h, t3(p), t3(n), u);
    out[0] = r.x;
    out[1] = r.y;
}
// tree_pick at `count` uniforms u[k] for one shading point: out[2k] the
// picked light row (as a float, as tree_pick returns it), out[2k+1] its pdf.
void or_tree_pick_many(void* h, const float* p, const float* n, const float* u, uint32_t count, float* out) {
    const Scene& sc = *(Scene*)h;
    v3 pp = t3(p), nn = t3(n);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)count; ++k) {
        F2 r = tree_pick(sc, pp, nn, u[k]);
        out[2 * k] = r.x;
        out[2 * k + 1] = r.y;
    }
}
uint32_t or_light_meta(void* h, uint32_t* looped, uint32_t* tree) {
    Scene* sc = (Scene*)h;
    *looped = sc->lights.looped;
    *tree = sc->lights.tree;
    return (uint32_t)sc->lights.rows.size();
}
// Instrumented traversal counts for one closest-hit query (algorithmic bytes)
// Closest hits of n rays (origins and directions, 3 floats each) as
// caller triangle ids (UINT32_MAX on a miss) and distances; with `any`,
// occlusion before `limit` instead (id 1 blocked, 0 clear).
void or_cast(void* h, const float* o, const float* d, uint32_t n, int any, float limit, uint32_t* ids, float* t) {
    Scene* sc = (Scene*)h;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        v3 ro = t3(o + 3 * i), rd = t3(d + 3 * i);
        if (any) {
            ids[i] = occluded(*sc, ro, rd, limit) ? 1u : 0u;
            t[i] = 0.0f;
            continue;
        }
        Hit hit = scene_hit(*sc, ro, rd, limit, false);
        bool missed = hit.t >= limit;
        ids[i] = missed ? 0xffffffffu : sc->laid.slot_prim[hit.surface];
        t[i] = missed ? 0.0f : hit.t;
    }
}
void or_trace_counts(void* h, const float* o, const float* d, uint64_t* out) {
    Counter c;
    scene_hit(*(Scene*)h, t3(o), t3(d), 1e20f, false, &c);
    out[0] = c.nodes;
    out[1] = c.prims;
}

}  // extern "C"
