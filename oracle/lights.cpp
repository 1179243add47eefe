// Oracle (test infrastructure only): local lights, the light tree, the
// emitter CDF, the sky CDF and adaptive rounds, restating
// crates/trace/src/lights.rs, tree.rs, sky.rs and adaptive.rs.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "oracle.h"
#include "select_nth.h"

namespace oracle {

using pfx::fm_acos;
using pfx::fm_cos;
using pfx::fm_sin;

static const float PI = 3.14159265358979f;
static const float FRAC_PI_2 = 1.57079632679490f;
static const float TAU_ = 6.28318530717959f;

// math.rs:29-36 (a zero vector maps to +z)
static v3 rnormalize(v3 a) {
    float len = sqrtf(pfx::dot(a, a));
    if (len <= 1e-12f) return pfx::mk(0.0f, 0.0f, 1.0f);
    return a * (1.0f / len);
}
static v3 load3(const float* p) { return pfx::mk(p[0], p[1], p[2]); }
// f32::to_radians
static float to_radians(float deg) { return deg * (PI / 180.0f); }

// lights.rs:289-291
float luminance(const float v[3]) { return 0.2126f * v[0] + 0.7152f * v[1] + 0.0722f * v[2]; }
// lights.rs:293-298
bool emits(const float e[3]) {
    for (int k = 0; k < 3; ++k)
        if (!(std::isfinite(e[k]) && e[k] >= 0.0f)) return false;
    return luminance(e) > 0.0f;
}
// lights.rs:97-99
bool light_lit(const pfxh_light& l) {
    return l.intensity > 0.0f && l.range > 0.0f && (l.colour[0] > 0.0f || l.colour[1] > 0.0f || l.colour[2] > 0.0f);
}
// lights.rs:101-142 (point and spot)
bool light_valid(const pfxh_light& l) {
    const float* scalars[] = {l.position, l.position + 1, l.position + 2, l.colour, l.colour + 1, l.colour + 2,
                              &l.intensity, &l.radius, &l.range};
    for (auto s : scalars)
        if (!std::isfinite(*s)) return false;
    for (int k = 0; k < 3; ++k)
        if (l.colour[k] < 0.0f) return false;
    if (l.intensity < 0.0f || l.radius < 0.0f || l.range < 0.0f) return false;
    if (l.kind == PFXH_LIGHT_POINT) return true;
    if (l.kind == PFXH_LIGHT_RECT) {  // lights.rs:128-135
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(l.half_u[k]) || !std::isfinite(l.half_v[k])) return false;
        float lu = pfx::length(load3(l.half_u)), lv = pfx::length(load3(l.half_v));
        return lu > 1e-6f && lv > 1e-6f && fabsf(pfx::dot(load3(l.half_u), load3(l.half_v))) <= 1e-4f * lu * lv;
    }
    if (l.kind == PFXH_LIGHT_DISC) {  // lights.rs:136-141
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(l.direction[k])) return false;
        return pfx::length(load3(l.direction)) > 1e-6f && std::isfinite(l.disc_radius) && l.disc_radius > 1e-6f;
    }
    if (l.kind != PFXH_LIGHT_SPOT) return false;
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(l.direction[k])) return false;
    return pfx::length(load3(l.direction)) > 1e-6f && std::isfinite(l.inner_deg) && std::isfinite(l.outer_deg) &&
           0.0f <= l.inner_deg && l.inner_deg <= l.outer_deg && l.outer_deg > 0.0f && l.outer_deg <= 90.0f;
}
// lights.rs:233-238
static void spot_terms(float inner_deg, float outer_deg, float* scale, float* offset) {
    float inner = fm_cos(to_radians(inner_deg));
    float outer = fm_cos(to_radians(outer_deg));
    *scale = 1.0f / pfx::maxf(inner - outer, 0.001f);
    *offset = -outer * *scale;
}
// lights.rs:154-163
float light_falloff(const pfxh_light& l, float distance) {
    if (distance >= l.range || distance <= 1e-6f) return 0.0f;
    float ratio = distance / l.range;
    float r2 = ratio * ratio;
    float window = pfx::fm_clamp(1.0f - r2 * r2, 0.0f, 1.0f);
    float reach = pfx::maxf(distance, l.radius);
    return l.intensity * window * window / (reach * reach);
}
// lights.rs:165-181
float light_cone(const pfxh_light& l, const float toward[3]) {
    if (l.kind != PFXH_LIGHT_SPOT) return 1.0f;
    float scale, offset;
    spot_terms(l.inner_deg, l.outer_deg, &scale, &offset);
    v3 axis = rnormalize(load3(l.direction));
    float cd = pfx::dot(rnormalize(load3(toward)), axis);
    float t = pfx::fm_clamp(cd * scale + offset, 0.0f, 1.0f);
    return t * t;
}
// lights.rs:146-152
float light_area(const pfxh_light& l) {
    if (l.kind == PFXH_LIGHT_RECT) return 4.0f * pfx::length(pfx::cross(load3(l.half_u), load3(l.half_v)));
    if (l.kind == PFXH_LIGHT_DISC) return PI * l.disc_radius * l.disc_radius;
    return 0.0f;
}
// lights.rs:183-231
GLight light_pack(const pfxh_light& l) {
    GLight g;
    memset(&g, 0, sizeof g);
    float shadow = l.shadow ? 1.0f : 0.0f;
    if (l.kind == PFXH_LIGHT_RECT || l.kind == PFXH_LIGHT_DISC) {  // :207-230
        float area = light_area(l);
        float sided = l.two_sided ? 1.0f : 0.0f;
        for (int k = 0; k < 3; ++k) {
            g.a[k] = l.position[k];
            g.b[k] = (l.colour[k] * l.intensity) / area;
        }
        g.b[3] = FLT_MAX;
        g.c[3] = sided;
        g.d[3] = shadow;
        if (l.kind == PFXH_LIGHT_RECT) {
            g.a[3] = 2.0f;
            for (int k = 0; k < 3; ++k) {
                g.c[k] = l.half_u[k];
                g.d[k] = l.half_v[k];
            }
        } else {
            g.a[3] = 3.0f;
            v3 nrm = rnormalize(load3(l.direction));
            g.c[0] = nrm.x;
            g.c[1] = nrm.y;
            g.c[2] = nrm.z;
            g.d[0] = l.disc_radius;
        }
        return g;
    }
    for (int k = 0; k < 3; ++k) {
        g.a[k] = l.position[k];
        g.b[k] = l.colour[k] * l.intensity;
    }
    g.b[3] = l.range;
    g.c[3] = l.radius;
    g.d[3] = shadow;
    if (l.kind == PFXH_LIGHT_SPOT) {
        float scale, offset;
        spot_terms(l.inner_deg, l.outer_deg, &scale, &offset);
        v3 axis = rnormalize(load3(l.direction));
        g.a[3] = 1.0f;
        g.c[0] = axis.x;
        g.c[1] = axis.y;
        g.c[2] = axis.z;
        g.d[0] = scale;
        g.d[1] = offset;
    }
    return g;
}

// ---- tree.rs ----------------------------------------------------------------
struct Cone {
    v3 axis;
    float theta_o, theta_e;
};
static const Cone CONE_EMPTY = {{0.0f, 0.0f, 0.0f}, FLT_MIN, FLT_MIN};
static bool cone_empty(const Cone& c) { return c.axis.x == 0.0f && c.axis.y == 0.0f && c.axis.z == 0.0f; }
// tree.rs:28-41
static float cone_measure(const Cone& c) {
    if (cone_empty(c)) return 0.0f;
    float theta_w = pfx::minf(c.theta_o + c.theta_e, PI);
    float cos_o = fm_cos(c.theta_o), sin_o = fm_sin(c.theta_o);
    return TAU_ * (1.0f - cos_o) +
           FRAC_PI_2 * (2.0f * theta_w * sin_o - fm_cos(c.theta_o - 2.0f * theta_w) - 2.0f * c.theta_o * sin_o + cos_o);
}
// tree.rs:84-91
static v3 orthogonal(v3 n) {
    v3 helper = fabsf(n.x) > 0.1f ? pfx::mk(0.0f, 1.0f, 0.0f) : pfx::mk(1.0f, 0.0f, 0.0f);
    return rnormalize(pfx::cross(n, helper));
}
// tree.rs:43-81
static Cone cone_merge(Cone self, Cone other) {
    if (cone_empty(self)) return other;
    if (cone_empty(other)) return self;
    Cone a = self, b = other;
    if (other.theta_o > self.theta_o) {
        a = other;
        b = self;
    }
    float cos_a_b = pfx::dot(a.axis, b.axis);
    float theta_d = fm_acos(pfx::fm_clamp(cos_a_b, -1.0f, 1.0f));
    float theta_e = pfx::maxf(a.theta_e, b.theta_e);
    if (a.theta_o + 5e-4f >= pfx::minf(PI, theta_d + b.theta_o)) return Cone{a.axis, a.theta_o, theta_e};
    float theta_o = (theta_d + a.theta_o + b.theta_o) * 0.5f;
    if (theta_o >= PI) return Cone{a.axis, PI, theta_e};
    v3 axis;
    if (cos_a_b < -0.9995f) {
        axis = orthogonal(a.axis);
    } else {
        float theta_r = theta_o - a.theta_o;
        v3 ortho = rnormalize(b.axis - a.axis * cos_a_b);
        axis = a.axis * fm_cos(theta_r) + ortho * fm_sin(theta_r);
    }
    return Cone{axis, theta_o, theta_e};
}
struct Measure {
    float min[3], max[3];
    Cone cone;
    float energy, range;
};
static Measure measure_empty() {
    Measure m;
    for (int k = 0; k < 3; ++k) {
        m.min[k] = FLT_MAX;
        m.max[k] = -FLT_MAX;
    }
    m.cone = CONE_EMPTY;
    m.energy = 0.0f;
    m.range = 0.0f;
    return m;
}
// tree.rs:111-122
static void measure_add(Measure& self, const Measure& other) {
    if (other.energy == 0.0f) return;
    for (int k = 0; k < 3; ++k) {
        self.min[k] = pfx::minf(self.min[k], other.min[k]);
        self.max[k] = pfx::maxf(self.max[k], other.max[k]);
    }
    self.cone = cone_merge(self.cone, other.cone);
    self.energy += other.energy;
    self.range = pfx::maxf(self.range, other.range);
}
// tree.rs:124-132
static float measure_calculate(const Measure& m) {
    if (m.energy == 0.0f) return 0.0f;
    float s0 = m.max[0] - m.min[0], s1 = m.max[1] - m.min[1], s2 = m.max[2] - m.min[2];
    float area = 2.0f * (s0 * s1 + s1 * s2 + s2 * s0);
    float area_measure = area == 0.0f ? sqrtf(s0 * s0 + s1 * s1 + s2 * s2) : area;
    return m.energy * area_measure * cone_measure(m.cone);
}
struct Emitter {
    size_t index;
    float centroid[3];
    Measure measure;
};
// tree.rs:142-175
static Emitter emitter_of(size_t index, const pfxh_light& l) {
    Emitter e;
    e.index = index;
    Cone cone;
    if (l.kind == PFXH_LIGHT_SPOT) {
        cone = Cone{rnormalize(load3(l.direction)), 0.0f, pfx::minf(to_radians(l.outer_deg), FRAC_PI_2)};
    } else {
        cone = Cone{pfx::mk(0.0f, 0.0f, -1.0f), PI, FRAC_PI_2};
    }
    float s[3];
    for (int k = 0; k < 3; ++k) {
        e.centroid[k] = l.position[k];
        s[k] = fabsf(l.colour[k] * l.intensity);
        e.measure.min[k] = l.position[k] - l.radius;
        e.measure.max[k] = l.position[k] + l.radius;
    }
    e.measure.cone = cone;
    e.measure.energy = (s[0] + s[1] + s[2]) / 3.0f;
    e.measure.range = l.range;
    return e;
}
struct Bucket {
    bool has;
    Measure measure;
    size_t count;
};
static Bucket bucket_default() { return Bucket{false, measure_empty(), 0}; }
// tree.rs:183-203
static void bucket_add(Bucket& b, const Measure& m) {
    if (!b.has) {
        b.has = true;
        b.measure = measure_empty();
    }
    measure_add(b.measure, m);
    b.count += 1;
}
static Bucket bucket_plus(const Bucket& self, const Bucket& other) {
    Measure m = self.has ? self.measure : measure_empty();
    if (other.has) measure_add(m, other.measure);
    return Bucket{true, m, self.count + other.count};
}
static float bucket_cost(const Bucket& b) { return b.has ? measure_calculate(b.measure) : 0.0f; }

static const size_t LOOPED = 4, LEAF = 8, BUCKETS = 12;
// tree.rs:205-207
static bool treed(const pfxh_light& l) { return light_lit(l) && (l.kind == PFXH_LIGHT_POINT || l.kind == PFXH_LIGHT_SPOT); }

// tree.rs:319-395; returns (split, middle, dim or -1)
static Measure should_split(const Emitter* emitters, size_t count, bool* split, size_t* middle_out, int* dim_out) {
    *dim_out = -1;
    if (count < 2) {
        *split = false;
        return emitters[0].measure;
    }
    size_t middle = count / 2;
    float low[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, high[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (size_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            low[k] = pfx::minf(low[k], emitters[i].centroid[k]);
            high[k] = pfx::maxf(high[k], emitters[i].centroid[k]);
        }
    float extent[3] = {high[0] - low[0], high[1] - low[1], high[2] - low[2]};
    float max_extent = pfx::maxf(pfx::maxf(pfx::maxf(extent[0], extent[1]), extent[2]), 0.0f);
    Measure measure = measure_empty();
    float total_cost = 0.0f, min_cost = FLT_MAX;
    int split_dim = -1;
    for (int dim = 0; dim < 3; ++dim) {
        Bucket buckets[BUCKETS];
        for (auto& b : buckets) b = bucket_default();
        float inv_extent;
        if (extent[dim] == 0.0f) {
            if (dim != 0) continue;
            inv_extent = FLT_MAX;
            for (size_t i = 0; i < count; ++i) bucket_add(buckets[0], emitters[i].measure);
        } else {
            inv_extent = 1.0f / extent[dim];
            for (size_t i = 0; i < count; ++i) {
                float f = (float)BUCKETS * (emitters[i].centroid[dim] - low[dim]) * inv_extent;
                long long at = (f != f) ? 0 : (long long)f;  // `as i64` saturates, NaN -> 0
                if (at < 0) at = 0;
                if (at > (long long)BUCKETS - 1) at = BUCKETS - 1;
                bucket_add(buckets[at], emitters[i].measure);
            }
        }
        Bucket left[BUCKETS - 1];
        left[0] = buckets[0];
        for (size_t i = 1; i < BUCKETS - 1; ++i) left[i] = bucket_plus(left[i - 1], buckets[i]);
        if (dim == 0) {
            Bucket all = bucket_plus(left[BUCKETS - 2], buckets[BUCKETS - 1]);
            measure = all.has ? all.measure : measure_empty();
            if (extent[0] == 0.0f && extent[1] == 0.0f && extent[2] == 0.0f) break;
            total_cost = measure_calculate(measure);
            if (total_cost == 0.0f) break;
            if (extent[dim] == 0.0f) continue;
        }
        Bucket right[BUCKETS - 1];
        right[BUCKETS - 2] = buckets[BUCKETS - 1];
        for (int i = (int)BUCKETS - 3; i >= 0; --i) right[i] = bucket_plus(right[i + 1], buckets[i + 1]);
        float regularization = max_extent * inv_extent;
        for (size_t s = 0; s < BUCKETS - 1; ++s) {
            float cost = regularization * (bucket_cost(left[s]) + bucket_cost(right[s]));
            if (cost < total_cost && cost < min_cost && left[s].count >= 1 && left[s].count < count) {
                min_cost = cost;
                split_dim = dim;
                middle = left[s].count;
            }
        }
    }
    *split = min_cost < total_cost || count > LEAF;
    *middle_out = middle;
    *dim_out = split_dim;
    return measure;
}
// tree.rs:295-317
static GLight tree_row(const Measure& m, size_t count, size_t link) {
    GLight g;
    for (int k = 0; k < 3; ++k) {
        g.a[k] = m.min[k];
        g.b[k] = m.max[k];
    }
    g.a[3] = m.energy;
    g.b[3] = m.range;
    g.c[0] = m.cone.axis.x;
    g.c[1] = m.cone.axis.y;
    g.c[2] = m.cone.axis.z;
    g.c[3] = m.cone.theta_o;
    g.d[0] = m.cone.theta_e;
    g.d[1] = (float)count;
    g.d[2] = (float)link;
    g.d[3] = 0.0f;
    return g;
}
// f32::total_cmp(a, b) == Less: IEEE 754 totalOrder on the bit patterns
// (negative values have all but the sign bit flipped, then a signed compare)
static bool total_less(float a, float b) {
    int32_t x, y;
    memcpy(&x, &a, 4);
    memcpy(&y, &b, 4);
    x ^= (int32_t)((uint32_t)(x >> 31) >> 1);
    y ^= (int32_t)((uint32_t)(y >> 31) >> 1);
    return x < y;
}
// When set, the split sorts the whole slice stably by centroid instead:
// another valid selection, which tests use to show that the leaf order
// depends only on which emitters fall on each side, not on the pivots.
static bool g_split_by_sort = false;
void set_tree_split_by_sort(bool on) { g_split_by_sort = on; }

// tree.rs:263-293; the split around `middle` is Rust's
// select_nth_unstable_by on the centroid (select_nth.h), so each side keeps
// the order that selection leaves, which becomes the leaf order.
static void grow(Emitter* emitters, size_t count, size_t start, size_t base, size_t root, std::vector<GLight>& nodes) {
    size_t at = nodes.size();
    nodes.push_back(GLight{});
    bool split;
    size_t middle;
    int dim;
    Measure measure = should_split(emitters, count, &split, &middle, &dim);
    if (split) {
        if (dim >= 0) {
            auto less = [dim](const Emitter& l, const Emitter& r) {
                return total_less(l.centroid[dim], r.centroid[dim]);
            };
            if (g_split_by_sort)
                std::stable_sort(emitters, emitters + count, less);
            else
                rsel::select_nth_unstable_by(emitters, count, middle, less);
        }
        grow(emitters, middle, start, base, root, nodes);
        size_t right_at = root + nodes.size();
        grow(emitters + middle, count - middle, start + middle, base, root, nodes);
        nodes[at] = tree_row(measure, 0, right_at);
    } else {
        nodes[at] = tree_row(measure, count, base + start);
    }
}
// tree.rs:220-261
LaidLights tree_lay(const pfxh_light* lights, size_t n) {
    LaidLights out;
    size_t treed_count = 0;
    for (size_t i = 0; i < n; ++i) treed_count += treed(lights[i]) ? 1 : 0;
    if (treed_count <= LOOPED) {
        for (size_t i = 0; i < n; ++i) out.rows.push_back(light_pack(lights[i]));
        out.looped = (uint32_t)n;
        out.tree = 0;
        return out;
    }
    for (size_t i = 0; i < n; ++i)
        if (!treed(lights[i])) out.rows.push_back(light_pack(lights[i]));
    size_t looped = out.rows.size();
    std::vector<Emitter> emitters;
    for (size_t i = 0; i < n; ++i)
        if (treed(lights[i])) emitters.push_back(emitter_of(i, lights[i]));
    std::vector<GLight> nodes;
    size_t count = emitters.size();
    if (count > 0) grow(emitters.data(), count, 0, looped, looped + count, nodes);
    for (auto& e : emitters) out.rows.push_back(light_pack(lights[e.index]));
    out.rows.insert(out.rows.end(), nodes.begin(), nodes.end());
    out.looped = (uint32_t)looped;
    out.tree = (uint32_t)treed_count;
    return out;
}

// lights.rs:300-305
static float triangle_area(const float v[9]) {
    v3 e1 = load3(v + 3) - load3(v), e2 = load3(v + 6) - load3(v);
    v3 c = pfx::cross(e1, e2);
    return 0.5f * sqrtf(pfx::dot(c, c));
}
// lights.rs:391-440
EmitterList build_emitters(const std::vector<EmitterCandidate>& cs, size_t looped, size_t tree) {
    std::vector<double> powers;
    double total = -0.0;
    for (auto& c : cs) {
        double p = (double)triangle_area(c.vertices) * (double)luminance(c.emission);
        powers.push_back(p);
        total = total + p;
    }
    EmitterList list;
    GEmitter head;
    memset(&head, 0, sizeof head);
    head.a[0] = (float)cs.size();
    head.a[1] = (float)looped;
    head.a[2] = (float)total;
    head.a[3] = (float)tree;
    list.entries.push_back(head);
    double running = 0.0;
    for (size_t i = 0; i < cs.size(); ++i) {
        running += powers[i];
        list.cdf.push_back((float)(running / total));
        float probability = (float)(powers[i] / total);
        GEmitter e;
        for (int k = 0; k < 3; ++k) {
            e.a[k] = cs[i].vertices[k];
            e.b[k] = cs[i].vertices[3 + k];
            e.c[k] = cs[i].vertices[6 + k];
            e.e[k] = cs[i].emission[k];
        }
        e.a[3] = 0.0f;
        e.b[3] = (float)cs[i].record;
        e.c[3] = triangle_area(cs[i].vertices);
        e.e[3] = probability;
        list.entries.push_back(e);
    }
    if (!list.cdf.empty()) list.cdf.back() = 1.0f;
    return list;
}

// ---- sky.rs -----------------------------------------------------------------
// sky_radiance for an HDR lat-long environment (core sky.rs SKY_MODEL_WGSL
// `sky_radiance`, textureSampleLevel with repeat-u / clamp-v bilinear).
v3 env_radiance(const Env& env, v3 direction) {
    if (env.width == 0 || env.height == 0) return pfx::mk(0.0f, 0.0f, 0.0f);
    v3 d = pfx::normalize(direction);
    float u = pfx::fm_atan2(d.x, -d.z) / 6.28318531f + 0.5f;
    u = u - floorf(u);
    float v = fm_acos(pfx::fm_clamp(d.y, -1.0f, 1.0f)) / 3.14159265f;
    float x = u * (float)env.width - 0.5f, y = v * (float)env.height - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float tx = x - x0, ty = y - y0;
    auto at = [&](float ix, float iy, int k) {
        int w = (int)env.width, h = (int)env.height;
        int col = ((int)ix % w + w) % w;
        int row = (int)iy;
        row = row < 0 ? 0 : (row > h - 1 ? h - 1 : row);
        return env.texels[((size_t)row * env.width + col) * 4 + k];
    };
    float out[3];
    for (int k = 0; k < 3; ++k) {
        float a = at(x0, y0, k), b = at(x0 + 1.0f, y0, k), c = at(x0, y0 + 1.0f, k), dd = at(x0 + 1.0f, y0 + 1.0f, k);
        out[k] = (a * (1.0f - tx) + b * tx) * (1.0f - ty) + (c * (1.0f - tx) + dd * tx) * ty;
    }
    return pfx::mk(out[0], out[1], out[2]);
}
// sky.rs:207-272 (no sun: the trace calls `.without_sun()`, gpu.rs:1795)
SkyCdf sky_build(const Env& env, uint32_t width, uint32_t height) {
    SkyCdf cdf;
    width = width < 1 ? 1 : width;
    height = height < 1 ? 1 : height;
    cdf.width = width;
    cdf.height = height;
    cdf.rows.assign(height, 0.0f);
    cdf.columns.assign((size_t)width * height, 0.0f);
    double total = 0.0;
    for (uint32_t row = 0; row < height; ++row) {
        float theta = PI * ((float)row + 0.5f) / (float)height;
        float sine = fm_sin(theta);
        double row_total = 0.0;
        for (uint32_t col = 0; col < width; ++col) {
            float phi = TAU_ * ((float)col + 0.5f) / (float)width;
            v3 d = pfx::mk(fm_sin(theta) * fm_cos(phi), fm_cos(theta), fm_sin(theta) * fm_sin(phi));
            v3 color = env_radiance(env, d);
            float c3[3] = {color.x, color.y, color.z};
            row_total += (double)pfx::maxf(luminance(c3), 0.0f) * (double)sine;
            cdf.columns[(size_t)row * width + col] = (float)row_total;
        }
        if (row_total > 0.0) {
            for (uint32_t col = 0; col < width; ++col) cdf.columns[(size_t)row * width + col] /= (float)row_total;
        } else {
            for (uint32_t col = 0; col < width; ++col)
                cdf.columns[(size_t)row * width + col] = (float)(col + 1) / (float)width;
        }
        total += row_total;
        cdf.rows[row] = (float)total;
    }
    if (total > 0.0) {
        for (auto& v : cdf.rows) v /= (float)total;
    } else {
        for (uint32_t i = 0; i < height; ++i) cdf.rows[i] = (float)(i + 1) / (float)height;
    }
    cdf.total = (float)total;
    return cdf;
}

// adaptive.rs:46-63
std::vector<uint32_t> adaptive_rounds(const pfxh_adaptive& a) {
    std::vector<uint32_t> rounds{std::max<uint32_t>(a.min_samples / 2, 1)};
    for (;;) {
        uint32_t last = rounds.back();
        if (last >= a.max_samples) return rounds;
        float grown = ceilf((float)last * a.growth);
        uint32_t next = grown >= 4294967296.0f ? UINT32_MAX : (uint32_t)grown;
        next = std::max(next, last + 1);
        next = std::max(next, rounds.size() == 1 ? a.min_samples : 0u);
        next = std::min(next, a.max_samples);
        rounds.push_back(next);
    }
}

}  // namespace oracle
