// Oracle (test infrastructure only): C entry points used by the tests to
// reach the restated layout functions and the shared math.
#include <string.h>

#include "oracle.h"
#include "select_nth.h"

using namespace oracle;

extern "C" {

uint32_t or_pcg(uint32_t v) { return pfx::pcg(v); }
float or_offset_axis(float position, float normal) { return pfx::offset_axis(position, normal); }
void or_spawn_point(const float* p, const float* dir, const float* origin, float t, float* out) {
    pfx::v3 r = pfx::spawn_point(pfx::mk(p[0], p[1], p[2]), pfx::mk(dir[0], dir[1], dir[2]),
                                 pfx::mk(origin[0], origin[1], origin[2]), t);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
// 0 sin, 1 cos, 2 acos, 3 atan2, 4 exp2, 5 log2, 6 pow, 7 asin, 8 cos(max(0, acos x - acos y))
float or_fm(int which, float x, float y) {
    switch (which) {
        case 8: return pfx::fm_cos_acos_gap(x, y);
        case 0: return pfx::fm_sin(x);
        case 1: return pfx::fm_cos(x);
        case 2: return pfx::fm_acos(x);
        case 3: return pfx::fm_atan2(x, y);
        case 4: return pfx::fm_exp2(x);
        case 5: return pfx::fm_log2(x);
        case 6: return pfx::fm_pow(x, y);
        default: return pfx::fm_asin(x);
    }
}
void or_fm_many(int which, const float* x, const float* y, float* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = or_fm(which, x[i], y ? y[i] : 0.0f);
}

// bvh.rs:95-105
uint32_t or_bvh(const pfxh_triangle* tris, uint32_t count, pfxh_bvh_node* nodes, uint32_t* order) {
    std::vector<float> v(count * 9);
    for (uint32_t i = 0; i < count; ++i) memcpy(&v[9 * i], tris[i].vertices, 36);
    Bvh bvh = bvh_build(v.data(), count);
    memcpy(nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(Node));
    memcpy(order, bvh.order.data(), bvh.order.size() * 4);
    return (uint32_t)bvh.nodes.size();
}
// bvh.rs:135-164 and :294-324
uint32_t or_bvh_intersect(const pfxh_triangle* tris, uint32_t count, const pfxh_bvh_node* nodes, uint32_t node_count,
                          const uint32_t* order, const float* o, const float* d) {
    std::vector<float> v(count * 9);
    for (uint32_t i = 0; i < count; ++i) memcpy(&v[9 * i], tris[i].vertices, 36);
    Bvh bvh;
    bvh.nodes.resize(node_count);
    memcpy(bvh.nodes.data(), nodes, node_count * sizeof(Node));
    bvh.order.assign(order, order + count);
    return bvh_intersect(bvh, v.data(), pfx::mk(o[0], o[1], o[2]), pfx::mk(d[0], d[1], d[2]), 1e30f);
}
float or_triangle_hit(const float* v9, const float* o, const float* d, float limit) {
    return triangle_hit(v9, pfx::mk(o[0], o[1], o[2]), pfx::mk(d[0], d[1], d[2]), limit);
}
// levels.rs:225-383; counts only when nodes == NULL
int or_levels(const pfxh_scene* scene, pfxh_bvh_node* nodes, uint32_t* node_count, uint32_t* order,
              uint32_t* order_count, uint32_t* slot_prim, char* err, size_t err_len) {
    Laid laid;
    std::string e = lay_scene(*scene, laid);
    if (!e.empty()) {
        if (err && err_len) {
            size_t k = e.size() < err_len - 1 ? e.size() : err_len - 1;
            memcpy(err, e.data(), k);
            err[k] = 0;
        }
        return 1;
    }
    *node_count = (uint32_t)laid.nodes.size();
    *order_count = (uint32_t)laid.world_order.size();
    if (nodes) memcpy(nodes, laid.nodes.data(), laid.nodes.size() * sizeof(Node));
    if (order) memcpy(order, laid.world_order.data(), laid.world_order.size() * 4);
    if (slot_prim) memcpy(slot_prim, laid.slot_prim.data(), laid.slot_prim.size() * 4);
    return 0;
}
// tree.rs:220-243
uint32_t or_lights(const pfxh_light* lights, uint32_t count, float* rows, uint32_t* looped, uint32_t* tree) {
    LaidLights l = tree_lay(lights, count);
    memcpy(rows, l.rows.data(), l.rows.size() * sizeof(GLight));
    *looped = l.looped;
    *tree = l.tree;
    return (uint32_t)l.rows.size();
}
// tree_lay's split: 1 = stable sort by centroid, 0 = Rust's selection
void or_set_tree_split(int sorted) { set_tree_split_by_sort(sorted != 0); }
// lights.rs pack / falloff / cone
void or_light_pack(const pfxh_light* l, float* out16) {
    GLight g = light_pack(*l);
    memcpy(out16, &g, 64);
}
float or_light_falloff(const pfxh_light* l, float distance) { return light_falloff(*l, distance); }
float or_light_cone(const pfxh_light* l, const float* toward) { return light_cone(*l, toward); }
int or_light_valid(const pfxh_light* l) { return light_valid(*l) ? 1 : 0; }
// lights.rs:391-440 over explicit triangle candidates (9 vertices + 3 emission each)
uint32_t or_emitters(const float* tris12, uint32_t count, uint32_t looped, uint32_t tree, float* entries, float* cdf) {
    std::vector<EmitterCandidate> cs(count);
    for (uint32_t i = 0; i < count; ++i) {
        memcpy(cs[i].vertices, tris12 + 12 * i, 36);
        memcpy(cs[i].emission, tris12 + 12 * i + 9, 12);
        cs[i].record = i;
    }
    EmitterList list = build_emitters(cs, looped, tree);
    memcpy(entries, list.entries.data(), list.entries.size() * sizeof(GEmitter));
    if (!list.cdf.empty()) memcpy(cdf, list.cdf.data(), list.cdf.size() * 4);
    return (uint32_t)list.entries.size();
}
// sky.rs:207-272 for a lat-long environment
void or_sky_cdf(const float* texels, uint32_t w, uint32_t h, uint32_t cw, uint32_t ch, float* rows, float* columns) {
    Env env;
    env.width = w;
    env.height = h;
    if (w && h) env.texels.assign(texels, texels + (size_t)w * h * 4);
    SkyCdf cdf = sky_build(env, cw, ch);
    memcpy(rows, cdf.rows.data(), cdf.rows.size() * 4);
    memcpy(columns, cdf.columns.data(), cdf.columns.size() * 4);
}
// adaptive.rs:46-63
uint32_t or_rounds(const pfxh_adaptive* a, uint32_t* out, uint32_t cap) {
    std::vector<uint32_t> r = adaptive_rounds(*a);
    uint32_t n = (uint32_t)r.size() < cap ? (uint32_t)r.size() : cap;
    memcpy(out, r.data(), n * 4);
    return (uint32_t)r.size();
}

// Rust's select_nth_unstable_by (select_nth.h) over (key, id) pairs, keys
// compared with f32::total_cmp as the light tree does (tree.rs:281-283);
// mode 1 runs one `partition` around keys[index] instead, mode 2 the
// median-of-medians fallback alone.
struct KeyId {
    float key;
    uint32_t id;
};
static bool key_total_less(const KeyId& a, const KeyId& b) {
    int32_t x, y;
    memcpy(&x, &a.key, 4);
    memcpy(&y, &b.key, 4);
    x ^= (int32_t)((uint32_t)(x >> 31) >> 1);
    y ^= (int32_t)((uint32_t)(y >> 31) >> 1);
    return x < y;
}
size_t or_select_nth(float* keys, uint32_t* ids, size_t len, size_t index, int mode) {
    std::vector<KeyId> v(len);
    for (size_t i = 0; i < len; ++i) v[i] = KeyId{keys[i], ids[i]};
    size_t r = index;
    if (mode == 1)
        r = rsel::partition(v.data(), len, index, key_total_less);
    else if (mode == 2)
        rsel::median_of_medians(v.data(), len, index, key_total_less);
    else
        rsel::select_nth_unstable_by(v.data(), len, index, key_total_less);
    for (size_t i = 0; i < len; ++i) {
        keys[i] = v[i].key;
        ids[i] = v[i].id;
    }
    return r;
}

}  // extern "C"
