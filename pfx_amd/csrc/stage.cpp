// Staging: indexed meshes and their placements in, a tracer scene out --
// the producer `Stage::build` of crates/trace/src/stage.rs:210-385 for the
// parts the still path traces (no content images or text, no motion, no
// clip planes: those layers are not built here).
//
// Every placement becomes an instance of the placement's mesh, triangulated
// (indexed corners gathered, triangles whose corners fall below the
// placement's alpha cutoff cut along the cutoff and fanned,
// stage.rs:736-822) and dressed in the placement's material. Placements of
// one mesh at one cutoff share the cut (stage.rs:286-290); placements that
// also share a material share the dressed triangles, so they share one
// bottom-level BVH in the trace (the reference's shared Arc,
// levels.rs:266-283). The staged scene has no world triangles; its detail
// is Detail::default() with the stage's lens and projection
// (stage.rs:369-381).
#include <math.h>
#include <string.h>

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pfxhip.h"
#include "trace_internal.h"

struct pfxh_staged {
    std::vector<pfxh_triangle_detail> triangles;
    std::vector<pfxh_mesh> meshes;
    std::vector<pfxh_instance> instances;
    std::vector<pfxh_material> materials;
    std::vector<pfxh_light> lights;
    std::vector<float> sky;
    pfxh_scene scene{};
    pfxh_detail detail{};
};

namespace {

const float ALPHA_CUTOFF = 0.5f;  // stage.rs:25

struct Corner {  // stage.rs:705-711 (no uv sets: the tracer shades none)
    float position[3];
    float normal[3];
    float alpha;
};

Corner lerp(const Corner& a, const Corner& b, float t, float alpha) {  // stage.rs:713-715
    Corner c;
    for (int k = 0; k < 3; ++k) {
        c.position[k] = a.position[k] + (b.position[k] - a.position[k]) * t;
        c.normal[k] = a.normal[k] + (b.normal[k] - a.normal[k]) * t;
    }
    c.alpha = alpha;
    return c;
}

// cut / cut_at (stage.rs:770-822): the part of a triangle at or above the
// cutoff, as a polygon of up to four corners.
std::vector<Corner> cut_at(const Corner (&corners)[3], float cutoff) {
    if (corners[0].alpha >= cutoff && corners[1].alpha >= cutoff && corners[2].alpha >= cutoff)
        return std::vector<Corner>(corners, corners + 3);
    std::vector<Corner> kept;
    for (int k = 0; k < 3; ++k) {
        const Corner& a = corners[k];
        const Corner& b = corners[(k + 1) % 3];
        bool inside_a = a.alpha >= cutoff, inside_b = b.alpha >= cutoff;
        if (inside_a) kept.push_back(a);
        if (inside_a != inside_b) kept.push_back(lerp(a, b, (cutoff - a.alpha) / (b.alpha - a.alpha), cutoff));
    }
    return kept;
}

std::string check_mesh(const pfxh_stage_mesh& m) {  // stage.rs:717-734
    if (m.vertex_count == 0 || m.index_count == 0 || m.index_count % 3 != 0 || !m.positions || !m.indices)
        return "mesh needs vertices and triangle indices";
    if (!m.normals) return "mesh attributes have different vertex counts";
    for (uint32_t i = 0; i < m.index_count; ++i)
        if (m.indices[i] >= m.vertex_count) return "mesh index is outside the vertex buffer";
    return "";
}

// triangles (stage.rs:736-768): fan of each kept polygon, material 0
std::vector<pfxh_triangle_detail> triangles(const pfxh_stage_mesh& m, float cutoff) {
    std::vector<pfxh_triangle_detail> out;
    out.reserve(m.index_count / 3);
    for (uint32_t f = 0; f < m.index_count; f += 3) {
        Corner corners[3];
        for (int k = 0; k < 3; ++k) {
            uint32_t i = m.indices[f + k];
            memcpy(corners[k].position, m.positions + 3 * (size_t)i, 12);
            memcpy(corners[k].normal, m.normals + 3 * (size_t)i, 12);
            corners[k].alpha = m.alpha ? m.alpha[i] : 1.0f;
        }
        std::vector<Corner> kept = cut_at(corners, cutoff);
        for (size_t k = 1; k + 1 < kept.size(); ++k) {
            const Corner* c[3] = {&kept[0], &kept[k], &kept[k + 1]};
            pfxh_triangle_detail t{};
            for (int v = 0; v < 3; ++v) {
                memcpy(t.vertices + 3 * v, c[v]->position, 12);
                memcpy(t.normals + 3 * v, c[v]->normal, 12);
            }
            out.push_back(t);
        }
    }
    return out;
}

uint32_t bits_of(float v) {
    uint32_t b;
    memcpy(&b, &v, 4);
    return b;
}

bool finite_all(const float* v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

}  // namespace

extern "C" {

pfxh_status pfxh_stage_build(const pfxh_stage* st, pfxh_staged** out, char* err, size_t err_len) {
    if (!st || !out) return PFXH_INVALID;
    *out = nullptr;
    auto fail = [&](const std::string& e) {
        put_err(err, err_len, e);
        return PFXH_SETUP;
    };
    if (st->material_count == 0 || !st->materials) return fail("staged scene needs at least one material");
    bool projection_ok = st->projection <= PFXH_PROJ_EQUIRECT &&
                         (st->projection != PFXH_PROJ_ORTHOGRAPHIC ||
                          (std::isfinite(st->ortho_size[0]) && std::isfinite(st->ortho_size[1]) &&
                           st->ortho_size[0] > 0.0f && st->ortho_size[1] > 0.0f));
    bool lens_ok = finite_all(st->lens_shift, 2) && std::isfinite(st->aperture) && st->aperture >= 0.0f &&
                   std::isfinite(st->focus_distance) && st->focus_distance > 0.0f;
    if (!projection_ok || !lens_ok) return fail("staged camera has an invalid projection or lens");
    if (st->mesh_count && !st->meshes) return fail("mesh needs vertices and triangle indices");
    for (uint32_t i = 0; i < st->mesh_count; ++i) {
        std::string e = check_mesh(st->meshes[i]);
        if (!e.empty()) return fail(e);
    }
    if (st->placement_count && !st->placements) return fail("staged instance names a missing mesh");
    return guarded(err, err_len, [&] {
        auto staged = std::make_unique<pfxh_staged>();
        std::map<std::pair<uint32_t, uint32_t>, std::vector<pfxh_triangle_detail>> cuts;
        std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> dressed;
        for (uint32_t k = 0; k < st->placement_count; ++k) {
            const pfxh_placement& p = st->placements[k];
            if (!(p.alpha_cutoff >= 0.0f && p.alpha_cutoff <= 1.0f))
                throw DeviceError{PFXH_SETUP, "staged alpha cutoff must be a fraction from 0 to 1"};
            if (p.mesh >= st->mesh_count) throw DeviceError{PFXH_SETUP, "staged instance names a missing mesh"};
            if (p.material >= st->material_count)
                throw DeviceError{PFXH_SETUP, "staged instance names a missing material"};
            uint32_t key = bits_of(p.alpha_cutoff);
            auto at = dressed.find({p.mesh, key, p.material});
            uint32_t mesh_index;
            if (at != dressed.end()) {
                mesh_index = at->second;
            } else {
                auto cut = cuts.find({p.mesh, key});
                if (cut == cuts.end())
                    cut = cuts.emplace(std::make_pair(p.mesh, key), triangles(st->meshes[p.mesh], p.alpha_cutoff)).first;
                pfxh_mesh range{(uint32_t)staged->triangles.size(), (uint32_t)cut->second.size()};
                for (pfxh_triangle_detail t : cut->second) {
                    t.material = p.material;
                    staged->triangles.push_back(t);
                }
                mesh_index = (uint32_t)staged->meshes.size();
                staged->meshes.push_back(range);
                dressed.emplace(std::make_tuple(p.mesh, key, p.material), mesh_index);
            }
            pfxh_instance inst{};
            inst.mesh = mesh_index;
            inst.word = (p.shadow_only ? 1u : 0u) | (p.casts_shadow ? 0u : 2u) | (p.two_sided ? 0u : (1u << 23));
            memcpy(inst.transform, p.model, sizeof inst.transform);
            staged->instances.push_back(inst);
        }
        staged->materials.assign(st->materials, st->materials + st->material_count);
        if (st->light_count) staged->lights.assign(st->lights, st->lights + st->light_count);
        if (st->sky.width && st->sky.height && st->sky.texels)
            staged->sky.assign(st->sky.texels, st->sky.texels + (size_t)st->sky.width * st->sky.height * 4);
        pfxh_scene& sc = staged->scene;
        sc.mesh_triangles = staged->triangles.data();
        sc.mesh_triangle_count = (uint32_t)staged->triangles.size();
        sc.meshes = staged->meshes.data();
        sc.mesh_count = (uint32_t)staged->meshes.size();
        sc.instances = staged->instances.data();
        sc.instance_count = (uint32_t)staged->instances.size();
        sc.materials = staged->materials.data();
        sc.material_count = (uint32_t)staged->materials.size();
        sc.lights = staged->lights.empty() ? nullptr : staged->lights.data();
        sc.light_count = (uint32_t)staged->lights.size();
        sc.sky.width = staged->sky.empty() ? 0u : st->sky.width;
        sc.sky.height = staged->sky.empty() ? 0u : st->sky.height;
        sc.sky.texels = staged->sky.empty() ? nullptr : staged->sky.data();
        sc.camera = st->camera;
        sc.sun = st->sun;
        pfxh_detail& d = staged->detail;
        d.bounces = pfxh_bounces{16, 4, 8, 12};  // Bounces::REFERENCE (detail.rs:422-427)
        d.lens_shift[0] = st->lens_shift[0];
        d.lens_shift[1] = st->lens_shift[1];
        d.aperture = st->aperture;
        d.focus_distance = st->focus_distance;
        d.projection = st->projection;
        d.ortho_size[0] = st->ortho_size[0];
        d.ortho_size[1] = st->ortho_size[1];
        *out = staged.release();
    });
}

const pfxh_scene* pfxh_staged_scene(const pfxh_staged* staged) { return staged ? &staged->scene : nullptr; }

pfxh_status pfxh_staged_detail(const pfxh_staged* staged, pfxh_detail* out) {
    if (!staged || !out) return PFXH_INVALID;
    *out = staged->detail;
    return PFXH_OK;
}

pfxh_status pfxh_staged_trace(const pfxh_staged* staged, const pfxh_detail* detail, uint32_t width, uint32_t height,
                              pfxh_trace** out, char* err, size_t err_len) {
    if (!staged) return PFXH_INVALID;
    return pfxh_trace_new(&staged->scene, detail ? detail : &staged->detail, width, height, out, err, err_len);
}

void pfxh_staged_free(pfxh_staged* staged) { delete staged; }

}  // extern "C"
