// Oracle (test infrastructure only): BVH build, two-level layout and
// triangle packing, restating crates/trace/src/bvh.rs, levels.rs and the
// packing half of gpu.rs SceneBuffers::build_with.
#include <math.h>
#include <string.h>

#include <cmath>
#include <array>
#include <limits>

#include "oracle.h"

namespace oracle {

static const float INF = std::numeric_limits<float>::infinity();

// bvh.rs:47-53
Bounds bounds_empty() {
    Bounds b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = INF;
        b.hi[k] = -INF;
    }
    return b;
}
// bvh.rs:54-59 (f32::min / f32::max are IEEE minNum / maxNum, as fminf)
void bounds_point(Bounds& b, const float p[3]) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = pfx::minf(b.lo[k], p[k]);
        b.hi[k] = pfx::maxf(b.hi[k], p[k]);
    }
}
// bvh.rs:60-63
void bounds_grow(Bounds& b, const Bounds& o) {
    bounds_point(b, o.lo);
    bounds_point(b, o.hi);
}
// bvh.rs:64-71
static float bounds_area(const Bounds& b) {
    float d0 = b.hi[0] - b.lo[0], d1 = b.hi[1] - b.lo[1], d2 = b.hi[2] - b.lo[2];
    if (d0 < 0.0f) return 0.0f;
    return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
}

// bvh.rs:74-88
static const int BINS = 16;
struct Leaves {
    size_t size, most;
};
static const Leaves TRIANGLE_LEAVES = {4, 16};
static const Leaves INSTANCE_LEAVES = {1, 0};

// bvh.rs:268-270: `as usize` saturates and maps NaN to 0
static size_t bin(float value, float lo, float extent) {
    float x = ((value - lo) / extent) * (float)BINS;
    if (x != x) return 0;
    if (x < 0.0f) x = 0.0f;
    if (x > (float)(BINS - 1)) x = (float)(BINS - 1);
    return (size_t)x;
}

// bvh.rs:167-266
static void split(const std::vector<Bounds>& boxes, const std::vector<std::array<float, 3>>& centers,
                  std::vector<uint32_t>& order, size_t start, size_t end, std::vector<Node>& nodes,
                  Leaves leaves) {
    Bounds bounds = bounds_empty(), centroids = bounds_empty();
    for (size_t i = start; i < end; ++i) {
        bounds_grow(bounds, boxes[order[i]]);
        bounds_point(centroids, centers[order[i]].data());
    }
    size_t here = nodes.size();
    Node node;
    memset(&node, 0, sizeof node);
    for (int k = 0; k < 3; ++k) {
        node.lo[k] = bounds.lo[k];
        node.hi[k] = bounds.hi[k];
    }
    node.first = (uint32_t)start;
    node.count = (uint32_t)(end - start);
    nodes.push_back(node);
    size_t len = end - start;
    if (len <= leaves.size) return;
    float best_cost = INF;
    size_t best_axis = 0, best_index = 0;
    for (size_t axis = 0; axis < 3; ++axis) {
        float extent = centroids.hi[axis] - centroids.lo[axis];
        if (extent <= 1e-9f) continue;
        Bounds bin_box[BINS];
        size_t bin_count[BINS];
        for (int b = 0; b < BINS; ++b) {
            bin_box[b] = bounds_empty();
            bin_count[b] = 0;
        }
        for (size_t i = start; i < end; ++i) {
            size_t b = bin(centers[order[i]][axis], centroids.lo[axis], extent);
            bounds_grow(bin_box[b], boxes[order[i]]);
            bin_count[b] += 1;
        }
        float areas[BINS];
        size_t counts[BINS];
        Bounds bound = bounds_empty();
        size_t count = 0;
        for (int index = 0; index < BINS; ++index) {
            if (bin_count[index] > 0) bounds_grow(bound, bin_box[index]);
            count += bin_count[index];
            areas[index] = bounds_area(bound);
            counts[index] = count;
        }
        bound = bounds_empty();
        count = 0;
        for (int index = BINS - 1; index >= 1; --index) {
            if (bin_count[index] > 0) bounds_grow(bound, bin_box[index]);
            count += bin_count[index];
            float cost = areas[index - 1] * (float)counts[index - 1] + bounds_area(bound) * (float)count;
            if (counts[index - 1] > 0 && count > 0 && cost < best_cost) {
                best_cost = cost;
                best_axis = axis;
                best_index = (size_t)index;
            }
        }
    }
    if (best_cost >= bounds_area(bounds) * (float)len && len <= leaves.most) return;
    size_t mid;
    if (std::isfinite(best_cost)) {
        size_t axis = best_axis;
        float extent = centroids.hi[axis] - centroids.lo[axis];
        size_t left = 0, right = len;
        while (left < right) {
            if (bin(centers[order[start + left]][axis], centroids.lo[axis], extent) < best_index) {
                left += 1;
            } else {
                right -= 1;
                std::swap(order[start + left], order[start + right]);
            }
        }
        mid = (left == 0 || left == len) ? start + len / 2 : start + left;
    } else {
        mid = start + len / 2;
    }
    split(boxes, centers, order, start, mid, nodes, leaves);
    size_t right_at = nodes.size();
    split(boxes, centers, order, mid, end, nodes, leaves);
    nodes[here].first = (uint32_t)right_at;
    nodes[here].count = 0;
}

// bvh.rs:111-133
static Bvh over(const std::vector<Bounds>& boxes, Leaves leaves) {
    std::vector<std::array<float, 3>> centers(boxes.size());
    for (size_t i = 0; i < boxes.size(); ++i)
        for (int k = 0; k < 3; ++k) centers[i][k] = (boxes[i].lo[k] + boxes[i].hi[k]) * 0.5f;
    Bvh bvh;
    bvh.order.resize(boxes.size());
    for (size_t i = 0; i < boxes.size(); ++i) bvh.order[i] = (uint32_t)i;
    bvh.nodes.reserve(2 * boxes.size());
    if (!boxes.empty()) split(boxes, centers, bvh.order, 0, boxes.size(), bvh.nodes, leaves);
    return bvh;
}

// bvh.rs:95-105
Bvh bvh_build(const float* v9, size_t count) {
    std::vector<Bounds> boxes(count);
    for (size_t i = 0; i < count; ++i) {
        boxes[i] = bounds_empty();
        for (int v = 0; v < 3; ++v) bounds_point(boxes[i], v9 + 9 * i + 3 * v);
    }
    return over(boxes, TRIANGLE_LEAVES);
}

Bvh bvh_build_instances(const std::vector<Bounds>& boxes) { return over(boxes, INSTANCE_LEAVES); }

// bvh.rs:272-292
static bool box_hit_cpu(const Node& node, v3 o, v3 d, float limit) {
    float near_ = 0.0f, far_ = limit;
    for (int axis = 0; axis < 3; ++axis) {
        float dir = pfx::comp(d, axis), org = pfx::comp(o, axis);
        if (fabsf(dir) < 1e-12f) {
            if (org < node.lo[axis] || org > node.hi[axis]) return false;
        } else {
            float a = (node.lo[axis] - org) / dir;
            float b = (node.hi[axis] - org) / dir;
            near_ = pfx::maxf(near_, pfx::minf(a, b));
            far_ = pfx::minf(far_, pfx::maxf(a, b));
            if (near_ > far_) return false;
        }
    }
    return true;
}

// bvh.rs:294-324
float triangle_hit(const float* v, v3 o, v3 d, float limit) {
    v3 a = pfx::mk(v[0], v[1], v[2]), b = pfx::mk(v[3], v[4], v[5]), c = pfx::mk(v[6], v[7], v[8]);
    v3 e1 = b - a, e2 = c - a;
    v3 p = pfx::cross(d, e2);
    float det = pfx::dot(e1, p);
    if (fabsf(det) < 1e-9f) return -1.0f;
    float inv = 1.0f / det;
    v3 s = o - a;
    float u = pfx::dot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return -1.0f;
    v3 q = pfx::cross(s, e1);
    float vv = pfx::dot(d, q) * inv;
    if (vv < 0.0f || u + vv > 1.0f) return -1.0f;
    float distance = pfx::dot(e2, q) * inv;
    if (distance <= 1e-5f || distance >= limit) return -1.0f;
    return distance;
}

// bvh.rs:135-164
uint32_t bvh_intersect(const Bvh& bvh, const float* v9, v3 o, v3 d, float limit) {
    if (bvh.nodes.empty()) return UINT32_MAX;
    uint32_t best = UINT32_MAX;
    float best_t = limit;
    std::vector<size_t> stack{0};
    while (!stack.empty()) {
        size_t index = stack.back();
        stack.pop_back();
        const Node& node = bvh.nodes[index];
        if (!box_hit_cpu(node, o, d, best_t)) continue;
        if (node.count == 0) {
            stack.push_back(node.first);
            stack.push_back(index + 1);
        } else {
            for (uint32_t j = node.first; j < node.first + node.count; ++j) {
                uint32_t i = bvh.order[j];
                float t = triangle_hit(v9 + 9 * (size_t)i, o, d, best_t);
                if (t >= 0.0f) {
                    best = i;
                    best_t = t;
                }
            }
        }
    }
    return best;
}

// ---- detail::Pose (crates/trace/src/detail.rs:184-287) --------------------
struct Pose {
    float m[4][4];  // column-major: m[column][row]
};
static double pa(const Pose& p, int row, int column) { return (double)p.m[column][row]; }
static double cofactor(const Pose& p, int row, int column) {
    int r0 = (row + 1) % 3, r1 = (row + 2) % 3, c0 = (column + 1) % 3, c1 = (column + 2) % 3;
    return pa(p, r0, c0) * pa(p, r1, c1) - pa(p, r0, c1) * pa(p, r1, c0);
}
static std::string pose_new(const float t[16], Pose& out) {
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(t[i])) return "instance transform is not finite";
    memcpy(out.m, t, sizeof out.m);
    if (out.m[0][3] != 0.0f || out.m[1][3] != 0.0f || out.m[2][3] != 0.0f || out.m[3][3] != 1.0f)
        return "instance transform is not affine";
    double det = -0.0;
    for (int c = 0; c < 3; ++c) det = det + pa(out, 0, c) * cofactor(out, 0, c);
    double volume = 1.0;
    for (int c = 0; c < 3; ++c) {
        double s = -0.0;
        for (int r = 0; r < 3; ++r) s = s + pa(out, r, c) * pa(out, r, c);
        volume = volume * std::sqrt(s);
    }
    if (std::fabs(det) <= volume * 1e-6) return "instance transform is singular";
    return "";
}
// detail.rs:232-240
static void pose_point(const Pose& p, const float in[3], float out[3]) {
    for (int row = 0; row < 3; ++row) {
        float s = -0.0f;
        for (int c = 0; c < 3; ++c) s = s + p.m[c][row] * in[c];
        out[row] = p.m[3][row] + s;
    }
}
// detail.rs:254-260
static void pose_rows(const Pose& p, float r[12]) {
    for (int row = 0; row < 3; ++row)
        for (int k = 0; k < 4; ++k) r[row * 4 + k] = p.m[k][row];
}
// detail.rs:262-286
static void pose_inverse_rows(const Pose& p, float r[12]) {
    double det = -0.0;
    for (int c = 0; c < 3; ++c) det = det + pa(p, 0, c) * cofactor(p, 0, c);
    double inv[3][3];
    for (int row = 0; row < 3; ++row)
        for (int c = 0; c < 3; ++c) inv[row][c] = cofactor(p, c, row) / det;
    double tr[3] = {p.m[3][0], p.m[3][1], p.m[3][2]};
    for (int row = 0; row < 3; ++row) {
        double shift = -0.0;
        for (int k = 0; k < 3; ++k) shift = shift + inv[row][k] * tr[k];
        for (int c = 0; c < 3; ++c) r[row * 4 + c] = (float)inv[row][c];
        r[row * 4 + 3] = (float)(-shift);
    }
}
// detail.rs:223-230
static float pose_inverse_determinant(const Pose& p) {
    auto a = [&](int row, int column) { return pa(p, row, column); };
    double det = a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) - a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
                 a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
    return (float)(1.0 / det);
}

static const uint32_t INSTANCE_LEAF = 1;
static const uint32_t NO_MATERIAL = 0xffffffffu;

static Surface surface_of(const uint32_t info[4], const float fwd[12], const float inv[12]) {
    Surface s;
    memset(&s, 0, sizeof s);
    memcpy(s.info, info, sizeof s.info);
    memcpy(s.forward, fwd, sizeof s.forward);
    memcpy(s.inverse, inv, sizeof s.inverse);
    s.transform[2] = 1.0f;  // InstanceSurface::default uv scale
    s.transform[3] = 1.0f;
    return s;
}

struct DetailTri {
    float v[9];
    float n[9];
    uint32_t material;
};

// gpu.rs:539-563 pack_triangle (uvs are zero in this API)
static Tri pack_triangle(const DetailTri& t) {
    Tri r;
    memset(&r, 0, sizeof r);
    for (int k = 0; k < 3; ++k) {
        r.a[k] = t.v[k];
        r.b[k] = t.v[3 + k];
        r.c[k] = t.v[6 + k];
        r.na[k] = t.n[k];
        r.nb[k] = t.n[3 + k];
        r.nc[k] = t.n[6 + k];
    }
    r.a[3] = (float)t.material;
    return r;
}

static bool same_bits(const DetailTri& a, const DetailTri& b) {
    return memcmp(a.v, b.v, sizeof a.v) == 0 && memcmp(a.n, b.n, sizeof a.n) == 0;
}

// levels.rs:225-383 + gpu.rs:582-911 (packing and records; emitters are
// marked by the renderer's scene preparation).
std::string lay_scene(const pfxh_scene& sc, Laid& out) {
    std::vector<DetailTri> world;
    world.reserve((size_t)sc.triangle_count + sc.world_detail_count);
    for (uint32_t i = 0; i < sc.triangle_count; ++i) {
        DetailTri t;
        memcpy(t.v, sc.triangles[i].vertices, sizeof t.v);
        memset(t.n, 0, sizeof t.n);
        t.material = sc.triangles[i].material;
        world.push_back(t);
    }
    for (uint32_t i = 0; i < sc.world_detail_count; ++i) {
        DetailTri t;
        memcpy(t.v, sc.world_details[i].vertices, sizeof t.v);
        memcpy(t.n, sc.world_details[i].normals, sizeof t.n);
        t.material = sc.world_details[i].material;
        world.push_back(t);
    }
    std::vector<float> wv(world.size() * 9);
    for (size_t i = 0; i < world.size(); ++i) memcpy(&wv[9 * i], world[i].v, 36);
    Bvh world_bvh = bvh_build(wv.data(), world.size());
    out = Laid();
    uint32_t prim_base_mesh = sc.triangle_count + sc.world_detail_count;
    uint32_t ident_info[4] = {0, NO_MATERIAL, 0, pfx::f2u(1.0f)};
    float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    out.surfaces.push_back(surface_of(ident_info, ident, ident));
    out.instance_of_record.push_back(0);
    if (sc.instance_count == 0) {
        out.nodes = world_bvh.nodes;
        out.world_order = world_bvh.order;
        for (uint32_t slot = 0; slot < world.size(); ++slot) {
            out.tris.push_back(pack_triangle(world[world_bvh.order[slot]]));
            out.slot_prim.push_back(world_bvh.order[slot]);
        }
        out.slots = (uint32_t)world.size();
        return "";
    }
    struct Mesh {
        uint32_t source;  // caller mesh index
        std::vector<uint32_t> order;
        std::vector<Node> nodes;
        uint32_t root, first;
        Bounds bounds;
    };
    struct Placed {
        uint32_t mesh, base, record, word, material;
        Bounds bounds;
        float inverse[12];
        Pose pose;
    };
    std::vector<Mesh> meshes;
    struct Known {
        uint32_t source;
        size_t mesh;
    };
    std::vector<Known> known;
    std::vector<int> placed_of(sc.instance_count, -1);
    std::vector<Placed> placed;
    uint32_t next = (uint32_t)world.size();
    uint32_t record = 0;
    auto tris_of = [&](uint32_t mesh_index) {
        const pfxh_mesh& m = sc.meshes[mesh_index];
        std::vector<DetailTri> ts(m.count);
        for (uint32_t j = 0; j < m.count; ++j) {
            const pfxh_triangle_detail& d = sc.mesh_triangles[m.first + j];
            memcpy(ts[j].v, d.vertices, 36);
            memcpy(ts[j].n, d.normals, 36);
            ts[j].material = d.material;
        }
        return ts;
    };
    for (uint32_t m = 0; m < sc.mesh_count; ++m)
        if ((uint64_t)sc.meshes[m].first + sc.meshes[m].count > sc.mesh_triangle_count)
            return "mesh range is outside mesh_triangles";
    for (uint32_t i = 0; i < sc.instance_count; ++i) {
        const pfxh_instance& inst = sc.instances[i];
        if (inst.mesh >= sc.mesh_count) return "instance names a missing mesh";
        std::vector<DetailTri> ts = tris_of(inst.mesh);
        if (ts.empty()) continue;
        Pose pose;
        std::string e = pose_new(inst.transform, pose);
        if (!e.empty()) return e;
        uint32_t material = ts[0].material;
        for (auto& t : ts)
            if (t.material != material) material = NO_MATERIAL;
        // instances sharing a mesh (the reference's shared Arc, or bit-equal
        // geometry found through its geometry hash) share one BLAS
        long found = -1;
        for (auto& k : known) {
            if (k.source == inst.mesh) {
                found = (long)k.mesh;
                break;
            }
        }
        if (found < 0) {
            for (auto& k : known) {
                const Mesh& mesh = meshes[k.mesh];
                std::vector<DetailTri> other = tris_of(mesh.source);
                if (other.size() != ts.size()) continue;
                bool same = true;
                for (size_t j = 0; j < ts.size() && same; ++j) {
                    uint32_t o = mesh.order[j];
                    same = same_bits(ts[o], other[o]) && (material != NO_MATERIAL || ts[o].material == other[o].material);
                }
                if (same) {
                    found = (long)k.mesh;
                    break;
                }
            }
        }
        size_t mesh_index;
        if (found >= 0) {
            mesh_index = (size_t)found;
        } else {
            std::vector<float> mv(ts.size() * 9);
            for (size_t j = 0; j < ts.size(); ++j) memcpy(&mv[9 * j], ts[j].v, 36);
            Bvh bvh = bvh_build(mv.data(), ts.size());
            Mesh m;
            m.source = inst.mesh;
            m.order = bvh.order;
            m.nodes = bvh.nodes;
            m.root = 0;
            m.first = next;
            for (int k = 0; k < 3; ++k) {
                m.bounds.lo[k] = bvh.nodes[0].lo[k];
                m.bounds.hi[k] = bvh.nodes[0].hi[k];
            }
            meshes.push_back(m);
            next += (uint32_t)ts.size();
            mesh_index = meshes.size() - 1;
            known.push_back(Known{inst.mesh, mesh_index});
        }
        record += 1;
        Placed p;
        p.mesh = (uint32_t)mesh_index;
        p.base = meshes[mesh_index].first;
        p.record = record;
        p.word = inst.word;
        p.material = material;
        p.pose = pose;
        // levels.rs:110-125 placed_bounds
        p.bounds = bounds_empty();
        const Bounds& local = meshes[mesh_index].bounds;
        for (int corner = 0; corner < 8; ++corner) {
            float pt[3], w[3];
            for (int axis = 0; axis < 3; ++axis) pt[axis] = ((corner >> axis) & 1) == 0 ? local.lo[axis] : local.hi[axis];
            pose_point(pose, pt, w);
            bounds_point(p.bounds, w);
        }
        pose_inverse_rows(pose, p.inverse);
        placed_of[i] = (int)placed.size();
        placed.push_back(p);
    }
    // levels.rs:337-366
    bool has_world = !world_bvh.nodes.empty();
    size_t groups = (has_world ? 1 : 0) + placed.size();
    size_t top_len = groups == 0 ? 0 : 3 * groups - 1;
    uint32_t offset = (uint32_t)top_len;
    uint32_t world_root = 0;
    std::vector<Node> world_nodes = world_bvh.nodes;
    if (has_world) {
        world_root = offset;
        for (auto& n : world_nodes)
            if (n.count == 0) n.first += offset;
        offset += (uint32_t)world_nodes.size();
    }
    for (auto& m : meshes) {
        m.root = offset;
        for (auto& n : m.nodes)
            if (n.count == 0) n.first += offset;
        offset += (uint32_t)m.nodes.size();
    }
    // levels.rs:421-470 Levels::top
    struct Group {
        Bounds bounds;
        uint32_t root, base, record;
        bool has_inverse;
        float inverse[12];
    };
    std::vector<Group> gs;
    if (has_world) {
        Group g;
        for (int k = 0; k < 3; ++k) {
            g.bounds.lo[k] = world_bvh.nodes[0].lo[k];
            g.bounds.hi[k] = world_bvh.nodes[0].hi[k];
        }
        g.root = world_root;
        g.base = 0;
        g.record = 0;
        g.has_inverse = false;
        gs.push_back(g);
    }
    for (auto& p : placed) {
        Group g;
        g.bounds = p.bounds;
        g.root = meshes[p.mesh].root;
        g.base = p.base;
        g.record = p.record;
        g.has_inverse = true;
        memcpy(g.inverse, p.inverse, sizeof g.inverse);
        gs.push_back(g);
    }
    std::vector<Node> nodes;
    if (!gs.empty()) {
        std::vector<Bounds> boxes;
        for (auto& g : gs) boxes.push_back(g.bounds);
        Bvh top = bvh_build_instances(boxes);
        uint32_t tree = (uint32_t)top.nodes.size();
        for (Node n : top.nodes) {
            if (n.count > 0) {
                uint32_t index = top.order[n.first];
                const Group& g = gs[index];
                n.first = g.root;
                n.count = g.base;
                n.lo[3] = pfx::u2f(g.record);
                n.pad[0] = INSTANCE_LEAF;
                n.pad[1] = g.has_inverse ? tree + index : 0;
            }
            nodes.push_back(n);
        }
        for (auto& g : gs) {
            float r[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
            if (g.has_inverse) memcpy(r, g.inverse, sizeof r);
            Node n;
            memcpy(n.lo, r, 16);
            memcpy(n.hi, r + 4, 16);
            n.first = pfx::f2u(r[8]);
            n.count = pfx::f2u(r[9]);
            n.pad[0] = pfx::f2u(r[10]);
            n.pad[1] = pfx::f2u(r[11]);
            nodes.push_back(n);
        }
    }
    Node zero;
    memset(&zero, 0, sizeof zero);
    nodes.resize(top_len, zero);
    nodes.insert(nodes.end(), world_nodes.begin(), world_nodes.end());
    for (auto& m : meshes) nodes.insert(nodes.end(), m.nodes.begin(), m.nodes.end());
    out.nodes = nodes;
    out.world_order = world_bvh.order;
    // gpu.rs:816-827 triangles in BVH order: world slots then meshes
    for (uint32_t slot = 0; slot < world.size(); ++slot) {
        out.tris.push_back(pack_triangle(world[world_bvh.order[slot]]));
        out.slot_prim.push_back(world_bvh.order[slot]);
    }
    for (auto& m : meshes) {
        std::vector<DetailTri> ts = tris_of(m.source);
        for (uint32_t j = 0; j < m.order.size(); ++j) {
            out.tris.push_back(pack_triangle(ts[m.order[j]]));
            out.slot_prim.push_back(prim_base_mesh + sc.meshes[m.source].first + m.order[j]);
        }
    }
    out.slots = next;
    // gpu.rs:836-847 records: world, then one per placed instance
    for (uint32_t i = 0; i < sc.instance_count; ++i) {
        if (placed_of[i] < 0) continue;
        const Placed& p = placed[placed_of[i]];
        uint32_t info[4] = {p.word, p.material, pfx::f2u(0.5f), pfx::f2u(pose_inverse_determinant(p.pose))};
        float fwd[12];
        pose_rows(p.pose, fwd);
        out.surfaces.push_back(surface_of(info, fwd, p.inverse));
        out.instance_of_record.push_back(i + 1);
    }
    return "";
}

}  // namespace oracle
