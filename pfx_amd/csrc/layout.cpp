// Host scene layout: the reference's CPU scene build, then device formats.
//
// The build steps keep the reference's float operation order so node and
// triangle order are bit-identical to it: Bvh::try_build
// (crates/trace/src/bvh.rs:95-270), levels::lay / Levels::top
// (crates/trace/src/levels.rs:225-470), tree::lay (tree.rs:220-395),
// lights::build_emitters (lights.rs:391-440) and SkyCdf::build
// (sky.rs:207-272). Compiled with -ffp-contract=off.
#include "layout.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>

#include "vec3.h"

namespace pfxh {

using pfx::fm_acos;
using pfx::fm_cos;
using pfx::fm_sin;
using pfx::v3;

namespace {

const float INF = std::numeric_limits<float>::infinity();
const float PI_F = 3.14159265358979f;
const float HALF_PI = 1.57079632679490f;
const float TAU_F = 6.28318530717959f;

struct Box {
    float lo[3], hi[3];
    void empty() {
        for (int k = 0; k < 3; ++k) {
            lo[k] = INF;
            hi[k] = -INF;
        }
    }
    void point(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = pfx::minf(lo[k], p[k]);
            hi[k] = pfx::maxf(hi[k], p[k]);
        }
    }
    void grow(const Box& b) {
        point(b.lo);
        point(b.hi);
    }
    float area() const {
        float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
        if (d0 < 0.0f) return 0.0f;
        return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
    }
};

constexpr int BINS = 16;

inline size_t bin_of(float value, float lo, float extent) {
    float x = ((value - lo) / extent) * (float)BINS;
    if (!(x == x)) return 0;
    x = x < 0.0f ? 0.0f : (x > (float)(BINS - 1) ? (float)(BINS - 1) : x);
    return (size_t)x;
}

// Binned SAH builder over precomputed boxes; one instance per build.
struct Builder {
    const std::vector<Box>& boxes;
    std::vector<float> cx, cy, cz;  // centres, SoA
    std::vector<uint32_t>& order;
    std::vector<pfxh_bvh_node>& nodes;
    size_t leaf_size, leaf_most;

    Builder(const std::vector<Box>& b, std::vector<uint32_t>& o, std::vector<pfxh_bvh_node>& n, size_t ls, size_t lm)
        : boxes(b), order(o), nodes(n), leaf_size(ls), leaf_most(lm) {
        size_t count = boxes.size();
        cx.resize(count);
        cy.resize(count);
        cz.resize(count);
        for (size_t i = 0; i < count; ++i) {
            cx[i] = (boxes[i].lo[0] + boxes[i].hi[0]) * 0.5f;
            cy[i] = (boxes[i].lo[1] + boxes[i].hi[1]) * 0.5f;
            cz[i] = (boxes[i].lo[2] + boxes[i].hi[2]) * 0.5f;
        }
    }
    const float* centres(size_t axis) const { return axis == 0 ? cx.data() : (axis == 1 ? cy.data() : cz.data()); }

    void split(size_t start, size_t end) {
        Box bounds, centroids;
        bounds.empty();
        centroids.empty();
        for (size_t i = start; i < end; ++i) {
            uint32_t k = order[i];
            bounds.grow(boxes[k]);
            float c[3] = {cx[k], cy[k], cz[k]};
            centroids.point(c);
        }
        size_t here = nodes.size();
        pfxh_bvh_node node;
        memset(&node, 0, sizeof node);
        for (int k = 0; k < 3; ++k) {
            node.lo[k] = bounds.lo[k];
            node.hi[k] = bounds.hi[k];
        }
        node.first = (uint32_t)start;
        node.count = (uint32_t)(end - start);
        nodes.push_back(node);
        size_t len = end - start;
        if (len <= leaf_size) return;
        float best_cost = INF;
        size_t best_axis = 0, best_index = 0;
        for (size_t axis = 0; axis < 3; ++axis) {
            float lo = centroids.lo[axis];
            float extent = centroids.hi[axis] - lo;
            if (extent <= 1e-9f) continue;
            const float* c = centres(axis);
            Box bin_box[BINS];
            size_t bin_count[BINS];
            for (int b = 0; b < BINS; ++b) {
                bin_box[b].empty();
                bin_count[b] = 0;
            }
            for (size_t i = start; i < end; ++i) {
                uint32_t k = order[i];
                size_t b = bin_of(c[k], lo, extent);
                bin_box[b].grow(boxes[k]);
                bin_count[b] += 1;
            }
            float areas[BINS];
            size_t counts[BINS];
            Box bound;
            bound.empty();
            size_t count = 0;
            for (int b = 0; b < BINS; ++b) {
                if (bin_count[b] > 0) bound.grow(bin_box[b]);
                count += bin_count[b];
                areas[b] = bound.area();
                counts[b] = count;
            }
            bound.empty();
            count = 0;
            for (int b = BINS - 1; b >= 1; --b) {
                if (bin_count[b] > 0) bound.grow(bin_box[b]);
                count += bin_count[b];
                float cost = areas[b - 1] * (float)counts[b - 1] + bound.area() * (float)count;
                if (counts[b - 1] > 0 && count > 0 && cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_index = (size_t)b;
                }
            }
        }
        if (best_cost >= bounds.area() * (float)len && len <= leaf_most) return;
        size_t mid;
        if (std::isfinite(best_cost)) {
            const float* c = centres(best_axis);
            float lo = centroids.lo[best_axis];
            float extent = centroids.hi[best_axis] - lo;
            uint32_t* slice = order.data() + start;
            size_t left = 0, right = len;
            while (left < right) {
                if (bin_of(c[slice[left]], lo, extent) < best_index) {
                    left += 1;
                } else {
                    right -= 1;
                    std::swap(slice[left], slice[right]);
                }
            }
            mid = (left == 0 || left == len) ? start + len / 2 : start + left;
        } else {
            mid = start + len / 2;
        }
        split(start, mid);
        size_t right_at = nodes.size();
        split(mid, end);
        nodes[here].first = (uint32_t)right_at;
        nodes[here].count = 0;
    }
};

RefBvh build_over(const std::vector<Box>& boxes, size_t leaf_size, size_t leaf_most) {
    RefBvh out;
    out.order.resize(boxes.size());
    for (size_t i = 0; i < boxes.size(); ++i) out.order[i] = (uint32_t)i;
    out.nodes.reserve(2 * boxes.size());
    if (!boxes.empty()) {
        Builder b(boxes, out.order, out.nodes, leaf_size, leaf_most);
        b.split(0, boxes.size());
    }
    return out;
}

// ---- instance pose (crates/trace/src/detail.rs:184-287) ---------------------
struct Pose {
    float m[4][4];  // m[column][row]
    double a(int row, int column) const { return (double)m[column][row]; }
    double cof(int row, int column) const {
        int r0 = (row + 1) % 3, r1 = (row + 2) % 3, c0 = (column + 1) % 3, c1 = (column + 2) % 3;
        return a(r0, c0) * a(r1, c1) - a(r0, c1) * a(r1, c0);
    }
    double det() const {
        double d = -0.0;
        for (int c = 0; c < 3; ++c) d = d + a(0, c) * cof(0, c);
        return d;
    }
    void point(const float in[3], float out[3]) const {
        for (int row = 0; row < 3; ++row) {
            float s = -0.0f;
            for (int c = 0; c < 3; ++c) s = s + m[c][row] * in[c];
            out[row] = m[3][row] + s;
        }
    }
    void inverse_rows(float r[12]) const {
        double d = det();
        double inv[3][3];
        for (int row = 0; row < 3; ++row)
            for (int c = 0; c < 3; ++c) inv[row][c] = cof(c, row) / d;
        double tr[3] = {m[3][0], m[3][1], m[3][2]};
        for (int row = 0; row < 3; ++row) {
            double shift = -0.0;
            for (int k = 0; k < 3; ++k) shift = shift + inv[row][k] * tr[k];
            for (int c = 0; c < 3; ++c) r[row * 4 + c] = (float)inv[row][c];
            r[row * 4 + 3] = (float)(-shift);
        }
    }
    float inverse_determinant() const {
        double d = a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) - a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
                   a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
        return (float)(1.0 / d);
    }
};

std::string make_pose(const float t[16], Pose& p) {
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(t[i])) return "instance transform is not finite";
    memcpy(p.m, t, sizeof p.m);
    if (p.m[0][3] != 0.0f || p.m[1][3] != 0.0f || p.m[2][3] != 0.0f || p.m[3][3] != 1.0f)
        return "instance transform is not affine";
    double volume = 1.0;
    for (int c = 0; c < 3; ++c) {
        double s = -0.0;
        for (int r = 0; r < 3; ++r) s = s + p.a(r, c) * p.a(r, c);
        volume = volume * std::sqrt(s);
    }
    if (std::fabs(p.det()) <= volume * 1e-6) return "instance transform is singular";
    return "";
}

struct Tri9 {
    float v[9];
    float n[9];
    uint32_t material;
};

std::vector<Box> boxes_of(const std::vector<Tri9>& tris) {
    std::vector<Box> boxes(tris.size());
    for (size_t i = 0; i < tris.size(); ++i) {
        boxes[i].empty();
        for (int k = 0; k < 3; ++k) boxes[i].point(tris[i].v + 3 * k);
    }
    return boxes;
}

bool finite_n(const float* p, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// Everything levels::lay produces, kept for conversion.
struct Laid {
    RefLevels ref;
    std::vector<Tri9> world;
    RefBvh world_bvh;
    struct Mesh {
        uint32_t source;
        RefBvh bvh;
        uint32_t root, first;
        Box bounds;
    };
    std::vector<Mesh> meshes;
    struct Placed {
        uint32_t instance, mesh, base, record, word, material;
        Box bounds;
        float inverse[12];
        float inv_det;
    };
    std::vector<Placed> placed;
    std::vector<Tri9> mesh_tris(const pfxh_scene& sc, uint32_t m) const {
        const pfxh_mesh& mm = sc.meshes[m];
        std::vector<Tri9> ts(mm.count);
        for (uint32_t j = 0; j < mm.count; ++j) {
            const pfxh_triangle_detail& d = sc.mesh_triangles[mm.first + j];
            memcpy(ts[j].v, d.vertices, 36);
            memcpy(ts[j].n, d.normals, 36);
            ts[j].material = d.material;
        }
        return ts;
    }
};

std::string lay(const pfxh_scene& sc, Laid& out) {
    out.world.clear();
    for (uint32_t i = 0; i < sc.triangle_count; ++i) {
        Tri9 t;
        memcpy(t.v, sc.triangles[i].vertices, 36);
        memset(t.n, 0, 36);
        t.material = sc.triangles[i].material;
        out.world.push_back(t);
    }
    for (uint32_t i = 0; i < sc.world_detail_count; ++i) {
        Tri9 t;
        memcpy(t.v, sc.world_details[i].vertices, 36);
        memcpy(t.n, sc.world_details[i].normals, 36);
        t.material = sc.world_details[i].material;
        out.world.push_back(t);
    }
    for (auto& t : out.world)
        if (!finite_n(t.v, 9) || !finite_n(t.n, 9) || t.material >= sc.material_count) return "invalid trace geometry";
    for (uint32_t j = 0; j < sc.mesh_triangle_count; ++j) {
        const pfxh_triangle_detail& d = sc.mesh_triangles[j];
        if (!finite_n(d.vertices, 9) || !finite_n(d.normals, 9) || d.material >= sc.material_count)
            return "invalid trace geometry";
    }
    for (uint32_t m = 0; m < sc.mesh_count; ++m)
        if ((uint64_t)sc.meshes[m].first + sc.meshes[m].count > sc.mesh_triangle_count) return "mesh range is outside mesh_triangles";
    out.world_bvh = build_over(boxes_of(out.world), 4, 16);
    uint32_t next = (uint32_t)out.world.size();
    uint32_t record = 0;
    std::vector<std::pair<uint32_t, size_t>> known;  // (caller mesh, laid mesh)
    for (uint32_t i = 0; i < sc.instance_count; ++i) {
        const pfxh_instance& inst = sc.instances[i];
        if (inst.mesh >= sc.mesh_count) return "instance names a missing mesh";
        if (inst.word & ~(1u | 2u | (1u << 23))) return "unsupported instance surface word";
        if (sc.meshes[inst.mesh].count == 0) continue;
        std::vector<Tri9> ts = out.mesh_tris(sc, inst.mesh);
        Pose pose;
        std::string e = make_pose(inst.transform, pose);
        if (!e.empty()) return e;
        uint32_t material = ts[0].material;
        for (auto& t : ts)
            if (t.material != material) material = 0xffffffffu;
        long found = -1;
        for (auto& k : known)
            if (k.first == inst.mesh) {
                found = (long)k.second;
                break;
            }
        if (found < 0) {
            // bit-equal geometry shares a bottom level (levels.rs:284-303)
            for (auto& k : known) {
                const Laid::Mesh& mesh = out.meshes[k.second];
                const pfxh_mesh& other = sc.meshes[mesh.source];
                if (other.count != ts.size()) continue;
                bool same = true;
                for (size_t j = 0; j < ts.size() && same; ++j) {
                    const pfxh_triangle_detail& o = sc.mesh_triangles[other.first + j];
                    same = memcmp(o.vertices, ts[j].v, 36) == 0 && memcmp(o.normals, ts[j].n, 36) == 0 &&
                           (material != 0xffffffffu || o.material == ts[j].material);
                }
                if (same) {
                    found = (long)k.second;
                    break;
                }
            }
        }
        size_t mesh_index;
        if (found >= 0) {
            mesh_index = (size_t)found;
        } else {
            Laid::Mesh m;
            m.source = inst.mesh;
            m.bvh = build_over(boxes_of(ts), 4, 16);
            m.root = 0;
            m.first = next;
            for (int k = 0; k < 3; ++k) {
                m.bounds.lo[k] = m.bvh.nodes[0].lo[k];
                m.bounds.hi[k] = m.bvh.nodes[0].hi[k];
            }
            out.meshes.push_back(std::move(m));
            if ((uint64_t)next + ts.size() > 0xffffffffull) return "trace scene has too many triangles";
            next += (uint32_t)ts.size();
            mesh_index = out.meshes.size() - 1;
            known.emplace_back(inst.mesh, mesh_index);
        }
        record += 1;
        Laid::Placed p;
        p.instance = i;
        p.mesh = (uint32_t)mesh_index;
        p.base = out.meshes[mesh_index].first;
        p.record = record;
        p.word = inst.word;
        p.material = material;
        p.bounds.empty();
        const Box& local = out.meshes[mesh_index].bounds;
        for (int corner = 0; corner < 8; ++corner) {
            float pt[3], w[3];
            for (int axis = 0; axis < 3; ++axis) pt[axis] = ((corner >> axis) & 1) == 0 ? local.lo[axis] : local.hi[axis];
            pose.point(pt, w);
            p.bounds.point(w);
        }
        pose.inverse_rows(p.inverse);
        p.inv_det = pose.inverse_determinant();
        out.placed.push_back(p);
    }
    // the reference's node buffer: top level, world, meshes (levels.rs:337-366)
    RefLevels& ref = out.ref;
    ref.world_order = out.world_bvh.order;
    ref.nodes.clear();
    if (out.placed.empty() && sc.instance_count == 0) {
        ref.nodes = out.world_bvh.nodes;
        return "";
    }
    bool has_world = !out.world_bvh.nodes.empty();
    size_t groups = (has_world ? 1 : 0) + out.placed.size();
    size_t top_len = groups == 0 ? 0 : 3 * groups - 1;
    uint32_t offset = (uint32_t)top_len;
    uint32_t world_root = offset;
    std::vector<pfxh_bvh_node> world_nodes = out.world_bvh.nodes;
    if (has_world) {
        for (auto& n : world_nodes)
            if (n.count == 0) n.first += offset;
        offset += (uint32_t)world_nodes.size();
    }
    std::vector<std::vector<pfxh_bvh_node>> mesh_nodes;
    for (auto& m : out.meshes) {
        m.root = offset;
        std::vector<pfxh_bvh_node> ns = m.bvh.nodes;
        for (auto& n : ns)
            if (n.count == 0) n.first += offset;
        offset += (uint32_t)ns.size();
        mesh_nodes.push_back(std::move(ns));
    }
    if (groups > 0) {
        std::vector<Box> boxes;
        if (has_world) {
            Box b;
            for (int k = 0; k < 3; ++k) {
                b.lo[k] = out.world_bvh.nodes[0].lo[k];
                b.hi[k] = out.world_bvh.nodes[0].hi[k];
            }
            boxes.push_back(b);
        }
        for (auto& p : out.placed) boxes.push_back(p.bounds);
        RefBvh top = build_over(boxes, 1, 0);
        uint32_t tree = (uint32_t)top.nodes.size();
        for (pfxh_bvh_node n : top.nodes) {
            if (n.count > 0) {
                uint32_t index = top.order[n.first];
                bool world = has_world && index == 0;
                const Laid::Placed* p = world ? nullptr : &out.placed[index - (has_world ? 1 : 0)];
                n.first = world ? world_root : out.meshes[p->mesh].root;
                n.count = world ? 0 : p->base;
                n.lo[3] = pfx::u2f(world ? 0u : p->record);
                n.pad[0] = 1u;
                n.pad[1] = world ? 0u : tree + index;
            }
            ref.nodes.push_back(n);
        }
        for (size_t g = 0; g < groups; ++g) {
            float r[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
            bool world = has_world && g == 0;
            if (!world) memcpy(r, out.placed[g - (has_world ? 1 : 0)].inverse, sizeof r);
            pfxh_bvh_node n;
            memcpy(n.lo, r, 16);
            memcpy(n.hi, r + 4, 16);
            n.first = pfx::f2u(r[8]);
            n.count = pfx::f2u(r[9]);
            n.pad[0] = pfx::f2u(r[10]);
            n.pad[1] = pfx::f2u(r[11]);
            ref.nodes.push_back(n);
        }
    }
    pfxh_bvh_node zero;
    memset(&zero, 0, sizeof zero);
    ref.nodes.resize(top_len, zero);
    ref.nodes.insert(ref.nodes.end(), world_nodes.begin(), world_nodes.end());
    for (auto& ns : mesh_nodes) ref.nodes.insert(ref.nodes.end(), ns.begin(), ns.end());
    return "";
}

// ---- conversion of a reference-format binary tree into Node4s ---------------
// Each Node4 stands for one binary interior node with both children expanded
// one more level, children kept in left-first depth-first order. A deferred
// slab test of each child at pop time then visits exactly the reference's
// leaves in the reference's order: slab entry distances are monotone in the
// box bounds, so a skipped intermediate box passes whenever a child inside
// it passes.
struct Collapse {
    std::vector<Node4>& out;
    const std::vector<pfxh_bvh_node>* nodes = nullptr;
    uint32_t (*leaf_ref)(const pfxh_bvh_node&, const void*) = nullptr;
    const void* ctx = nullptr;
    std::string* err = nullptr;
    uint32_t depth = 0;  // binary depth of the deepest leaf
    explicit Collapse(std::vector<Node4>& o) : out(o) {}

    bool inner(size_t n) const { return (*nodes)[n].count == 0; }
    // returns the Node4 index; *stack gets the worst-case entries this
    // subtree can hold on the traversal stack
    uint32_t build(size_t n, uint32_t level, uint32_t* stack) {
        uint32_t index = (uint32_t)out.size();
        out.emplace_back();
        size_t kids[4];
        uint32_t kid_level[4];
        int count = 0;
        const size_t two[2] = {n + 1, (*nodes)[n].first};
        for (size_t c : two) {
            if (inner(c)) {
                kids[count] = c + 1;
                kid_level[count++] = level + 2;
                kids[count] = (*nodes)[c].first;
                kid_level[count++] = level + 2;
            } else {
                kids[count] = c;
                kid_level[count++] = level + 1;
            }
        }
        uint32_t refs[4] = {REF_NONE, REF_NONE, REF_NONE, REF_NONE};
        uint32_t deepest = 0;
        for (int k = 0; k < count; ++k) {
            if (inner(kids[k])) {
                uint32_t sub = 0;
                refs[k] = build(kids[k], kid_level[k], &sub);
                deepest = std::max(deepest, sub);
            } else {
                refs[k] = leaf_ref((*nodes)[kids[k]], ctx);
                if (refs[k] == REF_NONE && err) *err = "bvh leaf does not fit its reference";
                depth = std::max(depth, kid_level[k]);
            }
        }
        // a node pushes every hit child but the first; a descendant's entries
        // stack on top of its later siblings'
        *stack = (uint32_t)(count - 1) + deepest;
        Node4& w = out[index];
        const float inf = std::numeric_limits<float>::infinity();
        for (int k = 0; k < 4; ++k) {
            bool live = k < count;
            const pfxh_bvh_node* c = live ? &(*nodes)[kids[k]] : nullptr;
            w.lo_x[k] = live ? c->lo[0] : inf;
            w.lo_y[k] = live ? c->lo[1] : inf;
            w.lo_z[k] = live ? c->lo[2] : inf;
            w.hi_x[k] = live ? c->hi[0] : inf;
            w.hi_y[k] = live ? c->hi[1] : inf;
            w.hi_z[k] = live ? c->hi[2] : inf;
            w.ref[k] = refs[k];
            w.pad[k] = 0;
        }
        return index;
    }
    uint32_t tree(const std::vector<pfxh_bvh_node>& ns, uint32_t (*leaf)(const pfxh_bvh_node&, const void*),
                  const void* c, std::string& e, uint32_t* stack) {
        *stack = 0;
        if (ns.empty()) return REF_NONE;
        nodes = &ns;
        leaf_ref = leaf;
        ctx = c;
        err = &e;
        if (!inner(0)) {
            depth = std::max(depth, 0u);
            return leaf(ns[0], c);
        }
        return build(0, 0, stack);
    }
};

uint32_t triangle_leaf(const pfxh_bvh_node& n, const void*) {
    if (n.count == 0 || n.count > 16 || n.first > REF_FIRST_MASK) return REF_NONE;
    return REF_LEAF | ((n.count - 1u) << 27) | n.first;
}
// top-level leaves hold one group each; groups are stored in the top tree's
// depth-first leaf order, so a leaf's position there is its group index
uint32_t group_leaf(const pfxh_bvh_node& n, const void*) {
    if (n.count != 1 || n.first > REF_INDEX_MASK) return REF_NONE;
    return REF_INST | n.first;
}

float to_radians(float deg) { return deg * (PI_F / 180.0f); }

v3 rnormalize(v3 a) {
    float len = sqrtf(pfx::dot(a, a));
    if (len <= 1e-12f) return pfx::mk(0.0f, 0.0f, 1.0f);
    return a * (1.0f / len);
}
v3 ld3(const float* p) { return pfx::mk(p[0], p[1], p[2]); }

// lights.rs:233-238
void spot_terms(float inner_deg, float outer_deg, float* scale, float* offset) {
    float inner = fm_cos(to_radians(inner_deg));
    float outer = fm_cos(to_radians(outer_deg));
    *scale = 1.0f / pfx::maxf(inner - outer, 0.001f);
    *offset = -outer * *scale;
}
// LocalLight::area (lights.rs:146-152)
float light_area(const pfxh_light& l) {
    if (l.kind == PFXH_LIGHT_RECT) return 4.0f * pfx::length(pfx::cross(ld3(l.half_u), ld3(l.half_v)));
    if (l.kind == PFXH_LIGHT_DISC) return 3.14159265358979f * l.disc_radius * l.disc_radius;
    return 0.0f;
}
// lights.rs:183-231, 16 floats: a, b, c, d
void pack_light(const pfxh_light& l, float* g) {
    memset(g, 0, 64);
    if (l.kind == PFXH_LIGHT_RECT || l.kind == PFXH_LIGHT_DISC) {
        // a = centre + kind, b = power / area + FLT_MAX, c = half_u or the
        // normal + two-sidedness, d = half_v or the radius + shadow
        float area = light_area(l);
        for (int k = 0; k < 3; ++k) {
            g[k] = l.position[k];
            g[4 + k] = (l.colour[k] * l.intensity) / area;
        }
        g[3] = l.kind == PFXH_LIGHT_RECT ? 2.0f : 3.0f;
        g[7] = FLT_MAX;
        g[11] = l.two_sided ? 1.0f : 0.0f;
        g[15] = l.shadow ? 1.0f : 0.0f;
        if (l.kind == PFXH_LIGHT_RECT) {
            for (int k = 0; k < 3; ++k) {
                g[8 + k] = l.half_u[k];
                g[12 + k] = l.half_v[k];
            }
        } else {
            v3 nrm = rnormalize(ld3(l.direction));
            g[8] = nrm.x;
            g[9] = nrm.y;
            g[10] = nrm.z;
            g[12] = l.disc_radius;
        }
        return;
    }
    for (int k = 0; k < 3; ++k) {
        g[k] = l.position[k];
        g[4 + k] = l.colour[k] * l.intensity;
    }
    g[7] = l.range;
    g[11] = l.radius;
    g[15] = l.shadow ? 1.0f : 0.0f;
    if (l.kind == PFXH_LIGHT_SPOT) {
        float scale, offset;
        spot_terms(l.inner_deg, l.outer_deg, &scale, &offset);
        v3 axis = rnormalize(ld3(l.direction));
        g[3] = 1.0f;
        g[8] = axis.x;
        g[9] = axis.y;
        g[10] = axis.z;
        g[12] = scale;
        g[13] = offset;
    }
}
bool lit_light(const pfxh_light& l) {
    return l.intensity > 0.0f && l.range > 0.0f && (l.colour[0] > 0.0f || l.colour[1] > 0.0f || l.colour[2] > 0.0f);
}

// ---- tree.rs ------------------------------------------------------------------
struct Cone {
    v3 axis;
    float theta_o, theta_e;
    bool empty() const { return axis.x == 0.0f && axis.y == 0.0f && axis.z == 0.0f; }
    float measure() const {  // tree.rs:28-41
        if (empty()) return 0.0f;
        float theta_w = pfx::minf(theta_o + theta_e, PI_F);
        float cos_o = fm_cos(theta_o), sin_o = fm_sin(theta_o);
        return TAU_F * (1.0f - cos_o) +
               HALF_PI * (2.0f * theta_w * sin_o - fm_cos(theta_o - 2.0f * theta_w) - 2.0f * theta_o * sin_o + cos_o);
    }
};
const Cone CONE_EMPTY = {{0.0f, 0.0f, 0.0f}, FLT_MIN, FLT_MIN};
Cone merge(Cone self, Cone other) {  // tree.rs:43-81
    if (self.empty()) return other;
    if (other.empty()) return self;
    Cone a = self, b = other;
    if (other.theta_o > self.theta_o) {
        a = other;
        b = self;
    }
    float cos_a_b = pfx::dot(a.axis, b.axis);
    float theta_d = fm_acos(pfx::fm_clamp(cos_a_b, -1.0f, 1.0f));
    float theta_e = pfx::maxf(a.theta_e, b.theta_e);
    if (a.theta_o + 5e-4f >= pfx::minf(PI_F, theta_d + b.theta_o)) return Cone{a.axis, a.theta_o, theta_e};
    float theta_o = (theta_d + a.theta_o + b.theta_o) * 0.5f;
    if (theta_o >= PI_F) return Cone{a.axis, PI_F, theta_e};
    v3 axis;
    if (cos_a_b < -0.9995f) {
        v3 helper = fabsf(a.axis.x) > 0.1f ? pfx::mk(0.0f, 1.0f, 0.0f) : pfx::mk(1.0f, 0.0f, 0.0f);
        axis = rnormalize(pfx::cross(a.axis, helper));
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
    static Measure empty() {
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
    void add(const Measure& o) {  // tree.rs:111-122
        if (o.energy == 0.0f) return;
        for (int k = 0; k < 3; ++k) {
            min[k] = pfx::minf(min[k], o.min[k]);
            max[k] = pfx::maxf(max[k], o.max[k]);
        }
        cone = merge(cone, o.cone);
        energy += o.energy;
        range = pfx::maxf(range, o.range);
    }
    float calculate() const {  // tree.rs:124-132
        if (energy == 0.0f) return 0.0f;
        float s0 = max[0] - min[0], s1 = max[1] - min[1], s2 = max[2] - min[2];
        float area = 2.0f * (s0 * s1 + s1 * s2 + s2 * s0);
        float area_measure = area == 0.0f ? sqrtf(s0 * s0 + s1 * s1 + s2 * s2) : area;
        return energy * area_measure * cone.measure();
    }
};
struct LightEntry {
    size_t index;
    float centroid[3];
    Measure measure;
};
struct Bucket {
    bool has = false;
    Measure measure = Measure::empty();
    size_t count = 0;
    void add(const Measure& m) {
        if (!has) {
            has = true;
            measure = Measure::empty();
        }
        measure.add(m);
        count += 1;
    }
    Bucket plus(const Bucket& o) const {
        Bucket r;
        r.has = true;
        r.measure = has ? measure : Measure::empty();
        if (o.has) r.measure.add(o.measure);
        r.count = count + o.count;
        return r;
    }
    float cost() const { return has ? measure.calculate() : 0.0f; }
};
constexpr size_t LOOPED = 4, LEAF = 8, BUCKETS = 12;

// IEEE total order (f32::total_cmp)
bool total_less(float a, float b) {
    int32_t x = (int32_t)pfx::f2u(a), y = (int32_t)pfx::f2u(b);
    x ^= (int32_t)((uint32_t)(x >> 31) >> 1);
    y ^= (int32_t)((uint32_t)(y >> 31) >> 1);
    return x < y;
}

Measure should_split(const LightEntry* e, size_t count, bool* split, size_t* middle_out, int* dim_out) {
    *dim_out = -1;
    if (count < 2) {
        *split = false;
        return e[0].measure;
    }
    size_t middle = count / 2;
    float low[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, high[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (size_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            low[k] = pfx::minf(low[k], e[i].centroid[k]);
            high[k] = pfx::maxf(high[k], e[i].centroid[k]);
        }
    float extent[3] = {high[0] - low[0], high[1] - low[1], high[2] - low[2]};
    float max_extent = pfx::maxf(pfx::maxf(pfx::maxf(extent[0], extent[1]), extent[2]), 0.0f);
    Measure measure = Measure::empty();
    float total_cost = 0.0f, min_cost = FLT_MAX;
    int split_dim = -1;
    for (int dim = 0; dim < 3; ++dim) {
        Bucket buckets[BUCKETS];
        float inv_extent;
        if (extent[dim] == 0.0f) {
            if (dim != 0) continue;
            inv_extent = FLT_MAX;
            for (size_t i = 0; i < count; ++i) buckets[0].add(e[i].measure);
        } else {
            inv_extent = 1.0f / extent[dim];
            for (size_t i = 0; i < count; ++i) {
                float f = (float)BUCKETS * (e[i].centroid[dim] - low[dim]) * inv_extent;
                long long at = (f != f) ? 0 : (long long)f;
                at = at < 0 ? 0 : (at > (long long)BUCKETS - 1 ? (long long)BUCKETS - 1 : at);
                buckets[at].add(e[i].measure);
            }
        }
        Bucket left[BUCKETS - 1];
        left[0] = buckets[0];
        for (size_t i = 1; i < BUCKETS - 1; ++i) left[i] = left[i - 1].plus(buckets[i]);
        if (dim == 0) {
            Bucket all = left[BUCKETS - 2].plus(buckets[BUCKETS - 1]);
            measure = all.has ? all.measure : Measure::empty();
            if (extent[0] == 0.0f && extent[1] == 0.0f && extent[2] == 0.0f) break;
            total_cost = measure.calculate();
            if (total_cost == 0.0f) break;
            if (extent[dim] == 0.0f) continue;
        }
        Bucket right[BUCKETS - 1];
        right[BUCKETS - 2] = buckets[BUCKETS - 1];
        for (int i = (int)BUCKETS - 3; i >= 0; --i) right[i] = right[i + 1].plus(buckets[i + 1]);
        float regularization = max_extent * inv_extent;
        for (size_t s = 0; s < BUCKETS - 1; ++s) {
            float cost = regularization * (left[s].cost() + right[s].cost());
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
void tree_row(const Measure& m, size_t count, size_t link, float* g) {
    for (int k = 0; k < 3; ++k) {
        g[k] = m.min[k];
        g[4 + k] = m.max[k];
    }
    g[3] = m.energy;
    g[7] = m.range;
    g[8] = m.cone.axis.x;
    g[9] = m.cone.axis.y;
    g[10] = m.cone.axis.z;
    g[11] = m.cone.theta_o;
    g[12] = m.cone.theta_e;
    g[13] = (float)count;
    g[14] = (float)link;
    g[15] = 0.0f;
}
// Rust's slice::select_nth_unstable_by (rust-version 1.97, Cargo.toml:7;
// library/core/src/slice/select.rs with sort/shared/pivot.rs,
// sort/unstable/quicksort.rs and sort/shared/smallsort.rs), which the tree
// build uses to split its emitters (tree.rs:281-283). The order it leaves on
// each side becomes the leaf order, so the same moves are made here:
// introselect with at most 16 partition rounds on pseudo-median pivots
// (median of three below 64 elements, recursive ninther above), insertion
// sort at 16 elements or fewer, the branchless cyclic Lomuto partition (the
// 72-byte Rust element is under the 96-byte limit for it), a run equal to
// the previous pivot split off whole, and the median-of-medians fallback.
template <class E, class KeyOf>
struct Selector {
    E* v;
    KeyOf key;

    bool less(const E& a, const E& b) const { return total_less(key(a), key(b)); }
    // smallsort.rs: insertion sort of [lo, hi), shifting each new element left
    void insertion(size_t lo, size_t hi) const {
        for (size_t t = lo + 1; t < hi; ++t) {
            if (!less(v[t], v[t - 1])) continue;
            E held = v[t];
            size_t j = t;
            do {
                v[j] = v[j - 1];
                --j;
            } while (j > lo && less(held, v[j - 1]));
            v[j] = held;
        }
    }
    // pivot.rs median3 / median3_rec / choose_pivot, as indices
    size_t median3(size_t a, size_t b, size_t c) const {
        bool x = less(v[a], v[b]), y = less(v[a], v[c]);
        if (x != y) return a;
        return (less(v[b], v[c]) != x) ? c : b;
    }
    size_t ninther_rec(size_t a, size_t b, size_t c, size_t n) const {
        if (n * 8 >= 64) {
            size_t s = n / 8;
            a = ninther_rec(a, a + 4 * s, a + 7 * s, s);
            b = ninther_rec(b, b + 4 * s, b + 7 * s, s);
            c = ninther_rec(c, c + 4 * s, c + 7 * s, s);
        }
        return median3(a, b, c);
    }
    // (the pivot's position relative to lo)
    size_t pivot_of(size_t lo, size_t len) const {
        size_t e = len / 8;
        if (len < 64) return median3(lo, lo + 4 * e, lo + 7 * e) - lo;
        return ninther_rec(lo, lo + 4 * e, lo + 7 * e, e) - lo;
    }
    // quicksort.rs partition: pivot to the front, the cyclic Lomuto pass over
    // the rest (equal in effect to swap-Lomuto over it rotated left by one),
    // pivot to the boundary; `or_equal` partitions by !(pivot < x)
    size_t partition(size_t lo, size_t len, size_t p, bool or_equal) const {
        std::swap(v[lo], v[lo + p]);
        const E pivot = v[lo];
        auto goes_left = [&](const E& x) { return or_equal ? !less(pivot, x) : less(x, pivot); };
        size_t base = lo + 1, n = len - 1, lt = 0;
        if (n > 0) {
            E lifted = v[base];
            size_t gap = base;
            for (size_t r = 1; r <= n; ++r) {
                E next = r == n ? lifted : v[base + r];
                bool left = goes_left(next);
                v[gap] = v[base + lt];
                v[base + lt] = next;
                gap = base + r;  // unused after the last step
                lt += left ? 1 : 0;
            }
        }
        std::swap(v[lo], v[lo + lt]);
        return lt;
    }
    size_t first_min(size_t lo, size_t len) const {
        size_t m = lo;
        for (size_t i = lo + 1; i < lo + len; ++i)
            if (less(v[i], v[m])) m = i;
        return m;
    }
    size_t first_max(size_t lo, size_t len) const {
        size_t m = lo;
        for (size_t i = lo + 1; i < lo + len; ++i)
            if (less(v[m], v[i])) m = i;
        return m;
    }
    // select.rs median_idx / ninther / median_of_ninthers / median_of_medians
    size_t median_idx(size_t a, size_t b, size_t c) const {
        if (less(v[c], v[a])) std::swap(a, c);
        if (less(v[c], v[b])) return c;
        if (less(v[b], v[a])) return a;
        return b;
    }
    void ninther(size_t a, size_t b, size_t c, size_t d, size_t e, size_t f, size_t g, size_t h, size_t i) const {
        b = median_idx(a, b, c);
        h = median_idx(g, h, i);
        if (less(v[h], v[b])) std::swap(b, h);
        if (less(v[f], v[d])) std::swap(d, f);
        if (!less(v[e], v[d])) {
            if (less(v[f], v[e])) {
                d = f;
            } else {
                if (less(v[e], v[b]))
                    std::swap(v[e], v[b]);
                else if (less(v[h], v[e]))
                    std::swap(v[e], v[h]);
                return;
            }
        }
        if (less(v[d], v[b]))
            d = b;
        else if (less(v[h], v[d]))
            d = h;
        std::swap(v[d], v[e]);
    }
    size_t median_of_ninthers(size_t lo, size_t len) const {
        size_t frac = len <= 1024 ? len / 12 : (len <= 131072 ? len / 64 : len / 1024);
        size_t half = frac / 2, mlo = len / 2 - half, mhi = frac + mlo, gap = (len - 9 * frac) / 4;
        size_t a = mlo - 4 * frac - gap, b = mhi + gap;
        for (size_t i = mlo; i < mhi; ++i, a += 3, b += 3)
            ninther(lo + a, lo + i - frac, lo + b, lo + a + 1, lo + i, lo + b + 1, lo + a + 2, lo + i + frac,
                    lo + b + 2);
        median_of_medians(lo + mlo, frac, half);
        return partition(lo, len, mlo + half, false);
    }
    void median_of_medians(size_t lo, size_t len, size_t k) const {
        for (;;) {
            if (len <= 16) {
                insertion(lo, lo + len);
                return;
            }
            if (k == len - 1) {
                std::swap(v[first_max(lo, len)], v[lo + k]);
                return;
            }
            if (k == 0) {
                std::swap(v[first_min(lo, len)], v[lo]);
                return;
            }
            size_t p = median_of_ninthers(lo, len);
            if (p == k) return;
            if (p > k) {
                len = p;
            } else {
                lo += p + 1;
                len -= p + 1;
                k -= p + 1;
            }
        }
    }
    // select.rs partition_at_index + partition_at_index_loop
    void select(size_t len, size_t index) const {
        if (index >= len) return;
        if (index == len - 1) {
            std::swap(v[first_max(0, len)], v[index]);
            return;
        }
        if (index == 0) {
            std::swap(v[first_min(0, len)], v[0]);
            return;
        }
        size_t lo = 0;
        bool floor_set = false;
        E floor{};
        for (int rounds = 16;; --rounds) {
            if (len <= 16) {
                insertion(lo, lo + len);
                return;
            }
            if (rounds == 0) {
                median_of_medians(lo, len, index);
                return;
            }
            size_t p = pivot_of(lo, len);
            if (floor_set && !less(floor, v[lo + p])) {
                size_t mid = partition(lo, len, p, true) + 1;
                if (mid > index) return;
                lo += mid;
                len -= mid;
                index -= mid;
                floor_set = false;
                continue;
            }
            size_t mid = partition(lo, len, p, false);
            if (mid < index) {
                floor = v[lo + mid];
                floor_set = true;
                lo += mid + 1;
                len -= mid + 1;
                index -= mid + 1;
            } else if (mid > index) {
                len = mid;
            } else {
                return;
            }
        }
    }
};

struct CentroidOf {
    int dim;
    float operator()(const LightEntry& e) const { return e.centroid[dim]; }
};

// tree.rs:263-293
void grow(LightEntry* e, size_t count, size_t start, size_t base, size_t root, std::vector<float>& nodes) {
    size_t at = nodes.size() / 16;
    nodes.resize(nodes.size() + 16, 0.0f);
    bool split;
    size_t middle;
    int dim;
    Measure measure = should_split(e, count, &split, &middle, &dim);
    if (split) {
        if (dim >= 0)
            Selector<LightEntry, CentroidOf>{e, CentroidOf{dim}}.select(count, middle);
        grow(e, middle, start, base, root, nodes);
        size_t right_at = root + nodes.size() / 16;
        grow(e + middle, count - middle, start + middle, base, root, nodes);
        tree_row(measure, 0, right_at, &nodes[at * 16]);
    } else {
        tree_row(measure, count, base + start, &nodes[at * 16]);
    }
}

// sky_radiance for a lat-long HDR environment (bilinear, repeat u, clamp v)
v3 env_lookup(const std::vector<float>& tex, uint32_t W, uint32_t H, v3 direction) {
    if (W == 0 || H == 0) return pfx::mk(0, 0, 0);
    v3 d = pfx::normalize(direction);
    float u = pfx::fm_atan2(d.x, -d.z) / 6.28318531f + 0.5f;
    u = u - floorf(u);
    float v = fm_acos(pfx::fm_clamp(d.y, -1.0f, 1.0f)) / 3.14159265f;
    float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float tx = x - x0, ty = y - y0;
    auto at = [&](float ix, float iy, int k) {
        int w = (int)W, h = (int)H;
        int col = ((int)ix % w + w) % w;
        int row = (int)iy;
        row = row < 0 ? 0 : (row > h - 1 ? h - 1 : row);
        return tex[((size_t)row * W + col) * 4 + k];
    };
    float o[3];
    for (int k = 0; k < 3; ++k) {
        float a = at(x0, y0, k), b = at(x0 + 1.0f, y0, k), c = at(x0, y0 + 1.0f, k), dd = at(x0 + 1.0f, y0 + 1.0f, k);
        o[k] = (a * (1.0f - tx) + b * tx) * (1.0f - ty) + (c * (1.0f - tx) + dd * tx) * ty;
    }
    return pfx::mk(o[0], o[1], o[2]);
}
float lum(const float v[3]) { return 0.2126f * v[0] + 0.7152f * v[1] + 0.0722f * v[2]; }
bool emits(const float e[3]) {
    for (int k = 0; k < 3; ++k)
        if (!(std::isfinite(e[k]) && e[k] >= 0.0f)) return false;
    return lum(e) > 0.0f;
}
float tri_area(const float* v) {
    v3 c = pfx::cross(ld3(v + 3) - ld3(v), ld3(v + 6) - ld3(v));
    return 0.5f * sqrtf(pfx::dot(c, c));
}

}  // namespace

size_t select_keys(float* keys, uint32_t* ids, size_t len, size_t index, int mode) {
    struct KeyId {
        float key;
        uint32_t id;
    };
    struct KeyOf {
        float operator()(const KeyId& k) const { return k.key; }
    };
    std::vector<KeyId> v(len);
    for (size_t i = 0; i < len; ++i) v[i] = KeyId{keys[i], ids[i]};
    Selector<KeyId, KeyOf> sel{v.data(), KeyOf{}};
    size_t r = index;
    if (index < len) {
        if (mode == 1)
            r = sel.partition(0, len, index, false);
        else if (mode == 2)
            sel.median_of_medians(0, len, index);
        else
            sel.select(len, index);
    }
    for (size_t i = 0; i < len; ++i) {
        keys[i] = v[i].key;
        ids[i] = v[i].id;
    }
    return r;
}

RefBvh build_bvh(const float* v9, size_t count) {
    std::vector<Box> boxes(count);
    for (size_t i = 0; i < count; ++i) {
        boxes[i].empty();
        for (int k = 0; k < 3; ++k) boxes[i].point(v9 + 9 * i + 3 * k);
    }
    return build_over(boxes, 4, 16);
}

std::string lay_levels(const pfxh_scene& sc, RefLevels& out, std::vector<uint32_t>* slot_prim) {
    Laid laid;
    std::string e = lay(sc, laid);
    if (!e.empty()) return e;
    out = laid.ref;
    if (slot_prim) {
        slot_prim->clear();
        for (uint32_t j : laid.world_bvh.order) slot_prim->push_back(j);
        uint32_t mesh_base = sc.triangle_count + sc.world_detail_count;
        for (auto& m : laid.meshes)
            for (uint32_t j : m.bvh.order) slot_prim->push_back(mesh_base + sc.meshes[m.source].first + j);
    }
    return "";
}

bool light_valid(const pfxh_light& l) {  // lights.rs:101-142
    const float s[9] = {l.position[0], l.position[1], l.position[2], l.colour[0], l.colour[1], l.colour[2],
                        l.intensity, l.radius, l.range};
    if (!finite_n(s, 9)) return false;
    for (int k = 0; k < 3; ++k)
        if (l.colour[k] < 0.0f) return false;
    if (l.intensity < 0.0f || l.radius < 0.0f || l.range < 0.0f) return false;
    if (l.kind == PFXH_LIGHT_POINT) return true;
    if (l.kind == PFXH_LIGHT_RECT) {
        if (!finite_n(l.half_u, 3) || !finite_n(l.half_v, 3)) return false;
        float lu = pfx::length(ld3(l.half_u)), lv = pfx::length(ld3(l.half_v));
        return lu > 1e-6f && lv > 1e-6f && fabsf(pfx::dot(ld3(l.half_u), ld3(l.half_v))) <= 1e-4f * lu * lv;
    }
    if (l.kind == PFXH_LIGHT_DISC)
        return finite_n(l.direction, 3) && pfx::length(ld3(l.direction)) > 1e-6f && std::isfinite(l.disc_radius) &&
               l.disc_radius > 1e-6f;
    if (l.kind != PFXH_LIGHT_SPOT) return false;
    if (!finite_n(l.direction, 3)) return false;
    return pfx::length(ld3(l.direction)) > 1e-6f && std::isfinite(l.inner_deg) && std::isfinite(l.outer_deg) &&
           0.0f <= l.inner_deg && l.inner_deg <= l.outer_deg && l.outer_deg > 0.0f && l.outer_deg <= 90.0f;
}

void light_cones(const std::vector<float>& rows, uint32_t lights, std::vector<float>& out) {
    size_t count = rows.size() / 16;
    out.assign(count * 8, 0.0f);
    for (size_t r = 0; r < count; ++r) {
        const float* g = &rows[r * 16];
        float theta_o, theta_e;
        if (r < lights) {
            theta_o = 0.0f;
            theta_e = (uint32_t)g[3] == 1u ? pfx::fm_acos(pfx::fm_clamp(-g[13] / g[12], -1.0f, 1.0f))
                                           : 3.141592653589793f / 2.0f;
        } else {
            theta_o = g[11];
            theta_e = g[12];
        }
        float* c = &out[r * 8];
        c[0] = pfx::fm_cos(theta_o);
        c[1] = pfx::fm_sin(theta_o);
        c[2] = pfx::fm_cos(theta_o + theta_e);
        c[3] = pfx::fm_cos(theta_e - theta_o);
        c[4] = theta_o;
        c[5] = theta_e;
    }
}

void lay_lights(const pfxh_light* lights, size_t n, std::vector<float>& rows, uint32_t* looped, uint32_t* tree) {
    auto treed = [](const pfxh_light& l) {
        return lit_light(l) && (l.kind == PFXH_LIGHT_POINT || l.kind == PFXH_LIGHT_SPOT);
    };
    rows.clear();
    size_t treed_count = 0;
    for (size_t i = 0; i < n; ++i) treed_count += treed(lights[i]) ? 1 : 0;
    float g[16];
    if (treed_count <= LOOPED) {
        for (size_t i = 0; i < n; ++i) {
            pack_light(lights[i], g);
            rows.insert(rows.end(), g, g + 16);
        }
        *looped = (uint32_t)n;
        *tree = 0;
        return;
    }
    for (size_t i = 0; i < n; ++i)
        if (!treed(lights[i])) {
            pack_light(lights[i], g);
            rows.insert(rows.end(), g, g + 16);
        }
    size_t base = rows.size() / 16;
    std::vector<LightEntry> entries;
    for (size_t i = 0; i < n; ++i) {
        if (!treed(lights[i])) continue;
        const pfxh_light& l = lights[i];
        LightEntry e;
        e.index = i;
        Cone cone = l.kind == PFXH_LIGHT_SPOT
                        ? Cone{rnormalize(ld3(l.direction)), 0.0f, pfx::minf(to_radians(l.outer_deg), HALF_PI)}
                        : Cone{pfx::mk(0.0f, 0.0f, -1.0f), PI_F, HALF_PI};
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
        entries.push_back(e);
    }
    std::vector<float> nodes;
    grow(entries.data(), entries.size(), 0, base, base + entries.size(), nodes);
    for (auto& e : entries) {
        pack_light(lights[e.index], g);
        rows.insert(rows.end(), g, g + 16);
    }
    rows.insert(rows.end(), nodes.begin(), nodes.end());
    *looped = (uint32_t)base;
    *tree = (uint32_t)treed_count;
}

void build_emitters(const std::vector<float>& tris12, const std::vector<uint32_t>& records, uint32_t looped,
                    uint32_t tree, std::vector<float>& entries, std::vector<float>& cdf) {
    size_t n = records.size();
    std::vector<double> powers(n);
    double total = -0.0;
    for (size_t i = 0; i < n; ++i) {
        powers[i] = (double)tri_area(&tris12[12 * i]) * (double)lum(&tris12[12 * i + 9]);
        total = total + powers[i];
    }
    entries.assign(16 * (n + 1), 0.0f);
    cdf.assign(n, 0.0f);
    entries[0] = (float)n;
    entries[1] = (float)looped;
    entries[2] = (float)total;
    entries[3] = (float)tree;
    double running = 0.0;
    for (size_t i = 0; i < n; ++i) {
        running += powers[i];
        cdf[i] = (float)(running / total);
        float* e = &entries[16 * (i + 1)];
        const float* t = &tris12[12 * i];
        for (int k = 0; k < 3; ++k) {
            e[k] = t[k];
            e[4 + k] = t[3 + k];
            e[8 + k] = t[6 + k];
            e[12 + k] = t[9 + k];
        }
        e[3] = 0.0f;
        e[7] = (float)records[i];
        e[11] = tri_area(t);
        e[15] = (float)(powers[i] / total);
    }
    if (n) cdf[n - 1] = 1.0f;
}

// The environment texture and its sampling CDF, 256 x 128 (sky.rs:207-272,
// gpu.rs:1795); also what Trace::set_environment rebuilds (gpu.rs:2810-2815).
std::string lay_sky(const pfxh_environment& sky, HostScene& out) {
    out.env_w = sky.width;
    out.env_h = sky.height;
    out.env.clear();
    out.env_black = true;
    if (out.env_w) {
        out.env.assign(sky.texels, sky.texels + (size_t)out.env_w * out.env_h * 4);
        for (size_t i = 0; i < out.env.size(); ++i) {
            if (!std::isfinite(out.env[i])) return "the sky has non-finite texels";
            if ((i & 3) != 3 && out.env[i] != 0.0f) out.env_black = false;
        }
    } else {
        out.env.assign(4, 0.0f);
    }
    const uint32_t CW = 256, CH = 128;
    out.sky_rows.assign(CH, 0.0f);
    out.sky_columns.assign((size_t)CW * CH, 0.0f);
    double total = 0.0;
    for (uint32_t row = 0; row < CH; ++row) {
        float theta = PI_F * ((float)row + 0.5f) / (float)CH;
        float sine = fm_sin(theta);
        double row_total = 0.0;
        for (uint32_t col = 0; col < CW; ++col) {
            float phi = TAU_F * ((float)col + 0.5f) / (float)CW;
            v3 d = pfx::mk(fm_sin(theta) * fm_cos(phi), fm_cos(theta), fm_sin(theta) * fm_sin(phi));
            v3 c = env_lookup(out.env, out.env_w, out.env_h, d);
            float c3[3] = {c.x, c.y, c.z};
            row_total += (double)pfx::maxf(lum(c3), 0.0f) * (double)sine;
            out.sky_columns[(size_t)row * CW + col] = (float)row_total;
        }
        if (row_total > 0.0) {
            for (uint32_t col = 0; col < CW; ++col) out.sky_columns[(size_t)row * CW + col] /= (float)row_total;
        } else {
            for (uint32_t col = 0; col < CW; ++col) out.sky_columns[(size_t)row * CW + col] = (float)(col + 1) / (float)CW;
        }
        total += row_total;
        out.sky_rows[row] = (float)total;
    }
    if (total > 0.0) {
        for (auto& v : out.sky_rows) v /= (float)total;
    } else {
        for (uint32_t i = 0; i < CH; ++i) out.sky_rows[i] = (float)(i + 1) / (float)CH;
    }
    return "";
}

std::string build_scene(const pfxh_scene& sc, HostScene& out, bool emitter_sampling, bool watertight) {
    if (sc.material_count == 0) return "invalid trace scene detail";
    for (uint32_t i = 0; i < sc.light_count; ++i)
        if (!light_valid(sc.lights[i])) return "invalid trace light";
    if ((sc.sky.width == 0) != (sc.sky.height == 0)) return "the sky needs both a width and a height";
    if (sc.sky.width > 0 && !sc.sky.texels) return "the sky has no texels";
    Laid laid;
    std::string e = lay(sc, laid);
    if (!e.empty()) return e;
    out.ref = laid.ref;
    // materials
    out.materials.resize(sc.material_count);
    for (uint32_t m = 0; m < sc.material_count; ++m) {
        const pfxh_material& pm = sc.materials[m];
        MatDev& d = out.materials[m];
        memset(&d, 0, sizeof d);
        memcpy(d.base, pm.base, 12);
        d.roughness = pm.roughness;
        d.metalness = pm.metalness;
        d.specular = pm.specular;
        d.clearcoat = pm.clearcoat;
        d.clearcoat_roughness = pm.clearcoat_roughness;
        d.sheen = pm.sheen;
        d.fresnel_power = pm.fresnel_power;
        memcpy(d.emission, pm.emission, 12);
        d.emits = emits(pm.emission) ? 1u : 0u;
        if (!finite_n(pm.base, 3) || !finite_n(&pm.roughness, 7) || !finite_n(pm.emission, 3) ||
            !std::isfinite(pm.fresnel_power))
            return "invalid trace material";
    }
    // triangle slots: world in BVH order, then meshes in BVH order (gpu.rs:816-827)
    out.tris.clear();
    out.attrs.clear();
    out.slot_prim.clear();
    auto put = [&](const Tri9& t, uint32_t prim) {
        TriPos p;
        memset(&p, 0, sizeof p);
        TriAttr a;
        memset(&a, 0, sizeof a);
        for (int k = 0; k < 3; ++k) {
            p.a[k] = t.v[k];
            p.e1[k] = watertight ? t.v[3 + k] : t.v[3 + k] - t.v[k];
            p.e2[k] = watertight ? t.v[6 + k] : t.v[6 + k] - t.v[k];
            a.na[k] = t.n[k];
            a.nb[k] = t.n[3 + k];
            a.nc[k] = t.n[6 + k];
        }
        a.material = t.material;
        out.tris.push_back(p);
        out.attrs.push_back(a);
        out.slot_prim.push_back(prim);
    };
    for (uint32_t j : laid.world_bvh.order) put(laid.world[j], j);
    uint32_t mesh_base = sc.triangle_count + sc.world_detail_count;
    for (auto& m : laid.meshes) {
        std::vector<Tri9> ts = laid.mesh_tris(sc, m.source);
        for (uint32_t j : m.bvh.order) {
            if (emitter_sampling && out.materials[ts[j].material].emits)
                return "emissive instanced meshes are not supported";
            put(ts[j], mesh_base + sc.meshes[m.source].first + j);
        }
    }
    out.slots = (uint32_t)out.tris.size();
    // emitters: emissive world triangles in slot order (gpu.rs:866-889)
    std::vector<float> cand;
    std::vector<uint32_t> cand_rec;
    for (uint32_t slot = 0; slot < laid.world.size(); ++slot) {
        const Tri9& t = laid.world[laid.world_bvh.order[slot]];
        if (emitter_sampling && out.materials[t.material].emits && tri_area(t.v) > 1e-12f) {
            cand.insert(cand.end(), t.v, t.v + 9);
            cand.insert(cand.end(), sc.materials[t.material].emission, sc.materials[t.material].emission + 3);
            cand_rec.push_back(0);
            out.attrs[slot].emitter = (uint32_t)cand_rec.size();
        }
    }
    lay_lights(sc.lights, sc.light_count, out.lights, &out.looped, &out.tree);
    light_cones(out.lights, out.looped + out.tree, out.light_cones);
    build_emitters(cand, cand_rec, out.looped, out.tree, out.emitters, out.emitter_cdf);
    out.emitter_count = (uint32_t)cand_rec.size();
    out.lit = out.emitter_count > 0 || sc.light_count > 0;
    if (out.lights.empty()) out.lights.assign(16, 0.0f);
    if (out.emitter_cdf.empty()) out.emitter_cdf.assign(1, 1.0f);
    std::string e_sky = lay_sky(sc.sky, out);
    if (!e_sky.empty()) return e_sky;
    // device trees: world BLAS, mesh BLASes, then the top level
    out.nodes.clear();
    Collapse conv(out.nodes);
    std::string err;
    uint32_t bottom_stack = 0, sub = 0;
    uint32_t world_root = conv.tree(laid.world_bvh.nodes, triangle_leaf, nullptr, err, &sub);
    bottom_stack = std::max(bottom_stack, sub);
    std::vector<uint32_t> mesh_roots;
    for (auto& m : laid.meshes) {
        mesh_roots.push_back(conv.tree(m.bvh.nodes, triangle_leaf, nullptr, err, &sub));
        bottom_stack = std::max(bottom_stack, sub);
    }
    uint32_t bottom_depth = conv.depth;
    if (!err.empty()) return err;
    out.groups.clear();
    Surface world_surface;
    memset(&world_surface, 0, sizeof world_surface);
    world_surface.material = 0xffffffffu;
    world_surface.inv_det = 1.0f;
    float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    memcpy(world_surface.inv, ident, sizeof ident);
    std::vector<Surface> surfaces{world_surface};
    out.record_instance.assign(1, 0);
    memset(&out.root, 0, sizeof out.root);
    out.root.ref = REF_NONE;
    uint32_t top_depth = 0;
    if (sc.instance_count == 0) {
        if (!laid.world_bvh.nodes.empty()) {
            const pfxh_bvh_node& r = laid.world_bvh.nodes[0];
            for (int k = 0; k < 3; ++k) {
                out.root.lo[k] = r.lo[k];
                out.root.hi[k] = r.hi[k];
            }
            out.root.ref = world_root;
        }
        out.root.top = 0;
    } else {
        bool has_world = !laid.world_bvh.nodes.empty();
        std::vector<Box> boxes;
        if (has_world) {
            Group g;
            memset(&g, 0, sizeof g);
            const pfxh_bvh_node& r = laid.world_bvh.nodes[0];
            Box b;
            for (int k = 0; k < 3; ++k) {
                g.lo[k] = b.lo[k] = r.lo[k];
                g.hi[k] = b.hi[k] = r.hi[k];
            }
            g.root = world_root;
            g.base = 0;
            g.record = 0;
            g.word = 0;
            g.inv_det = 1.0f;
            g.has_xform = 0;
            memcpy(g.inv, ident, sizeof ident);
            out.groups.push_back(g);
            boxes.push_back(b);
        }
        for (auto& p : laid.placed) {
            Group g;
            memset(&g, 0, sizeof g);
            const Laid::Mesh& m = laid.meshes[p.mesh];
            for (int k = 0; k < 3; ++k) {
                g.lo[k] = m.bounds.lo[k];
                g.hi[k] = m.bounds.hi[k];
            }
            g.root = mesh_roots[p.mesh];
            g.base = p.base;
            g.record = p.record;
            g.word = p.word;
            g.inv_det = p.inv_det;
            g.has_xform = 1;
            memcpy(g.inv, p.inverse, sizeof g.inv);
            out.groups.push_back(g);
            boxes.push_back(p.bounds);
            Surface s;
            memset(&s, 0, sizeof s);
            s.word = p.word;
            s.material = p.material;
            s.inv_det = p.inv_det;
            memcpy(s.inv, p.inverse, sizeof s.inv);
            surfaces.push_back(s);
            out.record_instance.push_back(p.instance + 1);
        }
        if (!out.groups.empty()) {
            RefBvh top = build_over(boxes, 1, 0);
            // groups renumbered in the top tree's depth-first leaf order: the
            // traversal breaks ties at equal distance by group index, in the
            // order the reference's left-first walk meets the groups
            std::vector<Group> dfs(out.groups.size());
            for (size_t k = 0; k < top.order.size(); ++k) dfs[k] = out.groups[top.order[k]];
            out.groups.swap(dfs);
            Collapse top_conv(out.nodes);
            uint32_t top_stack = 0;
            uint32_t root = top_conv.tree(top.nodes, group_leaf, nullptr, err, &top_stack);
            if (!err.empty()) return err;
            top_depth = top_conv.depth;
            out.max_stack = top_stack + bottom_stack;
            for (int k = 0; k < 3; ++k) {
                out.root.lo[k] = top.nodes[0].lo[k];
                out.root.hi[k] = top.nodes[0].hi[k];
            }
            out.root.ref = root;
            out.root.top = 1;
        }
    }
    if (sc.instance_count == 0) out.max_stack = bottom_stack;
    out.watertight = watertight;
    out.bounds_tame = true;
    auto tame = [](float v) { return v == 0.0f || std::isinf(v) || (std::fabs(v) >= 0x1p-72f && std::fabs(v) <= 0x1p60f); };
    for (const Node4& n : out.nodes)
        for (int k = 0; k < 4; ++k)
            out.bounds_tame = out.bounds_tame && tame(n.lo_x[k]) && tame(n.lo_y[k]) && tame(n.lo_z[k]) &&
                              tame(n.hi_x[k]) && tame(n.hi_y[k]) && tame(n.hi_z[k]);
    for (const Group& g : out.groups)
        for (int k = 0; k < 3; ++k) out.bounds_tame = out.bounds_tame && tame(g.lo[k]) && tame(g.hi[k]);
    for (int k = 0; k < 3; ++k) out.bounds_tame = out.bounds_tame && tame(out.root.lo[k]) && tame(out.root.hi[k]);
    out.max_depth = bottom_depth + top_depth;
    // the reference's 64-entry stack never overflows on these trees (it
    // would drop subtrees), and ours must hold the worst case
    if (out.max_depth >= 60) return "the bvh is too deep for the reference's traversal stack";
    if (out.max_stack >= (uint32_t)STACK_LIMIT) return "the bvh is too deep for the traversal stack";
    out.surfaces = surfaces;
    if (out.nodes.empty()) out.nodes.resize(1);
    if (out.tris.empty()) {
        out.tris.resize(1);
        out.attrs.resize(1);
    }
    if (out.groups.empty()) out.groups.resize(1);
    quantize_nodes(out);
    return "";
}

// ---- quantized nodes ---------------------------------------------------------------
// A decoded plane value the walks can take: finite, not subnormal, and for a
// scene whose bounds are tame (0 or of magnitude in [2^-72, 2^60], which
// the conservative box test's error bound assumes) tame itself.
static bool decodable(float v, bool tame) {
    if (!std::isfinite(v) || (v != 0.0f && std::fabs(v) < FLT_MIN)) return false;
    return !tame || v == 0.0f || (std::fabs(v) >= 0x1p-72f && std::fabs(v) <= 0x1p60f);
}

// One Node4 as a QNode4, per axis: origin = the least lo of the node's
// children, scale = the least power of two with 255 scale at least their
// extent plus four units in the last place of the largest bound (room for
// the rounding of the decode), then per child the largest q_lo whose decoded
// value fmaf(q_lo, scale, origin) -- the FP32 fused multiply-add the device
// evaluates, correctly rounded on both -- is at most lo, and the least q_hi
// whose decoded value is at least hi. False when some value cannot be
// bounded or decoded so.
static bool quantize_node(const Node4& n, bool tame, QNode4& q) {
    memset(&q, 0, sizeof(q));
    for (int c = 0; c < 4; ++c) q.ref[c] = n.ref[c];
    const float* lo[3] = {n.lo_x, n.lo_y, n.lo_z};
    const float* hi[3] = {n.hi_x, n.hi_y, n.hi_z};
    for (int a = 0; a < 3; ++a) {
        float least = INFINITY, most = -INFINITY;
        for (int c = 0; c < 4; ++c) {
            if (n.ref[c] == REF_NONE) continue;
            if (!std::isfinite(lo[a][c]) || !std::isfinite(hi[a][c]) || lo[a][c] > hi[a][c]) return false;
            least = std::fmin(least, lo[a][c]);
            most = std::fmax(most, hi[a][c]);
        }
        if (!(least <= most)) return false;  // no child
        const float top = std::fmax(std::fabs(least), std::fabs(most));
        const double ulp = top > 0.0f ? std::ldexp(1.0, std::ilogb(top) - 23) : std::ldexp(1.0, -149);
        const double need = ((double)most - (double)least + 4.0 * ulp) / 255.0;
        int e = 0;
        const double f = std::frexp(need, &e);  // need = f 2^e, f in [0.5, 1)
        const int k = std::max(f == 0.5 ? e - 1 : e, -126);
        if (k > 127) return false;
        const float scale = std::ldexp(1.0f, k), origin = least;
        q.origin[a] = origin;
        q.scale[a] = scale;
        for (int c = 0; c < 4; ++c) {
            if (n.ref[c] == REF_NONE) continue;
            double ql = std::floor(((double)lo[a][c] - origin) / scale), qh = std::ceil(((double)hi[a][c] - origin) / scale);
            ql = std::min(std::max(ql, 0.0), 255.0);
            qh = std::min(std::max(qh, 0.0), 255.0);
            while (ql > 0.0 && std::fmaf((float)ql, scale, origin) > lo[a][c]) ql -= 1.0;
            while (qh < 255.0 && std::fmaf((float)qh, scale, origin) < hi[a][c]) qh += 1.0;
            const float dl = std::fmaf((float)ql, scale, origin), dh = std::fmaf((float)qh, scale, origin);
            if (dl > lo[a][c] || dh < hi[a][c] || !decodable(dl, tame) || !decodable(dh, tame)) return false;
            q.q_lo[a] |= (uint32_t)ql << (8 * c);
            q.q_hi[a] |= (uint32_t)qh << (8 * c);
        }
    }
    return true;
}

void quantize_nodes(HostScene& out) {
    out.qnodes.resize(out.nodes.size());
    for (size_t i = 0; i < out.nodes.size(); ++i) {
        if (!quantize_node(out.nodes[i], out.bounds_tame, out.qnodes[i])) {
            out.qnodes.clear();  // the walks then read the Node4s
            return;
        }
    }
}

}  // namespace pfxh
