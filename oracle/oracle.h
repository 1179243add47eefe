// CPU oracle for the MI355X tracer -- TEST INFRASTRUCTURE ONLY.
//
// A plain restatement of the reference tracer's still-render path
// (gmrdad82/pfx, crates/trace) used as the checker for the HIP kernels:
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are the
// only code allowed to load it. Every function cites the reference
// file:line it follows. Built with -ffp-contract=off and no fast-math.
//
// Parity status per part (pins in tests/test_oracle.py, test_oracle_pins.py,
// test_select.py, test_fmath.py, test_oracle_stats.py):
//   bvh ................. pinned: brute force on the reference's own case
//                         (bvh.rs:330-381, LCG seed 17), deterministic rebuild
//   levels .............. pinned: the reference's two-level test restated
//                         (levels.rs:591-681, LCG seed 29): 3 bottom builds,
//                         130 slots, 3,000 rays == brute force over the
//                         flattened triangles
//   emitter CDF ......... pinned: lights.rs:465-515 known answers
//   light pack/falloff .. pinned: lights.rs:578-628 known answers
//   light tree build .... pinned: tree.rs:433-492 invariants; the split
//                         restates Rust core's select_nth_unstable_by
//                         (select_nth.h), pinned by its selection
//                         properties; with distinct centroids the leaf
//                         order equals a stable-sort split's (so it does
//                         not hang on the pivots); pivot order beyond 16
//                         elements, which places tied centroids and
//                         orders the node sums, has no Rust output to
//                         compare: parity unpinned there
//   light tree pick ..... pinned: enumeration over the uniform (each
//                         light's share == its pdf; pdfs + failed share
//                         == 1) and a chi-squared test
//   PCG / offset_ray .... pinned: closed form, math.rs:80-105 known answers
//   adaptive rounds ..... pinned: adaptive.rs:391-461 known answers
//   radiance ............ pinned statistically: white furnace, emitter
//                         NEE vs BSDF-only, tree vs light-by-light, the
//                         closed-form direct term, FP64 converged means
//                         (gpu.rs:3787-3923, 4238-4560)
//   transcendentals ..... our Cephes-method functions (fmath.h), pinned
//                         against numpy; the reference's driver's bits are
//                         parity unpinned
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "../include/pfxhip.h"
#include "../pfx_amd/csrc/vec3.h"

namespace oracle {

using pfx::v3;

// bvh::Node (crates/trace/src/bvh.rs:5-13)
struct Node {
    float lo[4];
    float hi[4];
    uint32_t first;
    uint32_t count;
    uint32_t pad[2];
};
static_assert(sizeof(Node) == 48, "Node is 48 bytes");

// bvh::Bounds (bvh.rs:41-72)
struct Bounds {
    float lo[3];
    float hi[3];
};
Bounds bounds_empty();
void bounds_point(Bounds& b, const float p[3]);
void bounds_grow(Bounds& b, const Bounds& o);

struct Bvh {
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
};
// Bvh::try_build (bvh.rs:95-105) over triangle vertex arrays (9 floats each)
Bvh bvh_build(const float* vertices9, size_t count);
// Bvh::build_instances (bvh.rs:107-109)
Bvh bvh_build_instances(const std::vector<Bounds>& boxes);
// bvh::triangle_hit (bvh.rs:294-324); returns distance or -1
float triangle_hit(const float* v9, v3 o, v3 d, float limit);
// Bvh::intersect (bvh.rs:135-164); returns triangle index or UINT32_MAX
uint32_t bvh_intersect(const Bvh& bvh, const float* vertices9, v3 o, v3 d, float limit);

// gpu.rs:17-26 GpuTriangle
struct Tri {
    float a[4], b[4], c[4], na[4], nb[4], nc[4];
};
// gpu.rs:29-38 GpuSurface
struct Surface {
    float clip0[4], clip1[4], transform[4], crop[4];
    uint32_t info[4];
    float forward[12];
    float inverse[12];
};
// lights.rs:245-261
struct GLight {
    float a[4], b[4], c[4], d[4];
};
struct GEmitter {
    float a[4], b[4], c[4], e[4];
};

// Laid-out two-level scene (levels::lay + SceneBuffers::build_with)
struct Laid {
    std::vector<Node> nodes;
    std::vector<uint32_t> world_order;   // world triangle index per world slot
    std::vector<Tri> tris;               // all slots, BVH order
    std::vector<uint32_t> slot_prim;     // caller primitive index per slot
    std::vector<Surface> surfaces;       // [world, placed...]
    std::vector<uint32_t> instance_of_record;  // record -> caller instance index + 1
    uint32_t slots = 0;
};
std::string lay_scene(const pfxh_scene& scene, Laid& out);

// lights.rs pack/spot_terms/falloff/cone; tree.rs lay
GLight light_pack(const pfxh_light& l);
bool light_valid(const pfxh_light& l);
bool light_lit(const pfxh_light& l);
float light_falloff(const pfxh_light& l, float distance);
float light_cone(const pfxh_light& l, const float toward[3]);
struct LaidLights {
    std::vector<GLight> rows;
    uint32_t looped = 0;
    uint32_t tree = 0;
};
LaidLights tree_lay(const pfxh_light* lights, size_t count);
void set_tree_split_by_sort(bool on);

struct EmitterList {
    std::vector<GEmitter> entries;
    std::vector<float> cdf;
};
struct EmitterCandidate {
    float vertices[9];
    uint32_t record;
    float emission[3];
};
EmitterList build_emitters(const std::vector<EmitterCandidate>& c, size_t looped, size_t tree);
float luminance(const float v[3]);
bool emits(const float e[3]);

// sky.rs SkyCdf (256 x 128)
struct SkyCdf {
    uint32_t width = 0, height = 0;
    std::vector<float> rows, columns;
    float total = 0.0f;
};
struct Env {
    uint32_t width = 0, height = 0;
    std::vector<float> texels;
};
v3 env_radiance(const Env& env, v3 d);
SkyCdf sky_build(const Env& env, uint32_t width, uint32_t height);

// adaptive.rs rounds
std::vector<uint32_t> adaptive_rounds(const pfxh_adaptive& a);

}  // namespace oracle
