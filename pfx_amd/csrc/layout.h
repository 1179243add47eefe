// Host-side scene layout for the MI355X tracer: the reference's CPU build
// (binned-SAH BVH, two-level instancing, light tree, emitter and sky CDFs)
// followed by conversion into the device formats the kernels read.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pfxhip.h"
#include "device_types.h"

namespace pfxh {

// Reference-format BVH (bvh::Node, crates/trace/src/bvh.rs:5-13)
struct RefBvh {
    std::vector<pfxh_bvh_node> nodes;
    std::vector<uint32_t> order;
};
RefBvh build_bvh(const float* vertices9, size_t count);

// levels::lay (crates/trace/src/levels.rs:225-383) in the reference node format
struct RefLevels {
    std::vector<pfxh_bvh_node> nodes;
    std::vector<uint32_t> world_order;
};

// Everything a trace needs on the device, still in host memory.
struct HostScene {
    // reference-format two-level layout (kept for the layout entry points)
    RefLevels ref;
    // device formats
    std::vector<Node4> nodes;
    std::vector<QNode4> qnodes;  // the nodes quantized (empty when some node cannot be: see quantize_nodes)
    std::vector<TriPos> tris;
    std::vector<TriAttr> attrs;
    std::vector<uint32_t> slot_prim;   // caller primitive index per triangle slot
    std::vector<Group> groups;         // top-level leaves
    std::vector<Surface> surfaces;     // one per instance-surface record (0 = world)
    std::vector<uint32_t> record_instance;  // surface record -> caller instance + 1
    std::vector<MatDev> materials;
    std::vector<float> lights;         // GpuLight rows, 16 floats each
    std::vector<float> light_cones;    // per light row: cos/sin of its cone angles, 8 floats (light_cones)
    std::vector<float> emitters;       // GpuEmitter rows, 16 floats each (header first)
    std::vector<float> emitter_cdf;
    std::vector<float> sky_rows, sky_columns;
    std::vector<float> env;            // lat-long RGBA32F texels
    uint32_t env_w = 0, env_h = 0;
    bool env_black = true;
    uint32_t looped = 0, tree = 0, emitter_count = 0;
    bool lit = false;
    RootEntry root;
    uint32_t slots = 0;
    uint32_t max_depth = 0;   // binary-tree depth (top + bottom)
    uint32_t max_stack = 0;   // worst-case traversal stack entries
    bool bounds_tame = true;  // every node box bound is 0 or >= 2^-72 in magnitude
    bool watertight = false;  // tris hold vertices a, b, c (watertight test), not a, b - a, c - a
};

// `emitter_sampling` false leaves emissive triangles unmarked, so they are
// reached only by BSDF sampling (gpu.rs:781).
std::string build_scene(const pfxh_scene& scene, HostScene& out, bool emitter_sampling = true,
                        bool watertight = false);
std::string lay_levels(const pfxh_scene& scene, RefLevels& out, std::vector<uint32_t>* slot_prim);
// The environment texels, black flag and sky sampling CDF of `out` from
// `sky` (a scene's sky, or a Trace::set_environment argument).
std::string lay_sky(const pfxh_environment& sky, HostScene& out);
// out.qnodes from out.nodes (device_types.h QNode4); left empty when some
// node's boxes cannot be bounded by decodable values (the walks then read
// the Node4s).
void quantize_nodes(HostScene& out);

// tree::lay (crates/trace/src/tree.rs:220-243)
// Rust's slice::select_nth_unstable_by as the light tree uses it, over
// (key, id) pairs compared by f32::total_cmp: mode 0 selects `index`, mode 1
// runs one partition around keys[index] and returns the number of smaller
// keys, mode 2 runs the median-of-medians fallback alone.
size_t select_keys(float* keys, uint32_t* ids, size_t len, size_t index, int mode);
// Per light row, the trigonometric terms the light tree's importance takes
// of the row's constant cone angles (lights.wgsl:204-237 tree_importance):
// cos θo, sin θo, cos(θo + θe), cos(θe − θo), θo, θe, 0, 0, with θo = c.w
// and θe = d.x for a tree node, and θo = 0 with θe = acos(−d.y / d.x) for a
// spot light or π/2 for a point light (lights.wgsl:255-271), each computed
// by the same functions the device would use, so the device reads them
// instead of recomputing them per shading point, bit for bit.
void light_cones(const std::vector<float>& rows, uint32_t lights, std::vector<float>& out);
void lay_lights(const pfxh_light* lights, size_t count, std::vector<float>& rows, uint32_t* looped,
                uint32_t* tree);
bool light_valid(const pfxh_light& l);
// lights::build_emitters (crates/trace/src/lights.rs:391-440)
void build_emitters(const std::vector<float>& tris12, const std::vector<uint32_t>& records, uint32_t looped,
                    uint32_t tree, std::vector<float>& entries, std::vector<float>& cdf);

}  // namespace pfxh
