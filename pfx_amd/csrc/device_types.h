// Device record formats, shared by the host layout code and the kernels.
// Every record is a whole number of 16-byte vectors so a wave fetches it
// with dwordx4 loads; node and triangle records are 128 and 48 bytes.
#pragma once
#include <stdint.h>

namespace pfxh {

// Child reference encoding used by Node4, RootEntry and Group.
//   inner node : index into the node array           (bits 31,30 clear)
//   triangles  : bit 31 | (count - 1) << 27 | first  (count <= 16, first < 2^27,
//                first relative to the group's triangle base)
//   instance   : bit 30 | group index                (top level only)
constexpr uint32_t REF_LEAF = 0x80000000u;
constexpr uint32_t REF_INST = 0x40000000u;
constexpr uint32_t REF_FIRST_MASK = 0x07ffffffu;
constexpr uint32_t REF_INDEX_MASK = 0x3fffffffu;
constexpr uint32_t REF_NONE = 0xffffffffu;

// Traversal stack entries per ray (LDS head + global spill); scene builds
// whose worst-case stack would not fit are refused.
constexpr int STACK_LIMIT = 96;

// A 4-wide node: the reference's binary BVH collapsed two levels at a time,
// children kept in the reference's left-first depth-first order. 128 bytes
// (one L2 line): the six child-box planes as 4-wide vectors, then the four
// child references. Empty slots have REF_NONE and a +inf box.
struct alignas(16) Node4 {
    float lo_x[4], lo_y[4], lo_z[4];
    float hi_x[4], hi_y[4], hi_z[4];
    uint32_t ref[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 is 128 bytes");

// A Node4 with its child boxes quantized outward to one byte per plane (64
// bytes, half a Node4; same index and child references as its Node4):
// plane value = fmaf(q, scale[a], origin[a]) in FP32 with origin the
// children's least lo and scale a power of two, each byte chosen so the
// decoded box contains the exact one. The default walks (whose box tests
// are conservative anyway) read these; the left-first walk keeps Node4s.
// q_lo[a] and q_hi[a] hold byte c for child c.
struct alignas(16) QNode4 {
    float origin[3];
    float scale[3];
    uint32_t q_lo[3];
    uint32_t q_hi[3];
    uint32_t ref[4];
};
static_assert(sizeof(QNode4) == 64, "QNode4 is 64 bytes");

// Positions in edge form: a, e1 = b - a, e2 = c - a (the same subtractions
// the Moller-Trumbore test makes, done once on the host). A scene built for
// the watertight test keeps the vertices themselves instead (e1 holds b, e2
// holds c): that test needs a vertex shared by two triangles to be the same
// bits in both.
struct alignas(16) TriPos {
    float a[3];
    float e1[3];
    float e2[3];
    float pad[3];
};
static_assert(sizeof(TriPos) == 48, "TriPos is 48 bytes");

// Shading attributes, fetched only for the closest hit.
struct alignas(16) TriAttr {
    float na[3];
    uint32_t material;
    float nb[3];
    uint32_t emitter;  // 1-based emitter row, 0 if the triangle does not emit
    float nc[3];
    uint32_t pad;
};
static_assert(sizeof(TriAttr) == 48, "TriAttr is 48 bytes");

// A top-level leaf: one placed mesh (or the world group).
struct alignas(16) Group {
    float lo[3];  // bottom-level root box, object space
    uint32_t root;
    float hi[3];
    uint32_t base;  // first triangle slot of the mesh
    uint32_t record;
    uint32_t word;
    float inv_det;
    uint32_t has_xform;
    float inv[12];  // world -> object rows
};
static_assert(sizeof(Group) == 96, "Group is 96 bytes");

// Per instance-surface record (record 0 is the world).
struct alignas(16) Surface {
    uint32_t word;
    uint32_t material;  // override, or 0xffffffff
    uint32_t pad;
    float inv_det;
    float inv[12];
};
static_assert(sizeof(Surface) == 64, "Surface is 64 bytes");

struct alignas(16) MatDev {
    float base[3];
    float roughness;
    float metalness, specular, clearcoat, clearcoat_roughness;
    float sheen, fresnel_power;
    float pad0, pad1;
    float emission[3];
    uint32_t emits;
};
static_assert(sizeof(MatDev) == 64, "MatDev is 64 bytes");

struct RootEntry {
    float lo[3];
    float hi[3];
    uint32_t ref;
    uint32_t top;  // 1 when ref is a top-level (instance) tree
};

}  // namespace pfxh
