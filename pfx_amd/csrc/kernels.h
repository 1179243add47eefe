// Kernel parameter blocks and launch entry points of the wavefront tracer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_types.h"

namespace pfxh {

// Scene constants and read-only device buffers (passed by value).
// SceneDev::geo_flags: GEO_TAME when every box bound is 0 or has magnitude
// >= 2^-72 (the fast exact slab applies); GEO_WATERTIGHT when the triangles
// hold vertices and are tested watertight; GEO_LEFT_FIRST when closest-hit
// walks take the children of a node left first, as the reference does.
constexpr uint32_t GEO_TAME = 1u, GEO_WATERTIGHT = 2u, GEO_LEFT_FIRST = 4u;

struct SceneDev {
    const Node4* nodes;
    const QNode4* qnodes;    // the nodes quantized, or null (the default walks then read nodes)
    const TriPos* tris;
    const TriAttr* attrs;
    const Group* groups;
    const Surface* surfaces;
    const MatDev* materials;
    const float4* lights;    // GpuLight rows, 4 float4 each
    const float4* light_cones;  // per light row: (cos θo, sin θo, cos(θo+θe), cos(θe−θo)), (θo, θe, 0, 0)
    const float4* emitters;  // GpuEmitter rows, 4 float4 each
    const float* cdf;
    const float* sky_rows;
    const float* sky_columns;
    const float* env;
    const uint32_t* slot_prim;
    const uint32_t* record_instance;
    RootEntry root;
    uint32_t env_w, env_h, env_black, lit;
    uint32_t area;  // a looped light row is a shadowed rect or disc light: rays can hit it (light_hits)
    const float* sobol_table;  // the Sobol sampler's patterns (4 floats a point), or null for PCG
    uint32_t sobol_len;        // points per pattern
    uint32_t looped, tree, emitter_count, light_slots;
    uint32_t entries;  // NEE entries a vertex may queue
    uint32_t geo_flags;    // GEO_TAME | GEO_WATERTIGHT | GEO_LEFT_FIRST
    uint32_t total, diffuse, glossy, transmission;
    float sun_dir[3], sun_color[3], sun_intensity, sun_w;
    float origin[3], forward[3], right[3], up[3], lens[4];
    uint32_t projection;  // PFXH_PROJ_*
    float half_view[2];   // orthographic half width and half height
    uint32_t window[4];
    uint32_t width, height;
};

// One wave of paths: SoA state in HBM.
struct WaveDev {
    float4* o;      // origin.xyz, previous pdf
    float4* d;      // direction.xyz, lobe counters (bits)
    float4* thr;    // throughput.xyz, rng state (bits)
    float4* res;    // result.xyz
    float4* alb;    // first albedo
    float4* nrm;    // first normal
    float4* hit;    // t, u, v, slot (bits)
    uint32_t* hit_rec;
    float4* e_o;    // per entry: shadow origin.xyz, limit
    float4* e_d;    // per entry: direction.xyz
    float4* e_c;    // per entry: contribution.xyz
    uint32_t* q_in;
    uint32_t* q_next;
    uint32_t* q_sh;      // paths with NEE entries: index | entries << 28
    uint32_t* q_sray;    // shadow rays to test: path index | entry << 28
    uint32_t* counters;  // [0] next, [1] shadow paths, [2] shadow rays, [3..] stats
    uint32_t* sid;       // batch: the still of each path
    uint2* spill;        // traversal stack overflow
    uint32_t capacity;   // paths per wave (entry arrays are strided by it)
    uint32_t spill_threads;
    uint32_t refill;     // idle lanes that make a traversal wave fetch new rays (1..64)
};

// Accumulation buffers (per pixel).
struct AccumDev {
    float4* sum;   // color (w = count), albedo, normal: 3 float4 per pixel
    float4* comp;  // Neumaier compensation, 3 float4 per pixel
    double* sum64; // 12 doubles per pixel in FP64 mode
    float* m2;     // per pixel: the sum over samples of ((r + g + b) / 3)^2 (plain FP32)
    uint32_t mode;
};

// Device counters, each hot one on its own 1 KiB slot: atomics from all
// XCDs on one address (or one line) serialize, so counters that are bumped
// at the same time must not share a line.
constexpr uint32_t SLOT = 256;   // uint32 per slot
constexpr uint32_t PARTS = 8;    // work parts of a traversal queue, one per XCD
enum Counter : uint32_t {
    C_NEXT = 0,                        // next-bounce queue length
    C_SHADOW = SLOT,                   // shadow path queue length
    C_SRAYS = 2 * SLOT,                // shadow ray queue length
    C_RUNNING = 3 * SLOT,
    C_ACTIVE = 4 * SLOT,
    C_NODES = 5 * SLOT,                // 64-bit statistics
    C_PRIMS = 5 * SLOT + 2,
    C_SNODES = 5 * SLOT + 4,
    C_SPRIMS = 5 * SLOT + 6,
    C_BOXES = 5 * SLOT + 8,            // closest-hit child boxes tested
    C_WAVE = 5 * SLOT + 10,            // closest-hit wave rounds (6 x 64-bit, WaveCount order)
    C_FETCH = 8 * SLOT,                // PARTS work counters of k_extend
    C_SFETCH = C_FETCH + PARTS * SLOT, // PARTS work counters of k_shadow
    C_N = C_SFETCH + PARTS * SLOT,     // batch: rays, shadow rays, shadow paths of the bounce
    C_STATS = C_N + SLOT,              // batch: 64-bit closest-hit and any-hit ray counts
    C_HITS = C_STATS + SLOT,           // paths of the bounce whose ray hit (the shading queue)
    C_COUNT = C_HITS + SLOT
};

// A batch renders many stills (own scene, size, sample count and seed each)
// with one launch per stage. A wave of paths is a list of units, each a run
// of samples of one still laid out as a single trace lays out its wave
// (sample-major over the still's pixels), so every pixel's samples are
// folded in sample order and each still's sums are those of its own trace.
struct Unit {
    uint32_t sid;     // still
    uint32_t first;   // first path of the unit in the wave
    uint32_t pfirst;  // first pixel slot of the unit in the wave (k_finish_batch)
    uint32_t npix;    // pixels of the still
    uint32_t ns;      // samples of each pixel in this unit
    uint32_t acc;     // sample index of the first of them
    uint32_t seed;
    uint32_t pad;
};
constexpr uint32_t MAX_UNITS = 64;
struct WavePlan {
    uint32_t units, paths, pixels, pad;
    Unit u[MAX_UNITS];
};
struct BatchDev {
    const SceneDev* scenes;  // per still
    const AccumDev* accums;  // per still
    const WavePlan* waves;
    uint32_t* wave;          // the wave being traced (advanced by k_next_wave)
};

constexpr int TRACE_BLOCK = 128;
constexpr int STACK_SHORT = 11;  // LDS head per lane: 11 KiB per block, 14 blocks per CU fit 160 KiB
constexpr int STACK_DEPTH = STACK_LIMIT;

void launch_camera(hipStream_t s, const SceneDev& S, const WaveDev& W, const uint32_t* active, uint32_t npix,
                   uint32_t nsamples, uint32_t acc, uint32_t seed);
void launch_extend(hipStream_t s, const SceneDev& S, const WaveDev& W, const uint32_t* queue, uint32_t n,
                   bool count, uint32_t grid);
void launch_hits(hipStream_t s, const WaveDev& W, const uint32_t* queue, uint32_t n, uint32_t* hitq);
// Shades queue[0, n), or queue[0, *count) when count is given (n then bounds it).
void launch_shade(hipStream_t s, const SceneDev& S, const WaveDev& W, const uint32_t* queue, uint32_t n,
                  uint32_t bounce, const uint32_t* count = nullptr);
void launch_shadow(hipStream_t s, const SceneDev& S, const WaveDev& W, uint32_t n, bool count, uint32_t grid);
void launch_fold(hipStream_t s, const WaveDev& W, uint32_t n);
void launch_finish(hipStream_t s, const WaveDev& W, const AccumDev& A, const uint32_t* active, uint32_t npix,
                   uint32_t nsamples, float count_after);
void launch_develop(hipStream_t s, const AccumDev& A, uint32_t pixels, float4* color, float4* albedo,
                    float4* normal);
void launch_clear(hipStream_t s, const AccumDev& A, uint32_t pixels);
// The standard error of each pixel's mean luminance ((r + g + b) / 3) over its
// samples, sqrt(max(m2 / n - mean^2, 0) / n); 0 for a pixel with no samples.
void launch_deviation(hipStream_t s, const AccumDev& A, uint32_t pixels, float* out);
void launch_active(hipStream_t s, const AccumDev& A, uint32_t pixels, uint32_t* active, uint32_t* counter);
void launch_probe(hipStream_t s, const SceneDev& S, const WaveDev& W, void* out);
void launch_cast(hipStream_t s, const SceneDev& S, const WaveDev& W, const float4* o, const float4* d, uint32_t n,
                 float limit, uint32_t* ids, float* t);
void launch_stream_copy(hipStream_t s, const float4* src, float4* dst, uint64_t n, uint32_t blocks);
void launch_measure(hipStream_t s, const AccumDev& A, float4* snaps, uint32_t width, uint32_t height);
void launch_mark(hipStream_t s, const AccumDev& A, const float4* snaps, uint32_t width, uint32_t height,
                 uint32_t min_samples, float threshold, uint32_t* running);
void launch_zero(hipStream_t s, uint32_t* p, uint32_t n);

void launch_camera_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t grid);
void launch_roll(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t phase, uint32_t bounce);
void launch_extend_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, const uint32_t* queue, uint32_t grid,
                         bool left_first,
                         bool quantized);
void launch_hits_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, const uint32_t* queue, uint32_t* hitq,
                       uint32_t grid);
// Shades the wave's queue (length on the device), or queue[0, *count) when count is given.
void launch_shade_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, const uint32_t* queue, uint32_t bounce,
                        uint32_t grid, const uint32_t* count, bool sobol);
void launch_shadow_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t grid, bool left_first,
                         bool quantized);
void launch_fold_batch(hipStream_t s, const WaveDev& W, uint32_t grid);
void launch_finish_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t grid);
void launch_next_wave(hipStream_t s, const BatchDev& B);

}  // namespace pfxh
