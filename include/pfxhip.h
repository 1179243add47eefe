/* pfxhip.h -- C ABI of the MI355X path tracer (libpfxhip.so).
 *
 * It replaces the still-render surface of the reference tracer,
 * `pfx_trace::Trace` in crates/trace/src/gpu.rs, for a host that binds it
 * over FFI. Plain C types only: pointers, sizes, fixed-layout structs.
 *
 * Ownership: every pfxh_* call copies what it needs out of caller memory
 * before returning and never keeps a caller pointer. A pfxh_trace owns all
 * of its device memory and one HIP stream. Calls on one trace must be
 * serialised by the caller (the reference's `&mut self` methods); distinct
 * traces may be used from distinct threads.
 *
 * Errors: every fallible call returns a pfxh_status and, when `err` is not
 * NULL, writes a NUL-terminated message of at most `err_len` bytes. Build
 * failures are split like `TraceError::{Refused, Setup}`
 * (crates/trace/src/gpu.rs:268-307): PFXH_REFUSED names a device limit or
 * memory shortfall, PFXH_SETUP an invalid scene or a failed device call.
 */
#ifndef PFXHIP_H
#define PFXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pfxh_status {
    PFXH_OK = 0,
    PFXH_REFUSED = 1, /* TraceError::Refused: limits or memory (gpu.rs:1254-1295) */
    PFXH_SETUP = 2,   /* TraceError::Setup: invalid scene or device failure */
    PFXH_INVALID = 3, /* runtime Err(String) of a setter or sampler call */
    PFXH_DEVICE = 4   /* HIP runtime error during sampling or readback */
} pfxh_status;

/* ---- scene input (crates/trace/src/gpu.rs:249-256, detail.rs) ---------- */

/* bvh::Triangle (crates/trace/src/bvh.rs:15-19): a flat world triangle. */
typedef struct pfxh_triangle {
    float vertices[9]; /* a.xyz, b.xyz, c.xyz */
    uint32_t material;
} pfxh_triangle;

/* detail::TriangleDetail without uv sets: vertices plus per-vertex normals
 * (all-zero normals mean "use the geometric normal"). */
typedef struct pfxh_triangle_detail {
    float vertices[9];
    float normals[9];
    uint32_t material;
} pfxh_triangle_detail;

/* A shared mesh: a range of pfxh_scene.mesh_triangles. Instances naming
 * the same mesh share one bottom-level BVH, as instances sharing one
 * `Arc<[TriangleDetail]>` do in crates/trace/src/levels.rs:266-283. */
typedef struct pfxh_mesh {
    uint32_t first;
    uint32_t count;
} pfxh_mesh;

/* detail::Instance (crates/trace/src/detail.rs:53-60): a placed mesh.
 * transform is column-major 4x4 (`Transform = [[f32; 4]; 4]`), affine. */
typedef struct pfxh_instance {
    uint32_t mesh;
    uint32_t word; /* surface flags: 1 shadow-only, 2 casts no shadow, 1<<23 one-sided */
    float transform[16];
} pfxh_instance;

/* The diffuse + GGX subset of pfx_materials::Material
 * (crates/materials/src/material.rs:493-575) the tracer shades. */
typedef struct pfxh_material {
    float base[3];
    float roughness;
    float metalness;
    float specular;
    float clearcoat;
    float clearcoat_roughness;
    float sheen;
    float emission[3];
    float fresnel_power; /* <= 0 means 5 */
} pfxh_material;

/* lights::LocalLight (crates/trace/src/lights.rs:30-39) with its four
 * LightShape kinds (:10-28): point, spot, and the rect and disc area
 * lights. `position` is the centre of an area light; `colour * intensity`
 * is a point or spot light's intensity and an area light's power, spread
 * over its area (lights.rs:207-230). Area lights are never put in the light
 * tree: each takes its own light slot, and camera and bounce rays that hit
 * a shadowed one see its radiance (lights.wgsl:64-91). LocalLight::rect and
 * ::disc (lights.rs:55-99) set radius 0, range FLT_MAX, shadow 1 and
 * two_sided 0. */
enum { PFXH_LIGHT_POINT = 0, PFXH_LIGHT_SPOT = 1, PFXH_LIGHT_RECT = 2, PFXH_LIGHT_DISC = 3 };
typedef struct pfxh_light {
    float position[3];
    float colour[3];
    float intensity;
    float radius;
    float range;
    uint32_t shadow;
    uint32_t kind;       /* PFXH_LIGHT_* */
    float direction[3];  /* spot axis, or the disc's normal */
    float inner_deg;     /* spot */
    float outer_deg;     /* spot */
    float half_u[3];     /* rect: half of each side; perpendicular; emits along half_u x half_v */
    float half_v[3];
    float disc_radius;   /* disc */
    uint32_t two_sided;  /* rect, disc: emit from both faces */
} pfxh_light;

/* gpu::Camera (crates/trace/src/gpu.rs:172-178) */
typedef struct pfxh_camera {
    float origin[3];
    float forward[3];
    float right[3];
    float up[3];
} pfxh_camera;

/* gpu::Sun (crates/trace/src/gpu.rs:230-235) */
typedef struct pfxh_sun {
    float direction[3];
    float color[3];
    float intensity;
} pfxh_sun;

/* sky::Environment::Hdr (crates/trace/src/sky.rs:11-15): a lat-long
 * RGBA32F texture, row-major, `width * height * 4` floats. width == 0 is a
 * black sky (the sky slot is still sampled, as the reference always does). */
typedef struct pfxh_environment {
    uint32_t width;
    uint32_t height;
    const float* texels;
} pfxh_environment;

typedef struct pfxh_scene {
    const pfxh_triangle* triangles;
    uint32_t triangle_count;
    const pfxh_triangle_detail* world_details; /* Detail::triangles */
    uint32_t world_detail_count;
    const pfxh_triangle_detail* mesh_triangles;
    uint32_t mesh_triangle_count;
    const pfxh_mesh* meshes;
    uint32_t mesh_count;
    const pfxh_instance* instances;
    uint32_t instance_count;
    const pfxh_material* materials;
    uint32_t material_count;
    const pfxh_light* lights;
    uint32_t light_count;
    pfxh_environment sky;
    pfxh_camera camera;
    pfxh_sun sun;
} pfxh_scene;

/* detail::Bounces (crates/trace/src/detail.rs:414-453) */
typedef struct pfxh_bounces {
    uint32_t total;
    uint32_t diffuse;
    uint32_t glossy;
    uint32_t transmission;
} pfxh_bounces;

enum { PFXH_ACCUM_COMPENSATED = 0, PFXH_ACCUM_PLAIN = 1 };

/* gpu::Projection (crates/trace/src/gpu.rs:181-189) */
enum { PFXH_PROJ_PERSPECTIVE = 0, PFXH_PROJ_ORTHOGRAPHIC = 1, PFXH_PROJ_EQUIRECT = 2 };

/* Ray-triangle test. PFXH_TRI_MOLLER_TRUMBORE is the reference's tri_hit /
 * tri_blocks (crates/trace/src/shaders/trace.wgsl:780-804, 1076-1095) and
 * the default, so hit ids match the reference. PFXH_TRI_WATERTIGHT is the
 * watertight test of Woop, Benthin and Wald (JCGT 2(1), 2013) in FP32: a
 * ray crossing a mesh through a shared edge or vertex hits one of the
 * triangles that meet there (no cracks). Same acceptance otherwise:
 * t > 1e-4 and closer than the best hit, one-sided culling by the sign of
 * the determinant, barycentrics u (weight of b) and v (weight of c). */
enum { PFXH_TRI_MOLLER_TRUMBORE = 0, PFXH_TRI_WATERTIGHT = 1 };

/* Closest-hit walk order. PFXH_WALK_NEAREST (the default) visits the hit
 * children of a node nearest first; it returns the reference's hit except
 * where the reference's own answer depends on its walk order (a ray through
 * a shared edge or vertex whose hit lies within an ulp of a box face).
 * PFXH_WALK_LEFT_FIRST is the reference's order (trace.wgsl:999-1075, left
 * child popped first, :1049-1054): it tests the reference's triangles in
 * the reference's order, so the closest hit is the reference's on every
 * ray. Any-hit (shadow) walks are left first in both modes. */
enum { PFXH_WALK_NEAREST = 0, PFXH_WALK_LEFT_FIRST = 1 };

/* Sampler (crates/trace/src/sobol.rs:8-31). PFXH_SAMPLER_PCG is
 * `Sampler::Pcg`, one PCG stream per path (trace.wgsl:423-441), what the
 * configs and PFX_TRACE_SAMPLER=pcg use. PFXH_SAMPLER_SOBOL is
 * `Sampler::Sobol { samples }`, the reference's default: 256 scrambled 4-D
 * patterns of sequence_size(samples) points (a power of two from 256 to
 * 8192), looked up per pixel, sample and dimension (sobol.wgsl). */
enum { PFXH_SAMPLER_PCG = 0, PFXH_SAMPLER_SOBOL = 1 };

/* The parts of detail::Detail (crates/trace/src/detail.rs:374-398) the
 * path uses. */
typedef struct pfxh_detail {
    pfxh_bounces bounces;
    float lens_shift[2];
    float aperture;
    float focus_distance;
    float sun_radius_deg;
    uint32_t accumulate; /* PFXH_ACCUM_COMPENSATED (Neumaier) or PFXH_ACCUM_PLAIN */
    uint32_t wave_paths; /* paths in flight per wave; 0 picks a default */
    uint32_t projection; /* PFXH_PROJ_*; Detail::projection */
    float ortho_size[2]; /* orthographic view width and height (PFXH_PROJ_ORTHOGRAPHIC) */
    uint32_t no_emitter_sampling; /* 1: emissive triangles are not light-sampled, as
                                   * Trace::build with emitter_sampling = false (gpu.rs:781) */
    uint32_t triangle_test;       /* PFXH_TRI_*; 0 is the reference's test */
    uint32_t walk_order;          /* PFXH_WALK_*; 0 is nearest first */
    uint32_t sampler;             /* PFXH_SAMPLER_*; 0 is PCG */
    uint32_t sobol_samples;       /* Sampler::Sobol { samples }: sizes the patterns (256..8192 points) */
} pfxh_detail;

/* adaptive::Adaptive (crates/trace/src/adaptive.rs:9-26) */
typedef struct pfxh_adaptive {
    float threshold;
    uint32_t min_samples;
    uint32_t max_samples;
    float growth;
} pfxh_adaptive;

/* adaptive::Converged + Round (crates/trace/src/adaptive.rs:66-83) */
#define PFXH_MAX_ROUNDS 64
typedef struct pfxh_converged {
    uint32_t round_count;
    uint32_t samples[PFXH_MAX_ROUNDS];
    uint64_t active[PFXH_MAX_ROUNDS];
    uint64_t pixels;
} pfxh_converged;

/* Probe (crates/trace/src/probe.rs:10-18) plus the hit primitive: the
 * caller's triangle index (world triangles first, then world_details, then
 * mesh_triangles by mesh-local index), or UINT32_MAX on a miss. */
typedef struct pfxh_probe_hit {
    float depth;
    float distance;
    uint32_t material_id; /* material + 1, 0 on a miss */
    uint32_t primitive;
    uint32_t instance;    /* 0 world, k for the k-th placed instance */
    float normal[3];
} pfxh_probe_hit;

/* Counters the kernels keep, for throughput and roofline reports. */
typedef struct pfxh_stats {
    uint64_t closest_rays;  /* extension rays traced */
    uint64_t shadow_rays;   /* any-hit rays traced */
    uint64_t paths;         /* camera paths started */
    uint64_t nodes_visited; /* closest-hit 64-byte nodes fetched, when counting is on */
    uint64_t prims_tested;  /* closest-hit triangles tested, when counting is on */
    uint64_t shadow_nodes;  /* the same two counts for any-hit rays */
    uint64_t shadow_prims;
    double trace_ms;        /* GPU time of closest-hit kernels (HIP events) */
    double shadow_ms;       /* GPU time of any-hit kernels */
    double shade_ms;        /* GPU time of raygen, shading and accumulation kernels */
    uint32_t launches;      /* closest-hit launches behind trace_ms */
    uint64_t boxes_tested;  /* closest-hit child boxes tested (the reference's binary node
                             * visits), when counting is on */
    /* Closest-hit wave rounds, when counting is on: a wave of 64 lanes steps
     * either inner nodes or leaves/instances per round. */
    uint64_t rounds_inner;  /* inner-node rounds */
    uint64_t rounds_leaf;   /* leaf and instance rounds */
    uint64_t lanes_inner;   /* lanes that stepped in inner rounds */
    uint64_t lanes_leaf;    /* lanes that stepped in leaf rounds */
    uint64_t lanes_live;    /* lanes holding a ray, over all rounds */
    uint64_t leaf_trips;    /* triangle-loop trips of leaf rounds (the wave's longest leaf) */
} pfxh_stats;

typedef struct pfxh_trace pfxh_trace;

/* Device limits traces are built against, the counterpart of the wgpu
 * limits Trace::new checks before it allocates (fits_buffer,
 * crates/trace/src/gpu.rs:1258-1278): a trace whose allocation would pass
 * one is refused with PFXH_REFUSED and the message "limit <name> of
 * <needed> is beyond the device's <has>" (GpuShortfall::Limit,
 * crates/gpu/src/floor.rs:349-351), name max_buffer_size (one allocation)
 * or max_memory (everything the trace, or a batch, holds). 0 means the
 * device's own limit (hipMalloc refuses what HBM cannot hold, also as
 * PFXH_REFUSED). Process-wide; for callers that share a GPU or want a
 * smaller budget, and for tests. */
typedef struct pfxh_device_limits {
    uint64_t max_buffer_size;
    uint64_t max_memory;
} pfxh_device_limits;
/* NULL restores the device's own limits. */
void pfxh_set_device_limits(const pfxh_device_limits* limits);
void pfxh_get_device_limits(pfxh_device_limits* out);

/* Trace::new_detailed (crates/trace/src/gpu.rs:1926-1934) */
pfxh_status pfxh_trace_new(const pfxh_scene* scene, const pfxh_detail* detail, uint32_t width,
                           uint32_t height, pfxh_trace** out, char* err, size_t err_len);
/* Trace::sample (gpu.rs:2942-2949): `count` more samples with `seed`,
 * accumulated onto the samples already taken. */
pfxh_status pfxh_trace_sample(pfxh_trace* trace, uint32_t count, uint32_t seed, char* err,
                              size_t err_len);
/* Trace::sample_adaptive (crates/trace/src/adaptive.rs:112-324) */
pfxh_status pfxh_trace_sample_adaptive(pfxh_trace* trace, const pfxh_adaptive* adaptive,
                                       uint32_t seed, pfxh_converged* out, char* err,
                                       size_t err_len);
/* The standard error of each pixel's mean luminance ((r + g + b) / 3) over
 * its samples, sqrt(max(E[l^2] - E[l]^2, 0) / n), width*height floats,
 * caller-allocated; 0 where a pixel has no samples. No counterpart in the
 * reference: the learned denoiser's noise channel. PFXCKPT1 holds no
 * moments: after pfxh_trace_restore the restored samples count at their
 * pixel's mean with no spread, so it reads 0 until new samples are taken
 * and understates the spread while the restored samples dominate. */
pfxh_status pfxh_trace_readback_deviation(pfxh_trace* trace, float* out, char* err, size_t err_len);
/* Trace::readback (gpu.rs:3236-3246): three RGBA32F images, row-major,
 * width*height*4 floats each, caller-allocated. Any pointer may be NULL. */
pfxh_status pfxh_trace_readback(pfxh_trace* trace, float* color, float* albedo, float* normal,
                                char* err, size_t err_len);
/* Trace::set_camera (gpu.rs:2817-2834); resets the sample count */
pfxh_status pfxh_trace_set_camera(pfxh_trace* trace, const pfxh_camera* camera, char* err,
                                  size_t err_len);
/* Trace::set_window (gpu.rs:2836-2851); resets the sample count */
pfxh_status pfxh_trace_set_window(pfxh_trace* trace, uint32_t full_width, uint32_t full_height,
                                  uint32_t x, uint32_t y, char* err, size_t err_len);
/* Trace::set_sun (gpu.rs:2912-2918); resets the sample count */
pfxh_status pfxh_trace_set_sun(pfxh_trace* trace, const pfxh_sun* sun, char* err,
                               size_t err_len);
/* Trace::set_lens (gpu.rs:2858-2870); resets the sample count */
pfxh_status pfxh_trace_set_lens(pfxh_trace* trace, float shift_x, float shift_y, float aperture,
                                float focus_distance, char* err, size_t err_len);
/* Trace::set_projection (gpu.rs:2872-2890): `kind` is a PFXH_PROJ_* value,
 * width/height the orthographic view size (ignored otherwise); resets the
 * sample count */
pfxh_status pfxh_trace_set_projection(pfxh_trace* trace, uint32_t kind, float width, float height,
                                      char* err, size_t err_len);
/* Trace::set_sun_radius (gpu.rs:2903-2910): the sun's angular radius in
 * degrees, 0 <= radius < 90 (0 is a point sun); resets the sample count */
pfxh_status pfxh_trace_set_sun_radius(pfxh_trace* trace, float radius_deg, char* err, size_t err_len);
/* Trace::set_environment (gpu.rs:2810-2815): a new lat-long sky (width 0 is
 * black) and its sampling CDF; resets the sample count. A trace built
 * under a black sky grows its shadow-ray buffers by one light slot when
 * given a lit one; a batch still cannot (PFXH_INVALID). */
pfxh_status pfxh_trace_set_environment(pfxh_trace* trace, const pfxh_environment* sky, char* err,
                                       size_t err_len);
/* Trace::samples (gpu.rs:3232-3234) */
uint32_t pfxh_trace_samples(const pfxh_trace* trace);
/* Trace::gpu_bytes (gpu.rs:2360) */
uint64_t pfxh_trace_gpu_bytes(const pfxh_trace* trace);
/* Bytes one node visit of the trace's closest-hit walk fetches: 64 when its
 * default walks read quantized nodes (one byte per child box plane on a
 * power-of-two grid, boxes rounded outward), 112 when they read 4-wide FP32
 * nodes (six box planes and four child references of a 128-byte node: the
 * left-first walk, or a scene whose boxes do not quantize); 0 for a null
 * trace. No counterpart in the reference: a measurement aid. */
uint32_t pfxh_trace_node_bytes(const pfxh_trace* trace);
/* Closest hits of `count` caller rays (origins and directions each hold
 * 3*count floats: x, y, z per ray, in world space) against the trace's scene, with the trace's triangle test:
 * primitive[k] is the caller's triangle id as in pfxh_probe_hit (UINT32_MAX
 * on a miss) and distance[k] the hit distance along the direction (0 on a
 * miss). Hits closer than 1e-4 do not count; limit bounds the distance.
 * No counterpart in the reference (its only ray query is the camera probe);
 * for callers and tests that aim rays themselves. */
pfxh_status pfxh_trace_cast(pfxh_trace* trace, const float* origins, const float* directions,
                            uint32_t count, float limit, uint32_t* primitive, float* distance,
                            char* err, size_t err_len);

/* Trace::probe (crates/trace/src/probe.rs:42; probe.wgsl): pixel-centre
 * primary rays. `out` holds width*height records. */
pfxh_status pfxh_trace_probe(pfxh_trace* trace, pfxh_probe_hit* out, char* err, size_t err_len);
/* Kernel counters since the trace was built or the counters were reset. */
pfxh_status pfxh_trace_stats(const pfxh_trace* trace, pfxh_stats* out);
void pfxh_trace_reset_stats(pfxh_trace* trace);
/* Count nodes visited and triangles tested per ray (slower; off by default). */
void pfxh_trace_set_counting(pfxh_trace* trace, int on);
/* Drop the trace and its device memory. NULL is allowed. */
void pfxh_trace_free(pfxh_trace* trace);

/* ---- staging (crates/trace/src/stage.rs) ---------------------------------
 * Stage::build (stage.rs:210-385): indexed meshes and their placements in,
 * a scene and detail out, for callers that hold meshes the way a loaded
 * scene does. Content images, text, motion and clip planes are not staged
 * (those layers are not built here). */

/* stage::Mesh (stage.rs:27-36) without uv sets: 3 floats per vertex of
 * positions and normals, an optional alpha per vertex (NULL: opaque), and
 * 3 indices per triangle. */
typedef struct pfxh_stage_mesh {
    const float* positions;
    const float* normals;
    const float* alpha;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t index_count;
} pfxh_stage_mesh;

/* stage::Placement (stage.rs:150-195). Placement::new sets casts_shadow 1,
 * shadow_only 0, two_sided 1 and alpha_cutoff 0.5 (ALPHA_CUTOFF). */
typedef struct pfxh_placement {
    uint32_t mesh;
    uint32_t material;
    float model[16]; /* column-major 4x4, affine */
    uint32_t casts_shadow;
    uint32_t shadow_only;
    uint32_t two_sided;
    float alpha_cutoff; /* corners below it are cut away (stage.rs:770-822) */
} pfxh_placement;

/* stage::Stage (stage.rs:196-203), plus the lights the trace's detail
 * takes. projection is a PFXH_PROJ_* value, ortho_size its view size; the
 * lens is detail::Lens (aperture, focus distance > 0). */
typedef struct pfxh_stage {
    const pfxh_stage_mesh* meshes;
    uint32_t mesh_count;
    const pfxh_placement* placements;
    uint32_t placement_count;
    const pfxh_material* materials;
    uint32_t material_count;
    const pfxh_light* lights;
    uint32_t light_count;
    pfxh_environment sky;
    pfxh_sun sun;
    pfxh_camera camera;
    uint32_t projection;
    float ortho_size[2];
    float lens_shift[2];
    float aperture;
    float focus_distance;
} pfxh_stage;

typedef struct pfxh_staged pfxh_staged;

/* Stage::build: PFXH_SETUP with the reference's message for a stage with
 * no material, an invalid projection or lens, a mesh without vertices or
 * indices or with an index outside its vertices, a placement naming a
 * missing mesh or material, or an alpha cutoff outside [0, 1]. Placements
 * of one mesh, cutoff and material share one mesh of the staged scene. */
pfxh_status pfxh_stage_build(const pfxh_stage* stage, pfxh_staged** out, char* err, size_t err_len);
/* Staged::scene: valid until pfxh_staged_free; no world triangles. */
const pfxh_scene* pfxh_staged_scene(const pfxh_staged* staged);
/* Staged::detail: Detail::default() (Bounces::REFERENCE 16/4/8/12) with the
 * stage's lens and projection. */
pfxh_status pfxh_staged_detail(const pfxh_staged* staged, pfxh_detail* out);
/* Staged::trace (stage.rs:383-387): pfxh_trace_new on the staged scene
 * with `detail` (NULL: the staged detail). */
pfxh_status pfxh_staged_trace(const pfxh_staged* staged, const pfxh_detail* detail, uint32_t width,
                              uint32_t height, pfxh_trace** out, char* err, size_t err_len);
void pfxh_staged_free(pfxh_staged* staged);

/* ---- host-only scene layout (no device needed) -------------------------- */

/* bvh::Node (crates/trace/src/bvh.rs:5-13), 48 bytes */
typedef struct pfxh_bvh_node {
    float lo[4];
    float hi[4];
    uint32_t first;
    uint32_t count;
    uint32_t pad[2];
} pfxh_bvh_node;

/* Bvh::try_build (crates/trace/src/bvh.rs:95-133): binned SAH build.
 * nodes needs room for 2*count nodes, order for count entries. */
pfxh_status pfxh_layout_bvh(const pfxh_triangle* triangles, uint32_t count, pfxh_bvh_node* nodes,
                            uint32_t* node_count, uint32_t* order, char* err, size_t err_len);

/* levels::lay + SceneBuffers node buffer (crates/trace/src/levels.rs:225-383,
 * 421-470): the reference's two-level node buffer and world order for a
 * whole scene. Call once with nodes == NULL to learn the counts. */
pfxh_status pfxh_layout_levels(const pfxh_scene* scene, pfxh_bvh_node* nodes,
                               uint32_t* node_count, uint32_t* order, uint32_t* order_count,
                               char* err, size_t err_len);

/* The traversal nodes of a scene's device layout, for tests: `count`
 * 4-wide FP32 nodes of 128 bytes (lo x, y, z[4], hi x, y, z[4], ref[4],
 * pad[4]: the reference's binary tree collapsed two levels at a time) and,
 * when the scene's boxes quantize (*quantized = 1), as many quantized nodes
 * of 64 bytes (origin[3], scale[3], q_lo[3], q_hi[3], ref[4]; child c's
 * plane is byte c of its word, value fmaf(q, scale, origin), a box that
 * contains the exact one). Call with null arrays to learn the count. No
 * counterpart in the reference (its nodes are binary, pfxh_layout_levels). */
pfxh_status pfxh_layout_nodes(const pfxh_scene* scene, void* nodes, void* qnodes, uint32_t* count,
                              uint32_t* quantized, char* err, size_t err_len);

/* tree::lay (crates/trace/src/tree.rs:220-243): light rows (16 floats
 * each, `GpuLight`) for [non-treed][treed in tree order][tree nodes].
 * rows needs room for 3*count rows (lights plus fewer than 2*count nodes). */
pfxh_status pfxh_layout_lights(const pfxh_light* lights, uint32_t count, float* rows,
                               uint32_t* row_count, uint32_t* looped, uint32_t* tree, char* err,
                               size_t err_len);

/* The split step of the light-tree build, emitters.select_nth_unstable_by
 * on the centroid (crates/trace/src/tree.rs:281-283): Rust core's
 * slice::select_nth_unstable_by (Rust 1.97) over `len` (key, id) pairs, keys
 * compared with f32::total_cmp, permuted in place. mode 0 selects `index`
 * (result = index); mode 1 runs one partition around keys[index] (result =
 * the count of smaller keys, now the pivot's position); mode 2 runs the
 * median-of-medians fallback alone. index < len. */
pfxh_status pfxh_layout_select(float* keys, uint32_t* ids, size_t len, size_t index, int mode,
                               size_t* result);

/* sobol::table (crates/trace/src/sobol.rs:137-150) for Sampler::Sobol
 * { samples }: *length = sequence_size(samples) points per pattern
 * (sobol.rs:33-39); with table non-NULL, the 256 patterns, 4 floats a
 * point, pattern-major (256 * length * 4 floats). */
pfxh_status pfxh_layout_sobol(uint32_t samples, float* table, uint32_t* length);
/* sobol::animated_seed (crates/trace/src/sobol.rs:105-107), what
 * still::Seed::Animated(frame).of(seed) (crates/still/src/trace.rs:77-82)
 * passes to Trace::sample for one frame of an animation: Cycles'
 * hash_uint2 (lookup3) of (frame, seed). */
uint32_t pfxh_sobol_animated_seed(uint32_t seed, uint32_t frame);

/* lights::build_emitters (crates/trace/src/lights.rs:391-440) for the
 * emissive triangles of a scene: entries (16 floats each, `GpuEmitter`,
 * header first) and the CDF. Call with entries == NULL to learn the count. */
pfxh_status pfxh_layout_emitters(const pfxh_scene* scene, float* entries, float* cdf,
                                 uint32_t* emitter_count, char* err, size_t err_len);

/* ---- checkpoints (crates/trace/src/checkpoint.rs) ------------------------ */

#define PFXH_NO_ROUNDS 0xffffffffu

/* checkpoint::Saved (checkpoint.rs:13-28): sizes and pointers to the blocks.
 * accum is 48 bytes per pixel (Accum: color.xyz + sample count, albedo,
 * normal sums), color/albedo/normal 16 bytes per pixel (RGBA32F), snaps 32
 * bytes per pixel and present only with adaptive rounds. */
typedef struct pfxh_checkpoint {
    uint32_t width;
    uint32_t height;
    uint32_t samples;
    uint32_t round_count; /* PFXH_NO_ROUNDS when the checkpoint has no adaptive rounds */
    const uint8_t* accum;
    const uint8_t* color;
    const uint8_t* albedo;
    const uint8_t* normal;
    const uint32_t* round_samples;
    const uint64_t* round_active;
    const uint8_t* snaps;
} pfxh_checkpoint;

/* Saved::write_to (checkpoint.rs:85-110): the PFXCKPT1 bytes under `key`.
 * With out == NULL only *len is set. Host only. */
pfxh_status pfxh_checkpoint_write(const pfxh_checkpoint* checkpoint, const char* key, uint8_t* out,
                                  size_t cap, size_t* len, char* err, size_t err_len);
/* Saved::read_from (checkpoint.rs:112-158): parse `bytes` made under `key`;
 * the block pointers of *out point into `bytes`, and up to round_cap rounds
 * are copied to round_samples/round_active. Host only. */
pfxh_status pfxh_checkpoint_read(const uint8_t* bytes, size_t len, const char* key,
                                 pfxh_checkpoint* out, uint32_t* round_samples,
                                 uint64_t* round_active, uint32_t round_cap, char* err,
                                 size_t err_len);
/* Trace::save (checkpoint.rs:163-174) + Saved::write_to: the trace's
 * accumulator, images and sample count as PFXCKPT1 bytes (no adaptive
 * rounds). A compensated or FP64 accumulator is saved as its FP32 sum.
 * With out == NULL only *len is set. */
pfxh_status pfxh_trace_save(pfxh_trace* trace, const char* key, uint8_t* out, size_t cap,
                            size_t* len, char* err, size_t err_len);
/* Saved::read_from + Trace::restore (checkpoint.rs:176-200): continue from
 * the checkpoint's accumulator and sample count; the images are developed
 * again from the accumulator at readback. Adaptive rounds in the
 * checkpoint are read and ignored, as Trace::restore does. */
pfxh_status pfxh_trace_restore(pfxh_trace* trace, const char* key, const uint8_t* bytes, size_t len,
                               char* err, size_t err_len);

/* ---- the tracer's own denoiser (crates/trace/src/denoise.rs) -------------- */

/* denoise::filter (denoise.rs:1-62): three edge-aware a-trous passes on
 * RGBA32F colour guided by albedo and normal (width*height*4 floats each),
 * for an image of `samples` spp; `out` gets RGBA32F (alpha from the colour).
 * PFXH_INVALID ("invalid denoise inputs") for zero samples. */
pfxh_status pfxh_atrous(uint32_t width, uint32_t height, const float* color, const float* albedo,
                        const float* normal, uint32_t samples, float* out, char* err, size_t err_len);
/* The same on a trace's current images and sample count, on the device. */
pfxh_status pfxh_trace_atrous(pfxh_trace* trace, float* out, char* err, size_t err_len);
/* pfxh_atrous on device memory, allocating nothing. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * color, albedo, normal and out are device pointers to width*height RGBA32F
 * pixels; out must not alias an input. The passes are queued on `stream`
 * (a hipStream_t; NULL is the null stream) and the call returns without
 * waiting: out is ready once the stream has run them. The middle pass's
 * image lives in one static device array per GPU of 2^21 pixels (32 MiB,
 * held until the process ends): larger images are refused with
 * PFXH_INVALID. Every call on a GPU shares that array, so calls that overlap
 * on one GPU (other streams, other threads) race; serialise them. The
 * result equals pfxh_atrous's bit for bit. */
pfxh_status pfxh_atrous_noalloc(uint32_t width, uint32_t height, const float* color, const float* albedo,
                                const float* normal, uint32_t samples, float* out, void* stream, char* err,
                                size_t err_len);

/* ---- the repo denoiser (crates/denoise: OIDN 2.5.1's RT filter) ------------ */

typedef struct pfxh_denoiser pfxh_denoiser;

/* tile::Tile (crates/denoise/src/tile.rs:4-12), [y, x] pairs */
typedef struct pfxh_denoise_tile {
    uint32_t src[2];
    uint32_t dst[2];
    uint32_t size[2];
    uint32_t out_src[2];
    uint32_t out_dst[2];
    uint32_t out_size[2];
} pfxh_denoise_tile;

/* Denoiser::new (crates/denoise/src/lib.rs:82-112): parse OIDN's RT weights
 * (the TZA blob pfx_amd/weights/rt_hdr_alb_nrm.tza, tza.rs:50-111), build
 * the U-Net (net.rs:202-232) and upload it. */
pfxh_status pfxh_denoiser_new(const uint8_t* weights, size_t len, pfxh_denoiser** out, char* err,
                              size_t err_len);
/* Denoiser::denoise (lib.rs:121-144): RGBA32F colour, albedo and normal
 * (row-major, width*height*4 floats each) in, the denoised RGBA32F image
 * out (alpha copied from the colour). */
pfxh_status pfxh_denoise(pfxh_denoiser* denoiser, uint32_t width, uint32_t height, const float* color,
                         const float* albedo, const float* normal, float* out, char* err,
                         size_t err_len);
/* Denoiser::denoise_textures on a trace's current colour, albedo and normal
 * (still::developed, crates/still/src/trace.rs:1299-1339), without a host
 * round trip for the inputs. `out` holds width*height*4 floats. */
pfxh_status pfxh_trace_denoise(pfxh_trace* trace, pfxh_denoiser* denoiser, float* out, char* err,
                               size_t err_len);
/* Largest tile in pixels (0, the default: one tile per image; the
 * reference uses 2160*2160, tile.rs:2). Results do not depend on it: tiles
 * overlap by more than half the receptive field and align to 16. */
void pfxh_denoiser_set_max_tile(pfxh_denoiser* denoiser, uint64_t pixels);
/* Layer and parameter counts of a weights blob (host only). */
pfxh_status pfxh_denoiser_layers(const uint8_t* weights, size_t len, uint32_t* layers,
                                 uint64_t* params, char* err, size_t err_len);
/* tile::plan (tile.rs:31-101) for a height x width image: the tile size and
 * up to `cap` tiles; *count gets the tile count (host only). */
pfxh_status pfxh_denoise_plan(uint32_t height, uint32_t width, uint64_t max_tile_pixels,
                              uint32_t* tile_h, uint32_t* tile_w, pfxh_denoise_tile* tiles,
                              uint32_t cap, uint32_t* count);
/* Drop the denoiser and its device memory. NULL is allowed. */
void pfxh_denoiser_free(pfxh_denoiser* denoiser);

/* ---- batches of stills -------------------------------------------------------
 * No counterpart in the reference, which renders one still per Trace; the
 * nearest is the bake's probe batch, many camera faces of one scene in one
 * dispatch (crates/bake/src/batch.rs:10-14,79-100). A batch renders many
 * stills, each with its own scene, size, sample count and seed, with one
 * launch per stage for all of them; the bounce loop of a wave is captured
 * in a HIP graph. Every still's result equals that of its own pfxh_trace at
 * the same seed, byte for byte. Calls on one batch, and on the traces
 * pfxh_batch_trace lends, must be serialised by the caller. */

typedef struct pfxh_batch pfxh_batch;

typedef struct pfxh_batch_still {
    const pfxh_scene* scene; /* read during pfxh_batch_new only */
    uint32_t width, height;
    uint32_t spp;  /* samples pfxh_batch_render adds */
    uint32_t seed; /* the seed pfxh_trace_sample would be given */
} pfxh_batch_still;

/* Launch shape; pfxh_batch_tune picks it by timing trial renders. */
typedef struct pfxh_batch_tuning {
    uint32_t wave_paths;          /* paths per wave (>= the largest still's pixels) */
    uint32_t refill;              /* idle lanes that make a traversal wave fetch rays, 1..64 */
    uint32_t trace_blocks_per_cu; /* persistent traversal blocks per CU, 1..16 */
    uint32_t shade_blocks_per_cu; /* shading blocks per CU, 1..64 */
    uint32_t graph;               /* 1: replay the captured bounce loop, 0: launch directly */
} pfxh_batch_tuning;

/* Build every still's scene (one `detail` for all) and the shared wave. */
pfxh_status pfxh_batch_new(const pfxh_batch_still* stills, uint32_t count, const pfxh_detail* detail,
                           pfxh_batch** out, char* err, size_t err_len);
/* Add each still's `spp` samples, as pfxh_trace_sample(trace, spp, seed)
 * would on each still's own trace. */
pfxh_status pfxh_batch_render(pfxh_batch* batch, char* err, size_t err_len);
/* The same with counts[i] samples for still i. */
pfxh_status pfxh_batch_render_samples(pfxh_batch* batch, const uint32_t* counts, char* err, size_t err_len);
/* Autotune: time trial renders (an eighth of every still's samples) over
 * wave sizes, refill thresholds and traversal blocks per CU, keep the
 * fastest; the stills are left with zero samples. `out` may be NULL. */
pfxh_status pfxh_batch_tune(pfxh_batch* batch, pfxh_batch_tuning* out, char* err, size_t err_len);
void pfxh_batch_get_tuning(const pfxh_batch* batch, pfxh_batch_tuning* out);
pfxh_status pfxh_batch_set_tuning(pfxh_batch* batch, const pfxh_batch_tuning* tuning, char* err,
                                  size_t err_len);
uint32_t pfxh_batch_count(const pfxh_batch* batch);
/* Still `index` as a trace (readback, denoise, probe, sample), owned by the
 * batch: do not free it. NULL when out of range. */
pfxh_trace* pfxh_batch_trace(pfxh_batch* batch, uint32_t index);
/* Rays and paths traced by pfxh_batch_render since the last reset. */
pfxh_status pfxh_batch_stats(const pfxh_batch* batch, pfxh_stats* out);
void pfxh_batch_reset_stats(pfxh_batch* batch);
void pfxh_batch_free(pfxh_batch* batch);

/* ---- the learned denoiser --------------------------------------------------
 * No counterpart in the reference, which runs only the fixed-weight filter
 * above (crates/denoise/src/lib.rs:15-17). A 5-level encoder-decoder of 16
 * 3x3 convolutions with skip connections (the U-Net shape of net.rs:202-232,
 * 32->128 channels, 1,239,651 parameters), BF16 operands on matrix cores
 * with FP32 sums, trained on the GPU (forward and backward kernels, L1 loss
 * on log radiance, Adam) from pairs the tracer renders. Inputs are the
 * still's colour, albedo, normal, depth and the standard error of each
 * pixel's mean luminance (11 channels); the output is
 * colour. Parameters cross the boundary in a canonical order: per layer,
 * weights [out][in][3][3] (the inputs being the first source's channels,
 * then the skip source's) and then the bias [out]. */

#define PFXH_LEARN_LAYERS 16
#define PFXH_LEARN_FP32 1u /* pfxh_learner_new: FP32 operands instead of BF16 */

typedef struct pfxh_learner pfxh_learner;

typedef struct pfxh_learn_layer_info {
    char name[16];
    int32_t in0;  /* tensor read (0: the encoded input, k + 1: layer k's output) */
    int32_t in1;  /* skip tensor concatenated after in0, or -1 */
    uint32_t in_channels, out_channels;
    uint32_t level;    /* resolution: full >> level */
    uint32_t upsample; /* in0 comes from level + 1, nearest 2x */
    uint32_t pool;     /* 2x2 max pool after the activation */
    uint32_t relu;
} pfxh_learn_layer_info;

/* The network's layers (host only). */
pfxh_status pfxh_learn_layer(uint32_t index, pfxh_learn_layer_info* out);
uint64_t pfxh_learn_param_count(void);
/* A network with zero parameters; `flags` PFXH_LEARN_FP32 or 0. */
pfxh_status pfxh_learner_new(uint32_t flags, pfxh_learner** out, char* err, size_t err_len);
pfxh_status pfxh_learner_set_params(pfxh_learner* learner, const float* params, uint64_t count,
                                    char* err, size_t err_len);
pfxh_status pfxh_learner_params(pfxh_learner* learner, float* out, uint64_t count, char* err,
                                size_t err_len);
/* Gradient of the loss at the last pfxh_learner_gradient or training step. */
pfxh_status pfxh_learner_grads(pfxh_learner* learner, float* out, uint64_t count, char* err,
                               size_t err_len);
/* The training loss: PFXH_LEARN_LOSS_L1_LOG (the default), the mean over
 * pixels and colour channels of |y - log(1 + target)| with y the network's
 * log radiance; or PFXH_LEARN_LOSS_REL_L2, the mean of
 * (x - target)^2 / (x^2 + 0.01) with x = expm1(y) and the denominator held
 * fixed in the gradient (the relative L2 of Noise2Noise, Lehtinen et al.
 * 2018: its minimiser over independent noisy targets is their mean, so
 * targets may be noisy renders), y capped at 40. */
#define PFXH_LEARN_LOSS_L1_LOG 0u
#define PFXH_LEARN_LOSS_REL_L2 1u
pfxh_status pfxh_learner_set_loss(pfxh_learner* learner, uint32_t mode);
/* Inference passes (self-ensemble over mirror images): 1 (the default), 2
 * (also the image mirrored in x) or 4 (also mirrored in y, and in both);
 * the frames run through the network as one batch and the denoised image is
 * the mean of their outputs, each mirrored back, in linear radiance. The
 * network was trained under these symmetries, so the mean keeps its bias
 * and averages out part of its error; the cost is the passes' convolutions
 * and their activations (the inference frame holds `passes` frames: about
 * 7 GB per pass at 3840x2160). Applies to pfxh_learned_denoise and the
 * trace entry points; other values
 * give PFXH_INVALID. No counterpart in the reference (it has no trained
 * denoiser). */
pfxh_status pfxh_learner_set_passes(pfxh_learner* learner, uint32_t passes);
/* Clear Adam's moments and step count. */
void pfxh_learner_reset_optimizer(pfxh_learner* learner);
/* Make room on the GPU for `capacity` training crops of side x side pixels
 * (a multiple of 16), dropping crops added before. */
pfxh_status pfxh_learner_reserve(pfxh_learner* learner, uint32_t side, uint32_t capacity, char* err,
                                 size_t err_len);
/* Add one crop: RGBA32F colour, albedo, normal and target (side*side*4
 * floats each), depth and deviation (side*side floats each, or NULL; the
 * deviation is the standard error of each pixel's mean luminance, as
 * pfxh_trace_readback_deviation gives it). */
pfxh_status pfxh_learner_add(pfxh_learner* learner, const float* color, const float* albedo,
                             const float* normal, const float* depth, const float* deviation,
                             const float* target, char* err, size_t err_len);
uint32_t pfxh_learner_crops(const pfxh_learner* learner);
/* Loss and gradient on `batch` crops, no update: picks[b] = crop << 3 |
 * symmetry (bit 0 mirrors x, bit 1 mirrors y, bit 2 transposes). */
pfxh_status pfxh_learner_gradient(pfxh_learner* learner, const uint32_t* picks, uint32_t batch,
                                  float* loss, char* err, size_t err_len);
/* `steps` Adam steps (beta 0.9, 0.999, eps 1e-8) on batches drawn from the
 * training set with `seed`; losses[k] gets step k's mean loss. */
pfxh_status pfxh_learner_train(pfxh_learner* learner, uint32_t steps, uint32_t batch, float lr,
                               uint64_t seed, float* losses, char* err, size_t err_len);
/* Denoise host images (RGBA32F colour, albedo, normal; depth and
 * deviation NULL or width*height floats); `out` gets RGBA32F, alpha from
 * the colour. */
pfxh_status pfxh_learned_denoise(pfxh_learner* learner, uint32_t width, uint32_t height,
                                 const float* color, const float* albedo, const float* normal,
                                 const float* depth, const float* deviation, float* out, char* err,
                                 size_t err_len);
/* The same on a trace's current colour, albedo and normal, its probe's
 * depth and its deviation, without a host round trip for the inputs. */
pfxh_status pfxh_trace_learned_denoise(pfxh_trace* trace, pfxh_learner* learner, float* out, char* err,
                                       size_t err_len);
/* The same with the repo denoiser as fallback: per `tile` x `tile` tile the
 * variance the network removed, mean (log(1 + noisy) - log(1 + learned))^2
 * over its pixels and colour channels, is checked against `threshold`;
 * tiles above it (or not finite) take the repo denoiser's output, which is
 * computed only when some tile needs it. *fallback_tiles gets their count. */
pfxh_status pfxh_trace_learned_denoise_checked(pfxh_trace* trace, pfxh_learner* learner,
                                               pfxh_denoiser* denoiser, float threshold, uint32_t tile,
                                               float* out, uint32_t* fallback_tiles, char* err,
                                               size_t err_len);
void pfxh_learner_free(pfxh_learner* learner);

/* Achievable HBM bandwidth: the best of `reps` timed device-to-device
 * streaming copies of `bytes` (a multiple of 16, at least 1 MiB; nontemporal
 * float4 loads and stores, 2, 4 and 8 blocks of 256 threads per CU, one
 * warm-up copy each); *gbs = (bytes read + bytes written) / time in GB/s. The
 * denominator the traversal's algorithmic rate is set against, beside the
 * published peak. No counterpart in the reference (it publishes no HBM
 * figures). */
pfxh_status pfxh_hbm_probe(uint64_t bytes, uint32_t reps, double* gbs, char* err, size_t err_len);

/* Library version string. */
const char* pfxh_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PFXHIP_H */
