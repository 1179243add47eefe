// libpfxhip: the C ABI of include/pfxhip.h over the wavefront kernels.
//
// A pfxh_trace mirrors `pfx_trace::Trace` (crates/trace/src/gpu.rs:309-342):
// it owns the scene buffers, the accumulation buffer and one HIP stream;
// `sample` accumulates onto the samples already taken and every setter
// resets the count (gpu.rs:2810-2940). The reference's ~10 ms pacing of
// submissions (gpu.rs:2992-3107) has no counterpart: the compute queue is
// not shared with a display, so a call runs its waves back to back.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <vector>

#include "trace_internal.h"

pfxh_status pfxh_denoise_device(pfxh_denoiser* d, uint32_t width, uint32_t height, const void* color,
                                const void* albedo, const void* normal, float* out, hipStream_t after, char* err,
                                size_t err_len, const void** device_out = nullptr);
pfxh_status pfxh_learned_denoise_device(pfxh_learner* l, uint32_t width, uint32_t height, const void* color,
                                        const void* albedo, const void* normal, const float* depth,
                                        uint32_t depth_stride, const float* deviation, float* out,
                                        hipStream_t after, char* err, size_t err_len, void** device_out = nullptr);
#include "learn.h"
hipError_t pfxh_atrous_device(hipStream_t s, float4* a, float4* b, const float4* albedo, const float4* normal,
                              uint32_t w, uint32_t h, uint32_t samples);
#include "sobol.h"
#include "vec3.h"

// Idle lanes out of 64 at which a traversal wave fetches new rays; a tuning
// knob (PFXH_REFILL), clamped to 1..64. 24 by default: on the 4K terrain it
// beat 16, 18, 20, 22, 26, 28 and 32 (658 ms per adaptive frame against
// 666 at 16 and 662 at 32).
uint32_t refill_lanes() {
    const char* v = std::getenv("PFXH_REFILL");
    long n = v ? std::strtol(v, nullptr, 10) : 24;
    return (uint32_t)std::min(64L, std::max(1L, n));
}

uint32_t spill_threads_for() {
    int device = 0, cus = 256;
    if (hipGetDevice(&device) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    // 16 traversal blocks per CU keep every SIMD fed; the overflow stack is
    // sized for exactly that many threads.
    return (uint32_t)cus * 16u * (uint32_t)TRACE_BLOCK;
}

namespace {

// Projection::valid and Projection::fits (gpu.rs:191-208)
bool projection_valid(uint32_t kind, float w, float h) {
    if (kind == PFXH_PROJ_ORTHOGRAPHIC) return std::isfinite(w) && std::isfinite(h) && w > 0.0f && h > 0.0f;
    return kind == PFXH_PROJ_PERSPECTIVE || kind == PFXH_PROJ_EQUIRECT;
}
bool projection_fits(uint32_t kind, const float* forward, const float* right, const float* up) {
    if (kind == PFXH_PROJ_PERSPECTIVE) return true;
    for (const float* a : {forward, right, up})
        if (!(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] > 1e-12f)) return false;
    return true;
}
// Projection::place (gpu.rs:210-217)
void place_projection(SceneDev& S, uint32_t kind, float w, float h) {
    S.projection = kind;
    S.half_view[0] = kind == PFXH_PROJ_ORTHOGRAPHIC ? w * 0.5f : 0.0f;
    S.half_view[1] = kind == PFXH_PROJ_ORTHOGRAPHIC ? h * 0.5f : 0.0f;
}

}  // namespace

void trace_setup(pfxh_trace* t, const pfxh_scene& scene, const pfxh_detail& detail, uint32_t width, uint32_t height,
                 bool with_wave) {
    std::string e = build_scene(scene, t->host, detail.no_emitter_sampling == 0u,
                                detail.triangle_test == PFXH_TRI_WATERTIGHT);
    if (!e.empty()) throw DeviceError{PFXH_SETUP, e};
    const HostScene& h = t->host;
    HIP_TRY(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    for (auto& ev : t->ev) HIP_TRY(hipEventCreate(&ev));
    HIP_TRY(hipHostMalloc((void**)&t->host_counters, C_COUNT * sizeof(uint32_t), hipHostMallocDefault));
    SceneDev& S = t->S;
    S.nodes = t->upload(h.nodes, "bvh nodes");
    S.qnodes = h.qnodes.empty() ? nullptr : t->upload(h.qnodes, "quantized bvh nodes");
    S.tris = t->upload(h.tris, "triangles");
    S.attrs = t->upload(h.attrs, "triangle attributes");
    S.groups = t->upload(h.groups, "instances");
    S.surfaces = t->upload(h.surfaces, "instance surfaces");
    S.materials = t->upload(h.materials, "materials");
    S.lights = (const float4*)t->upload(h.lights, "lights");
    S.light_cones = (const float4*)t->upload(h.light_cones, "light cones");
    S.emitters = (const float4*)t->upload(h.emitters, "emitters");
    S.cdf = t->upload(h.emitter_cdf, "emitter cdf");
    S.sky_rows = t->upload(h.sky_rows, "sky rows");
    S.sky_columns = t->upload(h.sky_columns, "sky columns");
    S.env = t->upload(h.env, "sky texels");
    S.slot_prim = t->upload(h.slot_prim, "primitive map");
    S.record_instance = t->upload(h.record_instance, "instance map");
    S.root = h.root;
    S.geo_flags = (h.bounds_tame ? GEO_TAME : 0u) | (h.watertight ? GEO_WATERTIGHT : 0u) |
                  (detail.walk_order == PFXH_WALK_LEFT_FIRST ? GEO_LEFT_FIRST : 0u);
    S.env_w = h.env_w;
    S.env_h = h.env_h;
    S.env_black = h.env_black ? 1u : 0u;
    S.lit = h.lit ? 1u : 0u;
    S.looped = h.looped;
    if (detail.sampler == PFXH_SAMPLER_SOBOL) {
        uint32_t len = pfx::sobol_length(detail.sobol_samples);
        std::vector<float> table((size_t)len * pfx::SOBOL_PATTERNS * 4);
        pfx::sobol_table(len, table.data());
        S.sobol_table = t->upload(table, "sobol patterns");
        S.sobol_len = len;
    }
    S.area = 0u;
    for (uint32_t i = 0; i < h.looped; ++i) {
        uint32_t kind = (uint32_t)h.lights[16 * i + 3];
        if ((kind == 2u || kind == 3u) && h.lights[16 * i + 15] >= 0.5f) S.area = 1u;
    }
    S.tree = h.tree;
    S.emitter_count = h.emitter_count;
    S.light_slots = h.lit ? 1u + h.looped + (h.tree > 0 ? 1u : 0u) : 0u;
    S.entries = (scene.sun.intensity > 0.0f ? 1u : 0u) + (h.env_black ? 0u : 1u) +
                (h.lit ? ((h.emitter_count > 0 ? 1u : 0u) + h.looped + (h.tree > 0 ? 1u : 0u)) : 0u);
    if (S.entries > 15u) throw DeviceError{PFXH_SETUP, "too many looped lights for one path vertex (at most 13)"};
    pfxh_bounces b = detail.bounces;
    if (b.total > 64 || b.diffuse > 64 || b.glossy > 64 || b.transmission > 64)
        throw DeviceError{PFXH_SETUP, "invalid trace bounces"};
    S.total = b.total;
    S.diffuse = b.diffuse;
    S.glossy = b.glossy;
    S.transmission = b.transmission;
    for (int k = 0; k < 3; ++k) {
        S.sun_dir[k] = scene.sun.direction[k];
        S.sun_color[k] = scene.sun.color[k];
        S.origin[k] = scene.camera.origin[k];
        S.forward[k] = scene.camera.forward[k];
        S.right[k] = scene.camera.right[k];
        S.up[k] = scene.camera.up[k];
    }
    S.sun_intensity = scene.sun.intensity;
    S.sun_w = pfx::fm_cos(detail.sun_radius_deg * (3.14159265358979f / 180.0f));
    S.lens[0] = detail.lens_shift[0];
    S.lens[1] = detail.lens_shift[1];
    S.lens[2] = detail.aperture * 0.5f;
    S.lens[3] = detail.focus_distance;
    place_projection(S, detail.projection, detail.ortho_size[0], detail.ortho_size[1]);
    S.window[0] = width;
    S.window[1] = height;
    S.window[2] = 0;
    S.window[3] = 0;
    S.width = width;
    S.height = height;
    t->width = width;
    t->height = height;
    size_t pixels = (size_t)width * height;
    AccumDev& A = t->A;
    A.mode = detail.accumulate;
    if (A.mode > 2u) throw DeviceError{PFXH_SETUP, "unknown accumulation mode"};
    A.sum = (float4*)t->alloc(pixels * 48, "accumulation");
    A.comp = A.mode == PFXH_ACCUM_COMPENSATED ? (float4*)t->alloc(pixels * 48, "accumulation compensation") : nullptr;
    A.sum64 = A.mode == 2u ? (double*)t->alloc(pixels * 96, "fp64 accumulation") : nullptr;
    A.m2 = (float*)t->alloc(pixels * 4, "luminance moments");
    t->images = (float4*)t->alloc(pixels * 48, "output images");
    launch_clear(t->stream, A, (uint32_t)pixels);
    if (!with_wave) {
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(t->stream));
        return;
    }
    // a wave holds every active pixel at least once
    // several samples of every pixel per wave keep late bounces wide enough
    // to fill the GPU; HBM is plentiful (about 180 bytes per path)
    size_t want = detail.wave_paths ? detail.wave_paths
                                    : std::max(pixels, std::min<size_t>((size_t)1 << 25, pixels * 64));
    want = std::max(want, pixels);
    if (want >= ((size_t)1 << 28)) throw DeviceError{PFXH_REFUSED, "a wave of more than 2^28 paths does not fit the queues"};
    t->capacity = (uint32_t)want;
    size_t cap = t->capacity;
    WaveDev& W = t->W;
    W.capacity = t->capacity;
    W.o = (float4*)t->alloc(cap * 16, "path origins");
    W.d = (float4*)t->alloc(cap * 16, "path directions");
    W.thr = (float4*)t->alloc(cap * 16, "path throughputs");
    W.res = (float4*)t->alloc(cap * 16, "path results");
    W.alb = (float4*)t->alloc(cap * 16, "path albedos");
    W.nrm = (float4*)t->alloc(cap * 16, "path normals");
    W.hit = (float4*)t->alloc(cap * 16, "path hits");
    W.hit_rec = (uint32_t*)t->alloc(cap * 4, "path hit records");
    size_t entries = std::max<uint32_t>(S.entries, 1u);
    W.e_o = (float4*)t->alloc(cap * entries * 16, "shadow ray origins");
    W.e_d = (float4*)t->alloc(cap * entries * 16, "shadow ray directions");
    W.e_c = (float4*)t->alloc(cap * entries * 16, "shadow ray contributions");
    t->qa = (uint32_t*)t->alloc(cap * 4, "ray queue");
    t->qb = (uint32_t*)t->alloc(cap * 4, "ray queue");
    t->qh = (uint32_t*)t->alloc(cap * 4, "hit queue");
    W.q_sh = (uint32_t*)t->alloc(cap * 4, "shadow path queue");
    W.q_sray = (uint32_t*)t->alloc(cap * entries * 4, "shadow ray queue");
    W.counters = (uint32_t*)t->alloc(C_COUNT * 4, "counters");
    W.spill_threads = spill_threads_for();
    W.refill = refill_lanes();
    W.spill = (uint2*)t->alloc((size_t)W.spill_threads * (STACK_DEPTH - STACK_SHORT) * 8, "traversal stacks");
    t->active = (uint32_t*)t->alloc(pixels * 4, "active pixels");
    launch_zero(t->stream, W.counters, C_COUNT);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(t->stream));
}

namespace {

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}

void sample(pfxh_trace* t, uint32_t count, uint32_t seed) {
    if (count == 0) return;
    hipStream_t s = t->stream;
    uint32_t pixels = t->width * t->height;
    if (pixels == 0) {
        t->samples += count;
        return;
    }
    uint32_t acc = t->samples;
    if (acc == 0) {
        launch_clear(s, t->A, pixels);
        t->any_frozen = false;
    }
    const uint32_t* active = nullptr;
    uint32_t npix = pixels;
    if (t->any_frozen) {
        launch_zero(s, t->W.counters + C_ACTIVE, 1);
        launch_active(s, t->A, pixels, t->active, t->W.counters + C_ACTIVE);
        HIP_TRY(hipMemcpyAsync(t->host_counters + C_ACTIVE, t->W.counters + C_ACTIVE, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        npix = t->host_counters[C_ACTIVE];
        active = t->active;
    }
    uint32_t grid = t->W.spill_threads / TRACE_BLOCK;
    uint32_t done = 0;
    while (done < count && npix > 0) {
        uint32_t ns = std::min<uint32_t>(count - done, std::max<uint32_t>(t->capacity / npix, 1u));
        launch_camera(s, t->S, t->W, active, npix, ns, acc + done, seed);
        t->stats.paths += (uint64_t)npix * ns;
        uint32_t n = npix * ns;
        const uint32_t* queue = nullptr;
        uint32_t* outq = t->qa;
        bool shadow_pending = false;
        for (uint32_t bounce = 0; bounce <= t->S.total + 1u && n > 0; ++bounce) {
            WaveDev W = t->W;
            W.q_next = outq;
            launch_zero(s, W.counters + C_FETCH, PARTS * SLOT);
            HIP_TRY(hipEventRecord(t->ev[0], s));
            launch_extend(s, t->S, W, queue, n, t->counting, grid);
            HIP_TRY(hipEventRecord(t->ev[1], s));
            // under a black sky only the paths that hit need shading
            const uint32_t* shade_queue = queue;
            const uint32_t* shade_count = nullptr;
            if (t->S.env_black && !t->S.area && t->qh) {
                launch_zero(s, W.counters + C_HITS, 1);
                launch_hits(s, W, queue, n, t->qh);
                shade_queue = t->qh;
                shade_count = W.counters + C_HITS;
            }
            launch_zero(s, W.counters, C_SRAYS + 1);
            launch_shade(s, t->S, W, shade_queue, n, bounce, shade_count);
            HIP_TRY(hipEventRecord(t->ev[2], s));
            HIP_TRY(hipMemcpyAsync(t->host_counters, W.counters, (C_SRAYS + 1) * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            HIP_TRY(hipGetLastError());
            t->stats.closest_rays += n;
            t->stats.trace_ms += elapsed(t->ev[0], t->ev[1]);
            t->stats.launches += 1;
            t->stats.shade_ms += elapsed(t->ev[1], t->ev[2]);
            if (shadow_pending) t->stats.shadow_ms += elapsed(t->ev[3], t->ev[4]);
            shadow_pending = false;
            uint32_t n_next = t->host_counters[C_NEXT], n_sh = t->host_counters[C_SHADOW];
            uint32_t n_rays = t->host_counters[C_SRAYS];
            if (n_sh) {
                launch_zero(s, W.counters + C_SFETCH, PARTS * SLOT);
                HIP_TRY(hipEventRecord(t->ev[3], s));
                launch_shadow(s, t->S, W, n_rays, t->counting, grid);
                launch_fold(s, W, n_sh);
                HIP_TRY(hipEventRecord(t->ev[4], s));
                t->stats.shadow_rays += n_rays;
                shadow_pending = true;
            }
            queue = outq;
            outq = (outq == t->qa) ? t->qb : t->qa;
            n = n_next;
        }
        HIP_TRY(hipEventRecord(t->ev[5], s));
        launch_finish(s, t->W, t->A, active, npix, ns, (float)(acc + done + ns));
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipGetLastError());
        if (shadow_pending) t->stats.shadow_ms += elapsed(t->ev[3], t->ev[4]);
        done += ns;
    }
    t->samples += count;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

}  // namespace

pfxh_status trace_check(const pfxh_scene* scene, const pfxh_detail* detail, uint32_t width, uint32_t height,
                        char* err, size_t err_len) {
    if (!scene || !detail) {
        put_err(err, err_len, "null argument");
        return PFXH_INVALID;
    }
    if (!finite3(scene->camera.origin) || !finite3(scene->camera.forward) || !finite3(scene->camera.right) ||
        !finite3(scene->camera.up)) {
        put_err(err, err_len, "camera contains invalid parameters");
        return PFXH_SETUP;
    }
    if ((uint64_t)width * height > 0xffffffffull) {
        put_err(err, err_len, "the frame is too large");
        return PFXH_REFUSED;
    }
    if (!projection_valid(detail->projection, detail->ortho_size[0], detail->ortho_size[1]) ||
        detail->triangle_test > PFXH_TRI_WATERTIGHT || detail->walk_order > PFXH_WALK_LEFT_FIRST ||
        detail->sampler > PFXH_SAMPLER_SOBOL) {
        put_err(err, err_len, "invalid trace detail");
        return PFXH_SETUP;
    }
    if (!projection_fits(detail->projection, scene->camera.forward, scene->camera.right, scene->camera.up)) {
        put_err(err, err_len, "an orthographic or equirect camera needs forward, right and up");
        return PFXH_SETUP;
    }
    return PFXH_OK;
}

namespace {
std::atomic<uint64_t> g_max_buffer{0}, g_max_memory{0};
}
uint64_t limit_max_buffer() { return g_max_buffer.load(); }
uint64_t limit_max_memory() { return g_max_memory.load(); }

extern "C" {

void pfxh_set_device_limits(const pfxh_device_limits* limits) {
    g_max_buffer.store(limits ? limits->max_buffer_size : 0);
    g_max_memory.store(limits ? limits->max_memory : 0);
}
void pfxh_get_device_limits(pfxh_device_limits* out) {
    if (!out) return;
    out->max_buffer_size = g_max_buffer.load();
    out->max_memory = g_max_memory.load();
}

const char* pfxh_version(void) { return "pfxhip 0.1.0 (gfx950 wavefront tracer)"; }

pfxh_status pfxh_trace_new(const pfxh_scene* scene, const pfxh_detail* detail, uint32_t width, uint32_t height,
                           pfxh_trace** out, char* err, size_t err_len) {
    if (!out) {
        put_err(err, err_len, "null argument");
        return PFXH_INVALID;
    }
    *out = nullptr;
    pfxh_status chk = trace_check(scene, detail, width, height, err, err_len);
    if (chk != PFXH_OK) return chk;
    pfxh_trace* t = new pfxh_trace();
    pfxh_status st = guarded(err, err_len, [&] { trace_setup(t, *scene, *detail, width, height, true); });
    if (st != PFXH_OK) {
        delete t;
        return st;
    }
    *out = t;
    return PFXH_OK;
}

pfxh_status pfxh_trace_sample(pfxh_trace* t, uint32_t count, uint32_t seed, char* err, size_t err_len) {
    if (!t) return PFXH_INVALID;
    return guarded(err, err_len, [&] { sample(t, count, seed); });
}

pfxh_status pfxh_trace_sample_adaptive(pfxh_trace* t, const pfxh_adaptive* a, uint32_t seed, pfxh_converged* out,
                                       char* err, size_t err_len) {
    if (!t || !a || !out) return PFXH_INVALID;
    if (!(std::isfinite(a->threshold) && a->threshold > 0.0f && a->min_samples >= 2 &&
          a->max_samples >= a->min_samples && std::isfinite(a->growth) && a->growth > 1.0f)) {
        put_err(err, err_len,
                "adaptive sampling needs a positive threshold, a growth over 1 and at least 2 samples");
        return PFXH_INVALID;
    }
    return guarded(err, err_len, [&] {
        uint32_t pixels = t->width * t->height;
        if (!t->snaps) t->snaps = (float4*)t->alloc((size_t)pixels * 32, "convergence snapshots");
        launch_zero(t->stream, (uint32_t*)t->snaps, pixels * 8u);
        t->samples = 0;
        memset(out, 0, sizeof *out);
        out->pixels = pixels;
        // adaptive.rs:46-63 rounds
        std::vector<uint32_t> rounds{std::max<uint32_t>(a->min_samples / 2, 1)};
        for (;;) {
            uint32_t last = rounds.back();
            if (last >= a->max_samples) break;
            float grown = ceilf((float)last * a->growth);
            uint32_t next = grown >= 4294967296.0f ? 0xffffffffu : (uint32_t)grown;
            next = std::max(next, last + 1);
            next = std::max(next, rounds.size() == 1 ? a->min_samples : 0u);
            next = std::min(next, a->max_samples);
            rounds.push_back(next);
        }
        for (size_t step = 0; step < rounds.size(); ++step) {
            uint32_t target = rounds[step];
            uint32_t count = target > t->samples ? target - t->samples : 0;
            sample(t, count, seed);
            launch_measure(t->stream, t->A, t->snaps, t->width, t->height);
            uint64_t running = pixels;
            if (step > 0) {
                launch_zero(t->stream, t->W.counters + C_RUNNING, 1);
                launch_mark(t->stream, t->A, t->snaps, t->width, t->height, a->min_samples, a->threshold,
                            t->W.counters + C_RUNNING);
                HIP_TRY(hipMemcpyAsync(t->host_counters + C_RUNNING, t->W.counters + C_RUNNING, 4,
                                       hipMemcpyDeviceToHost, t->stream));
                HIP_TRY(hipStreamSynchronize(t->stream));
                HIP_TRY(hipGetLastError());
                running = t->host_counters[C_RUNNING];
                t->any_frozen = running < pixels;
            }
            if (out->round_count < PFXH_MAX_ROUNDS) {
                out->samples[out->round_count] = target;
                out->active[out->round_count] = running;
                out->round_count += 1;
            }
            if (running == 0) break;
        }
    });
}

pfxh_status pfxh_trace_readback(pfxh_trace* t, float* color, float* albedo, float* normal, char* err, size_t err_len) {
    if (!t) return PFXH_INVALID;
    return guarded(err, err_len, [&] {
        size_t pixels = (size_t)t->width * t->height;
        if (!pixels) return;
        float4* c = t->images;
        float4* a = t->images + pixels;
        float4* n = t->images + 2 * pixels;
        launch_develop(t->stream, t->A, (uint32_t)pixels, c, a, n);
        if (color) HIP_TRY(hipMemcpyAsync(color, c, pixels * 16, hipMemcpyDeviceToHost, t->stream));
        if (albedo) HIP_TRY(hipMemcpyAsync(albedo, a, pixels * 16, hipMemcpyDeviceToHost, t->stream));
        if (normal) HIP_TRY(hipMemcpyAsync(normal, n, pixels * 16, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        HIP_TRY(hipGetLastError());
    });
}

pfxh_status pfxh_trace_set_camera(pfxh_trace* t, const pfxh_camera* c, char* err, size_t err_len) {
    if (!t || !c) return PFXH_INVALID;
    if (!finite3(c->origin) || !finite3(c->forward) || !finite3(c->right) || !finite3(c->up)) {
        put_err(err, err_len, "camera contains invalid parameters");
        return PFXH_INVALID;
    }
    if (!projection_fits(t->S.projection, c->forward, c->right, c->up)) {
        put_err(err, err_len, "an orthographic or equirect camera needs forward, right and up");
        return PFXH_INVALID;
    }
    for (int k = 0; k < 3; ++k) {
        t->S.origin[k] = c->origin[k];
        t->S.forward[k] = c->forward[k];
        t->S.right[k] = c->right[k];
        t->S.up[k] = c->up[k];
    }
    t->samples = 0;
    t->probe_fresh = false;
    return PFXH_OK;
}

pfxh_status pfxh_trace_set_window(pfxh_trace* t, uint32_t fw, uint32_t fh, uint32_t x, uint32_t y, char* err,
                                  size_t err_len) {
    if (!t) return PFXH_INVALID;
    if ((uint64_t)x + t->width > fw || (uint64_t)y + t->height > fh) {
        put_err(err, err_len, "the window must lie inside the full frame");
        return PFXH_INVALID;
    }
    if ((uint64_t)fw * fh > 0xffffffffull) {
        put_err(err, err_len, "the full frame is too large");
        return PFXH_INVALID;
    }
    t->S.window[0] = fw;
    t->S.window[1] = fh;
    t->S.window[2] = x;
    t->S.window[3] = y;
    t->samples = 0;
    t->probe_fresh = false;
    return PFXH_OK;
}

pfxh_status pfxh_trace_set_sun(pfxh_trace* t, const pfxh_sun* sun, char* err, size_t err_len) {
    if (!t || !sun) return PFXH_INVALID;
    if (!finite3(sun->direction) || !finite3(sun->color) || !std::isfinite(sun->intensity)) {
        put_err(err, err_len, "sun contains invalid parameters");
        return PFXH_INVALID;
    }
    uint32_t needed = (sun->intensity > 0.0f ? 1u : 0u) + (t->host.env_black ? 0u : 1u) +
                      (t->host.lit ? ((t->host.emitter_count > 0 ? 1u : 0u) + t->host.looped + (t->host.tree > 0 ? 1u : 0u)) : 0u);
    if (needed > t->S.entries) {
        put_err(err, err_len, "the trace was built without a sun; rebuild it with one to light it");
        return PFXH_INVALID;
    }
    for (int k = 0; k < 3; ++k) {
        t->S.sun_dir[k] = sun->direction[k];
        t->S.sun_color[k] = sun->color[k];
    }
    t->S.sun_intensity = sun->intensity;
    t->samples = 0;
    t->probe_fresh = false;
    return PFXH_OK;
}

pfxh_status pfxh_trace_set_sun_radius(pfxh_trace* t, float radius_deg, char* err, size_t err_len) {
    if (!t) return PFXH_INVALID;
    if (!std::isfinite(radius_deg) || !(radius_deg >= 0.0f && radius_deg < 90.0f)) {
        put_err(err, err_len, "invalid sun radius");
        return PFXH_INVALID;
    }
    t->S.sun_w = pfx::fm_cos(radius_deg * (3.14159265358979f / 180.0f));
    t->samples = 0;
    return PFXH_OK;
}

pfxh_status pfxh_trace_set_environment(pfxh_trace* t, const pfxh_environment* sky, char* err, size_t err_len) {
    if (!t || !sky) return PFXH_INVALID;
    if (sky->width && (!sky->height || !sky->texels)) {
        put_err(err, err_len, "invalid environment");
        return PFXH_INVALID;
    }
    HostScene next;
    std::string e = lay_sky(*sky, next);
    if (!e.empty()) {
        put_err(err, err_len, e);
        return PFXH_INVALID;
    }
    const HostScene& h = t->host;
    uint32_t needed = (t->S.sun_intensity > 0.0f ? 1u : 0u) + (next.env_black ? 0u : 1u) +
                      (h.lit ? ((h.emitter_count > 0 ? 1u : 0u) + h.looped + (h.tree > 0 ? 1u : 0u)) : 0u);
    if (needed > t->S.entries && (t->wave_shared || needed > 15u)) {
        put_err(err, err_len, "a batch still built under a black sky cannot take a lit one");
        return PFXH_INVALID;
    }
    return guarded(err, err_len, [&] {
        HIP_TRY(hipStreamSynchronize(t->stream));
        SceneDev& S = t->S;
        if (needed > S.entries && t->W.capacity) {
            // one more light slot per path vertex: the wave's entry buffers grow
            size_t cap = t->W.capacity, old = std::max<uint32_t>(S.entries, 1u);
            WaveDev& W = t->W;
            t->release(W.e_o, cap * old * 16);
            t->release(W.e_d, cap * old * 16);
            t->release(W.e_c, cap * old * 16);
            t->release(W.q_sray, cap * old * 4);
            W.e_o = W.e_d = W.e_c = nullptr;
            W.q_sray = nullptr;
            W.e_o = (float4*)t->alloc(cap * needed * 16, "shadow ray origins");
            W.e_d = (float4*)t->alloc(cap * needed * 16, "shadow ray directions");
            W.e_c = (float4*)t->alloc(cap * needed * 16, "shadow ray contributions");
            W.q_sray = (uint32_t*)t->alloc(cap * needed * 4, "shadow ray queue");
        }
        if (needed > S.entries) S.entries = needed;
        t->release(S.sky_rows, t->host.sky_rows.size() * 4);
        t->release(S.sky_columns, t->host.sky_columns.size() * 4);
        t->release(S.env, t->host.env.size() * 4);
        S.sky_rows = S.sky_columns = S.env = nullptr;
        t->host.env = std::move(next.env);
        t->host.env_w = next.env_w;
        t->host.env_h = next.env_h;
        t->host.env_black = next.env_black;
        t->host.sky_rows = std::move(next.sky_rows);
        t->host.sky_columns = std::move(next.sky_columns);
        S.sky_rows = t->upload(t->host.sky_rows, "sky rows");
        S.sky_columns = t->upload(t->host.sky_columns, "sky columns");
        S.env = t->upload(t->host.env, "sky texels");
        S.env_w = t->host.env_w;
        S.env_h = t->host.env_h;
        S.env_black = t->host.env_black ? 1u : 0u;
        t->samples = 0;
    });
}

pfxh_status pfxh_trace_set_lens(pfxh_trace* t, float sx, float sy, float aperture, float focus, char* err,
                                size_t err_len) {
    if (!t) return PFXH_INVALID;
    if (!std::isfinite(sx) || !std::isfinite(sy) || !std::isfinite(aperture) || !std::isfinite(focus) ||
        aperture < 0.0f || (aperture > 0.0f && !(focus > 0.0f))) {
        put_err(err, err_len, "invalid lens");
        return PFXH_INVALID;
    }
    t->S.lens[0] = sx;
    t->S.lens[1] = sy;
    t->S.lens[2] = aperture * 0.5f;
    t->S.lens[3] = focus;
    t->samples = 0;
    t->probe_fresh = false;
    return PFXH_OK;
}

pfxh_status pfxh_trace_set_projection(pfxh_trace* t, uint32_t kind, float w, float h, char* err, size_t err_len) {
    if (!t) return PFXH_INVALID;
    if (!projection_valid(kind, w, h)) {
        put_err(err, err_len, "invalid projection");
        return PFXH_INVALID;
    }
    if (!projection_fits(kind, t->S.forward, t->S.right, t->S.up)) {
        put_err(err, err_len, "an orthographic or equirect camera needs forward, right and up");
        return PFXH_INVALID;
    }
    place_projection(t->S, kind, w, h);
    t->samples = 0;
    t->probe_fresh = false;
    return PFXH_OK;
}

pfxh_status pfxh_trace_save(pfxh_trace* t, const char* key, uint8_t* out, size_t cap, size_t* len, char* err,
                            size_t err_len) {
    if (!t || !key || !len) return PFXH_INVALID;
    size_t pixels = (size_t)t->width * t->height;
    // sizing needs no device work
    std::vector<uint8_t> none(1);
    pfxh_checkpoint c{};
    c.width = t->width;
    c.height = t->height;
    c.samples = t->samples;
    c.round_count = PFXH_NO_ROUNDS;
    c.accum = c.color = c.albedo = c.normal = none.data();
    if (!out) return pfxh_checkpoint_write(&c, key, nullptr, 0, len, err, err_len);
    std::vector<float> accum(pixels * 12), images(pixels * 12);
    pfxh_status st = guarded(err, err_len, [&] {
        if (!pixels) return;
        HIP_TRY(hipMemcpyAsync(accum.data(), t->A.sum, pixels * 48, hipMemcpyDeviceToHost, t->stream));
        std::vector<float> comp;
        std::vector<double> wide;
        if (t->A.mode == PFXH_ACCUM_COMPENSATED) {
            comp.resize(pixels * 12);
            HIP_TRY(hipMemcpyAsync(comp.data(), t->A.comp, pixels * 48, hipMemcpyDeviceToHost, t->stream));
        } else if (t->A.mode == 2u) {
            wide.resize(pixels * 12);
            HIP_TRY(hipMemcpyAsync(wide.data(), t->A.sum64, pixels * 96, hipMemcpyDeviceToHost, t->stream));
        }
        float4* ci = t->images;
        launch_develop(t->stream, t->A, (uint32_t)pixels, ci, ci + pixels, ci + 2 * pixels);
        HIP_TRY(hipMemcpyAsync(images.data(), ci, pixels * 48, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        HIP_TRY(hipGetLastError());
        // fold the compensation (or the FP64 sums) into the FP32 sums; the
        // sample count (color.w) is stored as is
        for (size_t i = 0; i < pixels * 12; ++i) {
            if (i % 12 == 3) continue;
            if (!comp.empty()) accum[i] = accum[i] + comp[i];
            if (!wide.empty()) accum[i] = (float)wide[i];
        }
    });
    if (st != PFXH_OK) return st;
    c.accum = (const uint8_t*)accum.data();
    c.color = (const uint8_t*)images.data();
    c.albedo = c.color + pixels * 16;
    c.normal = c.color + pixels * 32;
    return pfxh_checkpoint_write(&c, key, out, cap, len, err, err_len);
}

pfxh_status pfxh_trace_restore(pfxh_trace* t, const char* key, const uint8_t* bytes, size_t len, char* err,
                               size_t err_len) {
    if (!t || !key || !bytes) return PFXH_INVALID;
    pfxh_checkpoint c{};
    pfxh_status st = pfxh_checkpoint_read(bytes, len, key, &c, nullptr, nullptr, 0, err, err_len);
    if (st != PFXH_OK) return st;
    if (c.width != t->width || c.height != t->height) {
        put_err(err, err_len, "the checkpoint is " + std::to_string(c.width) + "x" + std::to_string(c.height) +
                                  " and the trace " + std::to_string(t->width) + "x" + std::to_string(t->height));
        return PFXH_INVALID;
    }
    size_t pixels = (size_t)t->width * t->height;
    st = guarded(err, err_len, [&] {
        if (!pixels) return;
        HIP_TRY(hipMemcpy(t->A.sum, c.accum, pixels * 48, hipMemcpyHostToDevice));
        if (t->A.mode == PFXH_ACCUM_COMPENSATED) HIP_TRY(hipMemset(t->A.comp, 0, pixels * 48));
        // PFXCKPT1 holds no luminance moments: the restored samples enter the
        // moment sum at their pixel's mean luminance, with no spread, so the
        // deviation reads 0 until new samples add theirs
        std::vector<float> m2(pixels);
        for (size_t p = 0; p < pixels; ++p) {
            float s[4];
            memcpy(s, c.accum + p * 48, 16);
            const float n = fabsf(s[3]), l = (s[0] + s[1] + s[2]) / 3.0f;
            m2[p] = n > 0.0f ? l * l / n : 0.0f;
        }
        HIP_TRY(hipMemcpy(t->A.m2, m2.data(), pixels * 4, hipMemcpyHostToDevice));
        if (t->A.mode == 2u) {
            std::vector<double> wide(pixels * 12);
            for (size_t i = 0; i < wide.size(); ++i) {
                float v;
                memcpy(&v, c.accum + i * 4, 4);
                wide[i] = (double)v;
            }
            HIP_TRY(hipMemcpy(t->A.sum64, wide.data(), pixels * 96, hipMemcpyHostToDevice));
        }
    });
    if (st != PFXH_OK) return st;
    t->samples = c.samples;
    // a negative sample count marks a pixel frozen by adaptive sampling
    t->any_frozen = false;
    for (size_t p = 0; p < pixels && !t->any_frozen; ++p) {
        float w;
        memcpy(&w, c.accum + p * 48 + 12, 4);
        t->any_frozen = w < 0.0f;
    }
    return PFXH_OK;
}

pfxh_status pfxh_trace_readback_deviation(pfxh_trace* t, float* out, char* err, size_t err_len) {
    if (!t || !out) return PFXH_INVALID;
    size_t pixels = (size_t)t->width * t->height;
    return guarded(err, err_len, [&] {
        if (!pixels) return;
        if (!t->deviation) t->deviation = (float*)t->alloc(pixels * 4, "deviation image");
        launch_deviation(t->stream, t->A, (uint32_t)pixels, t->deviation);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, t->deviation, pixels * 4, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
    });
}

pfxh_status pfxh_trace_denoise(pfxh_trace* t, pfxh_denoiser* d, float* out, char* err, size_t err_len) {
    if (!t || !d || !out) return PFXH_INVALID;
    size_t pixels = (size_t)t->width * t->height;
    pfxh_status st = guarded(err, err_len, [&] {
        if (!pixels) return;
        float4* c = t->images;
        launch_develop(t->stream, t->A, (uint32_t)pixels, c, c + pixels, c + 2 * pixels);
        HIP_TRY(hipGetLastError());
    });
    if (st != PFXH_OK || !pixels) return st;
    float4* c = t->images;
    return pfxh_denoise_device(d, t->width, t->height, c, c + pixels, c + 2 * pixels, out, t->stream, err, err_len);
}

pfxh_status pfxh_trace_learned_denoise(pfxh_trace* t, pfxh_learner* l, float* out, char* err, size_t err_len) {
    if (!t || !l || !out) return PFXH_INVALID;
    size_t pixels = (size_t)t->width * t->height;
    pfxh_status st = guarded(err, err_len, [&] {
        if (!pixels) return;
        float4* c = t->images;
        launch_develop(t->stream, t->A, (uint32_t)pixels, c, c + pixels, c + 2 * pixels);
        if (!t->deviation) t->deviation = (float*)t->alloc(pixels * 4, "deviation image");
        launch_deviation(t->stream, t->A, (uint32_t)pixels, t->deviation);
        if (!t->probe) t->probe = (pfxh_probe_hit*)t->alloc(pixels * sizeof(pfxh_probe_hit), "probe");
        if (!t->probe_fresh) {  // depth depends on the view only: kept until a setter changes it
            launch_probe(t->stream, t->S, t->W, t->probe);
            t->probe_fresh = true;
        }
        HIP_TRY(hipGetLastError());
    });
    if (st != PFXH_OK || !pixels) return st;
    float4* c = t->images;
    static_assert(offsetof(pfxh_probe_hit, depth) == 0 && sizeof(pfxh_probe_hit) % 4 == 0, "probe layout");
    return pfxh_learned_denoise_device(l, t->width, t->height, c, c + pixels, c + 2 * pixels, (const float*)t->probe,
                                       (uint32_t)(sizeof(pfxh_probe_hit) / 4), t->deviation, out, t->stream, err,
                                       err_len);
}

pfxh_status pfxh_trace_learned_denoise_checked(pfxh_trace* t, pfxh_learner* l, pfxh_denoiser* d, float threshold,
                                               uint32_t tile, float* out, uint32_t* fallback_tiles, char* err,
                                               size_t err_len) {
    if (!t || !l || !d || !out || tile == 0) return PFXH_INVALID;
    size_t pixels = (size_t)t->width * t->height;
    if (fallback_tiles) *fallback_tiles = 0;
    if (!pixels) return PFXH_OK;
    pfxh_status st = guarded(err, err_len, [&] {
        float4* c = t->images;
        launch_develop(t->stream, t->A, (uint32_t)pixels, c, c + pixels, c + 2 * pixels);
        if (!t->deviation) t->deviation = (float*)t->alloc(pixels * 4, "deviation image");
        launch_deviation(t->stream, t->A, (uint32_t)pixels, t->deviation);
        if (!t->probe) t->probe = (pfxh_probe_hit*)t->alloc(pixels * sizeof(pfxh_probe_hit), "probe");
        if (!t->probe_fresh) {  // depth depends on the view only: kept until a setter changes it
            launch_probe(t->stream, t->S, t->W, t->probe);
            t->probe_fresh = true;
        }
        HIP_TRY(hipGetLastError());
    });
    if (st != PFXH_OK) return st;
    float4* c = t->images;
    void* learned = nullptr;
    st = pfxh_learned_denoise_device(l, t->width, t->height, c, c + pixels, c + 2 * pixels, (const float*)t->probe,
                                     (uint32_t)(sizeof(pfxh_probe_hit) / 4), t->deviation, nullptr, t->stream, err,
                                     err_len, &learned);
    if (st != PFXH_OK) return st;
    uint32_t tiles_x = (t->width + tile - 1) / tile, tiles = tiles_x * ((t->height + tile - 1) / tile);
    std::vector<float> r(tiles);
    std::vector<uint8_t> mask(tiles, 0);
    uint32_t marked = 0;
    st = guarded(err, err_len, [&] {
        // the learned output was written on the learner's stream, which
        // pfxh_learned_denoise_device has drained
        if (!t->tile_buf || t->tile_cap < tiles) {
            t->tile_buf = (float*)t->alloc((size_t)tiles * 8, "tile check");
            t->tile_cap = tiles;
        }
        pfxl::tile_residual(t->stream, c, (const float4*)learned, t->width, t->height, tile, t->tile_buf);
        HIP_TRY(hipMemcpyAsync(r.data(), t->tile_buf, tiles * 4, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        for (uint32_t k = 0; k < tiles; ++k)
            if (!(r[k] <= threshold)) {  // NaN falls back too
                mask[k] = 1;
                ++marked;
            }
    });
    if (st != PFXH_OK) return st;
    if (marked) {
        const void* repo = nullptr;
        st = pfxh_denoise_device(d, t->width, t->height, c, c + pixels, c + 2 * pixels, nullptr, t->stream, err,
                                 err_len, &repo);
        if (st != PFXH_OK) return st;
        st = guarded(err, err_len, [&] {
            HIP_TRY(hipStreamSynchronize(t->stream));
            uint8_t* m = (uint8_t*)(t->tile_buf + tiles);  // the second half of the tile buffer
            HIP_TRY(hipMemcpyAsync(m, mask.data(), tiles, hipMemcpyHostToDevice, t->stream));
            pfxl::tile_merge(t->stream, (float4*)learned, (const float4*)repo, m, t->width, t->height, tile);
            HIP_TRY(hipGetLastError());
        });
        if (st != PFXH_OK) return st;
    }
    st = guarded(err, err_len, [&] {
        HIP_TRY(hipMemcpyAsync(out, learned, pixels * 16, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
    });
    if (fallback_tiles) *fallback_tiles = marked;
    return st;
}

pfxh_status pfxh_trace_atrous(pfxh_trace* t, float* out, char* err, size_t err_len) {
    if (!t || !out) return PFXH_INVALID;
    size_t pixels = (size_t)t->width * t->height;
    return guarded(err, err_len, [&] {
        if (t->samples == 0) throw DeviceError{PFXH_INVALID, "invalid denoise inputs"};
        if (!pixels) return;
        float4* c = t->images;
        launch_develop(t->stream, t->A, (uint32_t)pixels, c, c + pixels, c + 2 * pixels);
        if (!t->atrous) t->atrous = (float4*)t->alloc(pixels * 32, "a-trous images");
        HIP_TRY(hipMemcpyAsync(t->atrous, c, pixels * 16, hipMemcpyDeviceToDevice, t->stream));
        HIP_TRY(pfxh_atrous_device(t->stream, t->atrous, t->atrous + pixels, c + pixels, c + 2 * pixels, t->width,
                                   t->height, t->samples));
        HIP_TRY(hipMemcpyAsync(out, t->atrous, pixels * 16, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        HIP_TRY(hipGetLastError());
    });
}

uint32_t pfxh_trace_samples(const pfxh_trace* t) { return t ? t->samples : 0; }
uint64_t pfxh_trace_gpu_bytes(const pfxh_trace* t) { return t ? t->gpu_bytes : 0; }
uint32_t pfxh_trace_node_bytes(const pfxh_trace* t) {
    if (!t) return 0;
    return t->S.qnodes && (t->S.geo_flags & GEO_LEFT_FIRST) == 0u ? (uint32_t)sizeof(QNode4) : 112u;
}

pfxh_status pfxh_hbm_probe(uint64_t bytes, uint32_t reps, double* gbs, char* err, size_t err_len) {
    if (!gbs || reps == 0 || bytes < (1u << 20) || bytes % 16 != 0) return PFXH_INVALID;
    *gbs = 0.0;
    void* src = nullptr;
    void* dst = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    pfxh_status st = guarded(err, err_len, [&] {
        HIP_TRY(hipMalloc(&src, bytes));
        HIP_TRY(hipMalloc(&dst, bytes));
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&ev[0]));
        HIP_TRY(hipEventCreate(&ev[1]));
        HIP_TRY(hipMemsetAsync(src, 0, bytes, s));
        int device = 0, cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDevice(&device) == hipSuccess && hipGetDeviceProperties(&prop, device) == hipSuccess &&
            prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
        uint64_t n = bytes / 16;
        float best = 0.0f;
        // the best over a few launch shapes: 2, 4 and 8 blocks of 256 per CU
        for (uint32_t per : {2u, 4u, 8u}) {
            uint32_t blocks = (uint32_t)cus * per;
            for (uint32_t r = 0; r <= reps; ++r) {  // the first copy of each shape warms up
                HIP_TRY(hipEventRecord(ev[0], s));
                launch_stream_copy(s, (const float4*)src, (float4*)dst, n, blocks);
                HIP_TRY(hipEventRecord(ev[1], s));
                HIP_TRY(hipEventSynchronize(ev[1]));
                HIP_TRY(hipGetLastError());
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
                if (r > 0 && (best == 0.0f || ms < best)) best = ms;
            }
        }
        *gbs = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
    });
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return st;
}

pfxh_status pfxh_trace_probe(pfxh_trace* t, pfxh_probe_hit* out, char* err, size_t err_len) {
    if (!t || !out) return PFXH_INVALID;
    return guarded(err, err_len, [&] {
        size_t pixels = (size_t)t->width * t->height;
        if (!pixels) return;
        if (!t->probe) t->probe = (pfxh_probe_hit*)t->alloc(pixels * sizeof(pfxh_probe_hit), "probe");
        launch_probe(t->stream, t->S, t->W, t->probe);
        t->probe_fresh = true;
        HIP_TRY(hipMemcpyAsync(out, t->probe, pixels * sizeof(pfxh_probe_hit), hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        HIP_TRY(hipGetLastError());
    });
}

pfxh_status pfxh_trace_cast(pfxh_trace* t, const float* origins, const float* directions, uint32_t count,
                            float limit, uint32_t* primitive, float* distance, char* err, size_t err_len) {
    if (!t || (count && (!origins || !directions || !primitive || !distance)) || !(limit > 0.0f))
        return PFXH_INVALID;
    if (!count) return PFXH_OK;
    void* buf = nullptr;
    pfxh_status st = guarded(err, err_len, [&] {
        size_t n = count;
        HIP_TRY(hipMalloc(&buf, n * 40));
        float4* o = (float4*)buf;
        float4* d = o + n;
        uint32_t* ids = (uint32_t*)(d + n);
        float* dist = (float*)(ids + n);
        std::vector<float4> ho(n), hd(n);
        for (size_t k = 0; k < n; ++k) {
            ho[k] = make_float4(origins[3 * k], origins[3 * k + 1], origins[3 * k + 2], 0.0f);
            hd[k] = make_float4(directions[3 * k], directions[3 * k + 1], directions[3 * k + 2], 0.0f);
        }
        HIP_TRY(hipMemcpyAsync(o, ho.data(), n * 16, hipMemcpyHostToDevice, t->stream));
        HIP_TRY(hipMemcpyAsync(d, hd.data(), n * 16, hipMemcpyHostToDevice, t->stream));
        launch_cast(t->stream, t->S, t->W, o, d, count, limit, ids, dist);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(primitive, ids, n * 4, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipMemcpyAsync(distance, dist, n * 4, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
    });
    if (buf) (void)hipFree(buf);
    return st;
}

pfxh_status pfxh_trace_stats(const pfxh_trace* t, pfxh_stats* out) {
    if (!t || !out) return PFXH_INVALID;
    *out = t->stats;
    uint32_t c[C_COUNT];
    if (hipMemcpy(c, t->W.counters, sizeof c, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return PFXH_DEVICE;
    }
    auto u64 = [&](int k) { return (uint64_t)c[k] | ((uint64_t)c[k + 1] << 32); };
    out->nodes_visited = u64(C_NODES);
    out->prims_tested = u64(C_PRIMS);
    out->boxes_tested = u64(C_BOXES);
    out->shadow_nodes = u64(C_SNODES);
    out->shadow_prims = u64(C_SPRIMS);
    out->rounds_inner = u64(C_WAVE);
    out->rounds_leaf = u64(C_WAVE + 2);
    out->lanes_inner = u64(C_WAVE + 4);
    out->lanes_leaf = u64(C_WAVE + 6);
    out->lanes_live = u64(C_WAVE + 8);
    out->leaf_trips = u64(C_WAVE + 10);
    return PFXH_OK;
}

void pfxh_trace_reset_stats(pfxh_trace* t) {
    if (!t) return;
    t->stats = pfxh_stats{};
    launch_zero(t->stream, t->W.counters, C_COUNT);
    (void)hipStreamSynchronize(t->stream);
}

void pfxh_trace_set_counting(pfxh_trace* t, int on) {
    if (t) t->counting = on != 0;
}

void pfxh_trace_free(pfxh_trace* t) { delete t; }

// ---- host-only layout -------------------------------------------------------------

pfxh_status pfxh_layout_bvh(const pfxh_triangle* tris, uint32_t count, pfxh_bvh_node* nodes, uint32_t* node_count,
                            uint32_t* order, char* err, size_t err_len) {
    if ((!tris && count) || !node_count) return PFXH_INVALID;
    std::vector<float> v((size_t)count * 9);
    for (uint32_t i = 0; i < count; ++i) {
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(tris[i].vertices[k])) {
                put_err(err, err_len, "invalid trace geometry");
                return PFXH_SETUP;
            }
        memcpy(&v[9 * (size_t)i], tris[i].vertices, 36);
    }
    RefBvh b = build_bvh(v.data(), count);
    *node_count = (uint32_t)b.nodes.size();
    if (nodes) memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(pfxh_bvh_node));
    if (order) memcpy(order, b.order.data(), b.order.size() * 4);
    return PFXH_OK;
}

pfxh_status pfxh_layout_nodes(const pfxh_scene* scene, void* nodes, void* qnodes, uint32_t* count,
                              uint32_t* quantized, char* err, size_t err_len) {
    if (!scene || !count || !quantized) return PFXH_INVALID;
    HostScene h;
    std::string e;
    try {
        e = build_scene(*scene, h);
    } catch (const std::bad_alloc&) {
        e = "host memory is exhausted";
    }
    if (!e.empty()) {
        put_err(err, err_len, e);
        return PFXH_SETUP;
    }
    *count = (uint32_t)h.nodes.size();
    *quantized = h.qnodes.empty() ? 0u : 1u;
    if (nodes) memcpy(nodes, h.nodes.data(), h.nodes.size() * sizeof(Node4));
    if (qnodes && !h.qnodes.empty()) memcpy(qnodes, h.qnodes.data(), h.qnodes.size() * sizeof(QNode4));
    return PFXH_OK;
}

pfxh_status pfxh_layout_levels(const pfxh_scene* scene, pfxh_bvh_node* nodes, uint32_t* node_count, uint32_t* order,
                               uint32_t* order_count, char* err, size_t err_len) {
    if (!scene || !node_count || !order_count) return PFXH_INVALID;
    RefLevels levels;
    std::string e;
    try {
        e = lay_levels(*scene, levels, nullptr);
    } catch (const std::bad_alloc&) {
        e = "host memory is exhausted";
    }
    if (!e.empty()) {
        put_err(err, err_len, e);
        return PFXH_SETUP;
    }
    *node_count = (uint32_t)levels.nodes.size();
    *order_count = (uint32_t)levels.world_order.size();
    if (nodes) memcpy(nodes, levels.nodes.data(), levels.nodes.size() * sizeof(pfxh_bvh_node));
    if (order) memcpy(order, levels.world_order.data(), levels.world_order.size() * 4);
    return PFXH_OK;
}

pfxh_status pfxh_layout_lights(const pfxh_light* lights, uint32_t count, float* rows, uint32_t* row_count,
                               uint32_t* looped, uint32_t* tree, char* err, size_t err_len) {
    if ((!lights && count) || !row_count || !looped || !tree) return PFXH_INVALID;
    for (uint32_t i = 0; i < count; ++i)
        if (!light_valid(lights[i])) {
            put_err(err, err_len, "invalid trace light");
            return PFXH_SETUP;
        }
    std::vector<float> r;
    lay_lights(lights, count, r, looped, tree);
    *row_count = (uint32_t)(r.size() / 16);
    if (rows) memcpy(rows, r.data(), r.size() * 4);
    return PFXH_OK;
}

pfxh_status pfxh_layout_select(float* keys, uint32_t* ids, size_t len, size_t index, int mode, size_t* result) {
    if ((len && (!keys || !ids)) || index >= len || mode < 0 || mode > 2 || !result) return PFXH_INVALID;
    *result = select_keys(keys, ids, len, index, mode);
    return PFXH_OK;
}

pfxh_status pfxh_layout_sobol(uint32_t samples, float* table, uint32_t* length) {
    if (!length) return PFXH_INVALID;
    *length = pfx::sobol_length(samples);
    if (table) pfx::sobol_table(*length, table);
    return PFXH_OK;
}

uint32_t pfxh_sobol_animated_seed(uint32_t seed, uint32_t frame) { return pfx::animated_seed(seed, frame); }

pfxh_status pfxh_layout_emitters(const pfxh_scene* scene, float* entries, float* cdf, uint32_t* emitter_count,
                                 char* err, size_t err_len) {
    if (!scene || !emitter_count) return PFXH_INVALID;
    HostScene h;
    std::string e;
    try {
        e = build_scene(*scene, h);
    } catch (const std::bad_alloc&) {
        e = "host memory is exhausted";
    }
    if (!e.empty()) {
        put_err(err, err_len, e);
        return PFXH_SETUP;
    }
    *emitter_count = h.emitter_count;
    if (entries) memcpy(entries, h.emitters.data(), h.emitters.size() * 4);
    if (cdf && h.emitter_count) memcpy(cdf, h.emitter_cdf.data(), (size_t)h.emitter_count * 4);
    return PFXH_OK;
}

}  // extern "C"
