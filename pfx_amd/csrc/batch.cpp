// Batches of stills: many scenes of uneven size, triangle count and sample
// count rendered with one launch per stage.
//
// Each still keeps its own scene buffers and accumulation (a pfxh_trace
// built without a wave of its own); the batch owns one wave of paths shared
// by all of them. The samples still to take are cut into units (a run of
// samples of one still) and the units packed into waves of at most
// `wave_paths` paths; a wave is traced with one k_camera_batch, the bounce
// loop (k_roll / k_extend / k_shade / k_shadow / k_fold per bounce, queue
// lengths kept on the device) and one k_finish_batch, so a wave can be
// captured once in a HIP graph and replayed for every wave. Units keep each
// still's samples in order and lay them out as a single trace does, so
// every still's sums equal those of its own trace at the same seed.
//
// The reference has no multi-scene batch; its closest relative is the
// bake's probe batch, one scene's 3,072 camera faces in one dispatch
// (crates/bake/src/batch.rs:10-14,79-100, batch.wgsl:11-41).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "trace_internal.h"

struct pfxh_batch {
    std::vector<pfxh_trace*> stills;
    std::vector<uint32_t> spp, seeds;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {};
    std::vector<void*> allocations;
    WaveDev W{};
    uint32_t *qa = nullptr, *qb = nullptr, *qh = nullptr, *active = nullptr;
    uint32_t capacity = 0, max_pixels = 0, total_bounces = 0, cus = 256;
    bool left_first = false;  // closest-hit walks in the reference's left-first order
    bool sobol = false;       // the stills use the Sobol sampler
    bool quantized = true;    // every still's nodes are quantized (QNode4)
    SceneDev* scenes = nullptr;
    AccumDev* accums = nullptr;
    WavePlan* waves = nullptr;
    uint32_t wave_cap = 0;
    uint32_t* wave = nullptr;
    pfxh_batch_tuning tune{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    pfxh_batch_tuning graph_tune{};
    uint64_t host_stats[2] = {0, 0};
    uint64_t paths = 0;

    uint64_t held = 0;  // bytes of the batch's own allocations
    void* alloc(size_t bytes, const char* what) {
        void* p = nullptr;
        std::string short_ = limit_shortfall(std::max<size_t>(bytes, 16), held);
        if (!short_.empty()) throw DeviceError{PFXH_REFUSED, short_ + " (the batch's " + what + ")"};
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw DeviceError{PFXH_REFUSED, std::string("the batch's ") + what + " need " + std::to_string(bytes) +
                                                " bytes on the GPU: " + hipGetErrorString(e)};
        }
        allocations.push_back(p);
        held += std::max<size_t>(bytes, 16);
        return p;
    }
    BatchDev dev() const { return BatchDev{scenes, accums, waves, wave}; }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~pfxh_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (pfxh_trace* t : stills) delete t;
        for (void* p : allocations) (void)hipFree(p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Units of the samples [done_i, done_i + count_i) of every still, packed in
// still order into waves of at most `wave_paths` paths.
std::vector<WavePlan> plan_waves(const pfxh_batch* b, const std::vector<uint32_t>& count, uint32_t wave_paths) {
    std::vector<WavePlan> out;
    WavePlan cur{};
    for (uint32_t i = 0; i < b->stills.size(); ++i) {
        const pfxh_trace* t = b->stills[i];
        uint32_t npix = t->width * t->height, left = count[i], acc = t->samples;
        if (npix == 0) continue;
        while (left > 0) {
            uint32_t room = wave_paths - cur.paths;
            uint32_t ns = std::min(left, room / npix);
            if (ns == 0 || cur.units == MAX_UNITS) {
                out.push_back(cur);
                cur = WavePlan{};
                continue;
            }
            Unit& u = cur.u[cur.units++];
            u.sid = i;
            u.first = cur.paths;
            u.pfirst = cur.pixels;
            u.npix = npix;
            u.ns = ns;
            u.acc = acc;
            u.seed = b->seeds[i];
            cur.paths += npix * ns;
            cur.pixels += npix;
            acc += ns;
            left -= ns;
        }
    }
    if (cur.units) out.push_back(cur);
    return out;
}

void check_tuning(const pfxh_batch* b, const pfxh_batch_tuning& t) {
    if (t.wave_paths < b->max_pixels || t.wave_paths > b->capacity)
        throw DeviceError{PFXH_INVALID, "wave_paths must hold the largest still once and fit the batch's wave (" +
                                            std::to_string(b->max_pixels) + ".." + std::to_string(b->capacity) + ")"};
    if (t.refill < 1 || t.refill > 64) throw DeviceError{PFXH_INVALID, "refill must be 1..64"};
    if (t.trace_blocks_per_cu < 1 || t.trace_blocks_per_cu > 16)
        throw DeviceError{PFXH_INVALID, "trace_blocks_per_cu must be 1..16"};
    if (t.shade_blocks_per_cu < 1 || t.shade_blocks_per_cu > 64)
        throw DeviceError{PFXH_INVALID, "shade_blocks_per_cu must be 1..64"};
}

// One wave: camera, the bounce loop, finish, next wave. Everything it reads
// that changes from wave to wave is on the device.
void record_wave(pfxh_batch* b, const pfxh_batch_tuning& t) {
    hipStream_t s = b->stream;
    BatchDev B = b->dev();
    WaveDev W = b->W;
    W.refill = t.refill;
    uint32_t trace_grid = b->cus * t.trace_blocks_per_cu, shade_grid = b->cus * t.shade_blocks_per_cu;
    uint32_t flat_grid = b->cus * 8u;
    launch_camera_batch(s, B, W, flat_grid);
    const uint32_t* queue = nullptr;
    uint32_t* outq = b->qa;
    for (uint32_t bounce = 0; bounce <= b->total_bounces + 1u; ++bounce) {
        WaveDev Wb = W;
        Wb.q_next = outq;
        launch_roll(s, B, Wb, 0, bounce);
        launch_extend_batch(s, B, Wb, queue, trace_grid, b->left_first, b->quantized);
        launch_roll(s, B, Wb, 1, bounce);
        // only paths that hit, or whose still has a sky, need shading
        launch_hits_batch(s, B, Wb, queue, b->qh, flat_grid);
        launch_shade_batch(s, B, Wb, b->qh, bounce, shade_grid, Wb.counters + C_HITS, b->sobol);
        launch_roll(s, B, Wb, 2, bounce);
        launch_shadow_batch(s, B, Wb, trace_grid, b->left_first, b->quantized);
        launch_fold_batch(s, Wb, flat_grid);
        queue = outq;
        outq = (outq == b->qa) ? b->qb : b->qa;
    }
    launch_finish_batch(s, B, W, flat_grid);
    launch_next_wave(s, B);
}

// Trace `count[i]` more samples of every still i.
void render(pfxh_batch* b, const std::vector<uint32_t>& count, const pfxh_batch_tuning& t) {
    check_tuning(b, t);
    for (pfxh_trace* s : b->stills)
        if (s->any_frozen) throw DeviceError{PFXH_INVALID, "a still with frozen pixels samples through its own trace"};
    std::vector<WavePlan> plans = plan_waves(b, count, t.wave_paths);
    hipStream_t s = b->stream;
    for (pfxh_trace* st : b->stills)  // trace streams may hold work on a still's buffers
        HIP_TRY(hipStreamSynchronize(st->stream));
    // the scene table again: a lent trace's setters (camera, sky, sun) change its SceneDev
    std::vector<SceneDev> rows;
    for (pfxh_trace* st : b->stills) rows.push_back(st->S);
    HIP_TRY(hipMemcpy(b->scenes, rows.data(), rows.size() * sizeof(SceneDev), hipMemcpyHostToDevice));
    for (uint32_t i = 0; i < b->stills.size(); ++i) {
        pfxh_trace* st = b->stills[i];
        if (st->samples == 0 && count[i] > 0) {
            launch_clear(s, st->A, st->width * st->height);
            st->any_frozen = false;
        }
    }
    if (plans.empty()) return;
    if (plans.size() > b->wave_cap) {
        uint32_t cap = (uint32_t)plans.size();
        b->waves = (WavePlan*)b->alloc((size_t)cap * sizeof(WavePlan), "wave plans");
        b->wave_cap = cap;
        b->drop_graph();  // the graph's kernels hold the old plan pointer
    }
    HIP_TRY(hipMemcpyAsync(b->waves, plans.data(), plans.size() * sizeof(WavePlan), hipMemcpyHostToDevice, s));
    launch_zero(s, b->wave, 1);
    launch_zero(s, b->W.counters + C_STATS, 4);
    bool use_graph = t.graph != 0;
    if (use_graph && (!b->exec || memcmp(&b->graph_tune, &t, sizeof t) != 0)) {
        b->drop_graph();
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        record_wave(b, t);
        hipError_t e = hipStreamEndCapture(s, &b->graph);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b->graph = nullptr;
            throw DeviceError{PFXH_DEVICE, std::string("capturing the bounce loop: ") + hipGetErrorString(e)};
        }
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_tune = t;
    }
    for (size_t w = 0; w < plans.size(); ++w) {
        if (use_graph)
            HIP_TRY(hipGraphLaunch(b->exec, s));
        else
            record_wave(b, t);
    }
    uint64_t stats[2];
    HIP_TRY(hipMemcpyAsync(stats, b->W.counters + C_STATS, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    b->host_stats[0] += stats[0];
    b->host_stats[1] += stats[1];
    for (uint32_t i = 0; i < b->stills.size(); ++i) {
        pfxh_trace* st = b->stills[i];
        st->samples += count[i];
        b->paths += (uint64_t)st->width * st->height * count[i];
    }
}

}  // namespace

extern "C" {

pfxh_status pfxh_batch_new(const pfxh_batch_still* stills, uint32_t count, const pfxh_detail* detail,
                           pfxh_batch** out, char* err, size_t err_len) {
    if (!stills || !detail || !out || count == 0) {
        put_err(err, err_len, "a batch needs stills and a detail");
        return PFXH_INVALID;
    }
    *out = nullptr;
    for (uint32_t i = 0; i < count; ++i) {
        pfxh_status st = trace_check(stills[i].scene, detail, stills[i].width, stills[i].height, err, err_len);
        if (st != PFXH_OK) return st;
    }
    pfxh_batch* b = new pfxh_batch();
    pfxh_status st = guarded(err, err_len, [&] {
        HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        for (auto& e : b->ev) HIP_TRY(hipEventCreate(&e));
        int device = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&device) == hipSuccess && hipGetDeviceProperties(&prop, device) == hipSuccess &&
            prop.multiProcessorCount > 0)
            b->cus = (uint32_t)prop.multiProcessorCount;
        uint32_t entries = 1;
        uint64_t work = 0;
        for (uint32_t i = 0; i < count; ++i) {
            pfxh_trace* t = new pfxh_trace();
            b->stills.push_back(t);
            trace_setup(t, *stills[i].scene, *detail, stills[i].width, stills[i].height, false);
            b->spp.push_back(stills[i].spp);
            b->seeds.push_back(stills[i].seed);
            b->max_pixels = std::max(b->max_pixels, stills[i].width * stills[i].height);
            entries = std::max(entries, t->S.entries);
            b->quantized = b->quantized && t->S.qnodes != nullptr;
            work += (uint64_t)stills[i].width * stills[i].height * stills[i].spp;
        }
        b->total_bounces = detail->bounces.total;
        b->left_first = detail->walk_order == PFXH_WALK_LEFT_FIRST;
        b->sobol = detail->sampler == PFXH_SAMPLER_SOBOL;
        // the wave: at least the largest still once, at most 2^25 paths
        size_t cap = std::max<size_t>(b->max_pixels, std::min<size_t>((size_t)1 << 25, std::max<uint64_t>(work, 1)));
        if (cap >= ((size_t)1 << 28)) throw DeviceError{PFXH_REFUSED, "a still of more than 2^28 pixels"};
        b->capacity = (uint32_t)cap;
        WaveDev& W = b->W;
        W.capacity = b->capacity;
        W.o = (float4*)b->alloc(cap * 16, "path origins");
        W.d = (float4*)b->alloc(cap * 16, "path directions");
        W.thr = (float4*)b->alloc(cap * 16, "path throughputs");
        W.res = (float4*)b->alloc(cap * 16, "path results");
        W.alb = (float4*)b->alloc(cap * 16, "path albedos");
        W.nrm = (float4*)b->alloc(cap * 16, "path normals");
        W.hit = (float4*)b->alloc(cap * 16, "path hits");
        W.hit_rec = (uint32_t*)b->alloc(cap * 4, "path hit records");
        W.sid = (uint32_t*)b->alloc(cap * 4, "path stills");
        W.e_o = (float4*)b->alloc(cap * entries * 16, "shadow ray origins");
        W.e_d = (float4*)b->alloc(cap * entries * 16, "shadow ray directions");
        W.e_c = (float4*)b->alloc(cap * entries * 16, "shadow ray contributions");
        b->qa = (uint32_t*)b->alloc(cap * 4, "ray queue");
        b->qb = (uint32_t*)b->alloc(cap * 4, "ray queue");
        b->qh = (uint32_t*)b->alloc(cap * 4, "hit queue");
        W.q_sh = (uint32_t*)b->alloc(cap * 4, "shadow path queue");
        W.q_sray = (uint32_t*)b->alloc(cap * entries * 4, "shadow ray queue");
        W.counters = (uint32_t*)b->alloc(C_COUNT * 4, "counters");
        W.spill_threads = spill_threads_for();
        W.refill = refill_lanes();
        W.spill = (uint2*)b->alloc((size_t)W.spill_threads * (STACK_DEPTH - STACK_SHORT) * 8, "traversal stacks");
        b->active = (uint32_t*)b->alloc((size_t)b->max_pixels * 4, "active pixels");
        b->wave = (uint32_t*)b->alloc(4, "wave index");
        launch_zero(b->stream, W.counters, C_COUNT);
        // every still can also sample, probe and read back as a trace of its
        // own, through the batch's wave
        std::vector<SceneDev> scenes;
        std::vector<AccumDev> accums;
        for (pfxh_trace* t : b->stills) {
            t->W = W;
            t->qa = b->qa;
            t->qb = b->qb;
            t->qh = b->qh;
            t->active = b->active;
            t->capacity = b->capacity;
            t->wave_shared = true;
            scenes.push_back(t->S);
            accums.push_back(t->A);
        }
        b->scenes = (SceneDev*)b->alloc(scenes.size() * sizeof(SceneDev), "scene table");
        b->accums = (AccumDev*)b->alloc(accums.size() * sizeof(AccumDev), "accumulation table");
        HIP_TRY(hipMemcpy(b->scenes, scenes.data(), scenes.size() * sizeof(SceneDev), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b->accums, accums.data(), accums.size() * sizeof(AccumDev), hipMemcpyHostToDevice));
        b->tune.wave_paths = b->capacity;
        b->tune.refill = W.refill;
        b->tune.trace_blocks_per_cu = 16;
        b->tune.shade_blocks_per_cu = 8;
        b->tune.graph = 1;
        HIP_TRY(hipStreamSynchronize(b->stream));
    });
    if (st != PFXH_OK) {
        delete b;
        return st;
    }
    *out = b;
    return PFXH_OK;
}

pfxh_status pfxh_batch_render(pfxh_batch* b, char* err, size_t err_len) {
    if (!b) return PFXH_INVALID;
    return guarded(err, err_len, [&] { render(b, b->spp, b->tune); });
}

pfxh_status pfxh_batch_render_samples(pfxh_batch* b, const uint32_t* counts, char* err, size_t err_len) {
    if (!b || !counts) return PFXH_INVALID;
    return guarded(err, err_len, [&] {
        std::vector<uint32_t> c(counts, counts + b->stills.size());
        render(b, c, b->tune);
    });
}

pfxh_status pfxh_batch_tune(pfxh_batch* b, pfxh_batch_tuning* out, char* err, size_t err_len) {
    if (!b) return PFXH_INVALID;
    return guarded(err, err_len, [&] {
        // trial workload: an eighth of every still's samples (at least one);
        // coordinate descent over the wave size, the refill threshold and the
        // traversal blocks per CU, keeping the fastest of each
        std::vector<uint32_t> trial;
        for (uint32_t n : b->spp) trial.push_back(std::max(1u, n / 8u));
        auto time_of = [&](const pfxh_batch_tuning& t) {
            for (pfxh_trace* s : b->stills) s->samples = 0;
            render(b, trial, t);  // once to capture the graph
            for (pfxh_trace* s : b->stills) s->samples = 0;
            HIP_TRY(hipEventRecord(b->ev[0], b->stream));
            render(b, trial, t);
            HIP_TRY(hipEventRecord(b->ev[1], b->stream));
            HIP_TRY(hipEventSynchronize(b->ev[1]));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
            return ms;
        };
        pfxh_batch_tuning best = b->tune;
        float best_ms = time_of(best);
        auto consider = [&](pfxh_batch_tuning t) {
            float ms = time_of(t);
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        };
        for (uint32_t w = 1u << 22; w <= b->capacity; w <<= 1) {
            pfxh_batch_tuning t = best;
            t.wave_paths = w;
            if (w >= b->max_pixels && w != best.wave_paths) consider(t);
        }
        for (uint32_t r : {8u, 24u, 32u}) {
            pfxh_batch_tuning t = best;
            t.refill = r;
            if (r != best.refill) consider(t);
        }
        for (uint32_t g : {8u, 12u}) {
            pfxh_batch_tuning t = best;
            t.trace_blocks_per_cu = g;
            if (g != best.trace_blocks_per_cu) consider(t);
        }
        b->tune = best;
        for (pfxh_trace* s : b->stills) s->samples = 0;  // trials leave nothing behind
        b->host_stats[0] = b->host_stats[1] = 0;
        b->paths = 0;
        if (out) *out = best;
    });
}

void pfxh_batch_get_tuning(const pfxh_batch* b, pfxh_batch_tuning* out) {
    if (b && out) *out = b->tune;
}

pfxh_status pfxh_batch_set_tuning(pfxh_batch* b, const pfxh_batch_tuning* t, char* err, size_t err_len) {
    if (!b || !t) return PFXH_INVALID;
    return guarded(err, err_len, [&] {
        check_tuning(b, *t);
        b->tune = *t;
    });
}

uint32_t pfxh_batch_count(const pfxh_batch* b) { return b ? (uint32_t)b->stills.size() : 0u; }

pfxh_trace* pfxh_batch_trace(pfxh_batch* b, uint32_t index) {
    if (!b || index >= b->stills.size()) return nullptr;
    return b->stills[index];
}

pfxh_status pfxh_batch_stats(const pfxh_batch* b, pfxh_stats* out) {
    if (!b || !out) return PFXH_INVALID;
    memset(out, 0, sizeof(*out));
    out->closest_rays = b->host_stats[0];
    out->shadow_rays = b->host_stats[1];
    out->paths = b->paths;
    return PFXH_OK;
}

void pfxh_batch_reset_stats(pfxh_batch* b) {
    if (!b) return;
    b->host_stats[0] = b->host_stats[1] = 0;
    b->paths = 0;
}

void pfxh_batch_free(pfxh_batch* b) { delete b; }

}  // extern "C"
