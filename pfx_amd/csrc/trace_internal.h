// Internal to libpfxhip: the trace object shared by the single-trace entry
// points (trace.cpp) and the batch (batch.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pfxhip.h"
#include "kernels.h"
#include "layout.h"

using namespace pfxh;

inline void put_err(char* err, size_t n, const std::string& s) {
    if (!err || n == 0) return;
    size_t k = std::min(n - 1, s.size());
    memcpy(err, s.data(), k);
    err[k] = 0;
}

struct DeviceError {
    pfxh_status status;
    std::string what;
};

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) throw DeviceError{PFXH_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// Process-wide device limits (pfxh_set_device_limits); 0 means the
// device's own.
uint64_t limit_max_buffer();
uint64_t limit_max_memory();
// The refusal for one allocation of `bytes` on top of `held` bytes already
// held, or an empty string when it fits the limits. Worded as the
// reference's GpuShortfall::Limit (crates/gpu/src/floor.rs:349-351).
inline std::string limit_shortfall(uint64_t bytes, uint64_t held) {
    uint64_t buffer = limit_max_buffer(), memory = limit_max_memory();
    if (buffer && bytes > buffer)
        return "limit max_buffer_size of " + std::to_string(bytes) + " is beyond the device's " +
               std::to_string(buffer);
    if (memory && held + bytes > memory)
        return "limit max_memory of " + std::to_string(held + bytes) + " is beyond the device's " +
               std::to_string(memory);
    return "";
}

struct pfxh_trace {
    HostScene host;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    std::vector<void*> allocations;
    uint64_t gpu_bytes = 0;
    SceneDev S{};
    WaveDev W{};
    AccumDev A{};
    uint32_t* qa = nullptr;
    uint32_t* qb = nullptr;
    uint32_t* qh = nullptr;  // hit queue of a bounce (black sky: misses need no shading)
    uint32_t* active = nullptr;
    uint32_t* host_counters = nullptr;  // pinned
    float4* snaps = nullptr;
    float4* images = nullptr;           // color, albedo, normal for readback
    pfxh_probe_hit* probe = nullptr;
    bool probe_fresh = false;           // probe holds the current view's hits
    float* tile_buf = nullptr;          // per-tile check of the learned denoiser
    float4* atrous = nullptr;           // a-trous passes: two images
    float* deviation = nullptr;         // per-pixel standard error of the mean luminance (lazily)
    uint32_t tile_cap = 0;
    uint32_t width = 0, height = 0, samples = 0;
    uint32_t capacity = 0;
    bool any_frozen = false;
    bool wave_shared = false;  // a batch still: W, qa, qb and active belong to the batch
    bool counting = false;
    pfxh_stats stats{};

    void* alloc(size_t bytes, const char* what) {
        void* p = nullptr;
        if (bytes == 0) bytes = 16;
        std::string short_ = limit_shortfall(bytes, gpu_bytes);
        if (!short_.empty()) throw DeviceError{PFXH_REFUSED, short_ + " (the tracer's " + what + ")"};
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw DeviceError{PFXH_REFUSED, std::string("the tracer's ") + what + " need " + std::to_string(bytes) +
                                                " bytes on the GPU: " + hipGetErrorString(e)};
        }
        allocations.push_back(p);
        gpu_bytes += bytes;
        return p;
    }
    // Free one allocation made by alloc (after the stream has drained).
    void release(const void* p, size_t bytes) {
        auto it = std::find(allocations.begin(), allocations.end(), p);
        if (it == allocations.end()) return;
        (void)hipFree(*it);
        allocations.erase(it);
        gpu_bytes -= std::min<uint64_t>(gpu_bytes, bytes == 0 ? 16 : bytes);
    }
    template <class T>
    T* upload(const std::vector<T>& v, const char* what) {
        T* p = (T*)alloc(v.size() * sizeof(T), what);
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    ~pfxh_trace() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : allocations) (void)hipFree(p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (host_counters) (void)hipHostFree(host_counters);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

template <class F>
pfxh_status guarded(char* err, size_t err_len, F&& f) {
    try {
        f();
        return PFXH_OK;
    } catch (const DeviceError& e) {
        put_err(err, err_len, e.what);
        return e.status;
    } catch (const std::bad_alloc&) {
        put_err(err, err_len, "host memory is exhausted");
        return PFXH_REFUSED;
    }
}

// Scene, accumulation and images of a trace; with `with_wave` also its own
// wave of paths and queues.
void trace_setup(pfxh_trace* t, const pfxh_scene& scene, const pfxh_detail& detail, uint32_t width, uint32_t height,
                 bool with_wave);
// The argument checks of pfxh_trace_new; PFXH_OK when they pass.
pfxh_status trace_check(const pfxh_scene* scene, const pfxh_detail* detail, uint32_t width, uint32_t height,
                        char* err, size_t err_len);
uint32_t spill_threads_for();
uint32_t refill_lanes();
