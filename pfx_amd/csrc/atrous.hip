// The reference's CPU denoiser `pfx_trace::denoise::filter`
// (crates/trace/src/denoise.rs:1-62) on the GPU: three edge-aware a-trous
// passes (steps 1, 2, 4) of the 5x5 kernel [3/8, 1/4, 1/16]; a tap's weight
// is the kernel times exp(-8 |albedo gap|_1 - 4 |normal gap|_1) times
// exp(-|luma gap| / max(tolerance, 1e-5)) with tolerance
// 2.8 / sqrt(spp) * (luma + 0.05). One thread per pixel per pass; the taps
// are summed in the reference's order (dy outer, dx inner) with the shared
// deterministic exp, so the oracle's CPU build gives the same bits.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/pfxhip.h"
#include "fmath.h"

namespace {

__device__ __forceinline__ float luma(float4 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
__device__ __forceinline__ float exp_of(float x) { return pfx::fm_exp2(x * 1.44269504088896341f); }

__global__ __launch_bounds__(256) void k_atrous(const float4* src, float4* dst, const float4* albedo,
                                                const float4* normal, uint32_t w, uint32_t h, int step,
                                                float tolerance_scale) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= w * h) return;
    const float kernel[3] = {0.375f, 0.25f, 0.0625f};
    int x = (int)(p % w), y = (int)(p / w);
    float4 center = src[p], base = albedo[p], n = normal[p];
    float l = luma(center);
    float tolerance = tolerance_scale * (l + 0.05f);  // 2.8 / sqrt(spp), formed on the host
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, wsum = 0.0f;
    for (int dy = -2; dy <= 2; ++dy) {
        for (int dx = -2; dx <= 2; ++dx) {
            int xx = x + dx * step, yy = y + dy * step;
            if (xx < 0 || yy < 0 || xx >= (int)w || yy >= (int)h) continue;
            uint32_t o = (uint32_t)yy * w + (uint32_t)xx;
            float4 b = albedo[o], m = normal[o], c = src[o];
            float base_gap = (fabsf(base.x - b.x) + fabsf(base.y - b.y)) + fabsf(base.z - b.z);
            float normal_gap = (fabsf(n.x - m.x) + fabsf(n.y - m.y)) + fabsf(n.z - m.z);
            float light_gap = fabsf(luma(c) - l);
            float guide = exp_of(-base_gap * 8.0f - normal_gap * 4.0f);
            float light = exp_of(-light_gap / fmaxf(tolerance, 1e-5f));
            float weight = kernel[dx < 0 ? -dx : dx] * kernel[dy < 0 ? -dy : dy] * guide * light;
            s0 += c.x * weight;
            s1 += c.y * weight;
            s2 += c.z * weight;
            wsum += weight;
        }
    }
    float d = fmaxf(wsum, 1e-8f);
    dst[p] = make_float4(s0 / d, s1 / d, s2 / d, center.w);
}

void put_err(char* err, size_t n, const std::string& s) {
    if (!err || n == 0) return;
    size_t k = std::min(n - 1, s.size());
    memcpy(err, s.data(), k);
    err[k] = 0;
}

}  // namespace

// Internal: the three passes on device images; the result lands in `a`
// (which holds the colour on entry; `b` is scratch of the same size).
hipError_t pfxh_atrous_device(hipStream_t s, float4* a, float4* b, const float4* albedo, const float4* normal,
                              uint32_t w, uint32_t h, uint32_t samples) {
    uint32_t blocks = (w * h + 255u) / 256u;
    float scale = 2.8f / sqrtf((float)samples);
    float4* src = a;
    float4* dst = b;
    for (int step : {1, 2, 4}) {
        hipLaunchKernelGGL(k_atrous, dim3(blocks), dim3(256), 0, s, src, dst, albedo, normal, w, h, step, scale);
        std::swap(src, dst);
    }
    // three passes leave the result in b; bring it back to a
    return hipMemcpyAsync(a, src, (size_t)w * h * 16, hipMemcpyDeviceToDevice, s);
}

// ---- the allocation-free variant ------------------------------------------
// The middle pass's image lives in a file-scope device array instead of
// device memory the call allocates. The array's name is a host-side stand-in
// in host code: its device address comes from hipGetSymbolAddress, once per
// device, with that device current, and only that address is handed to the
// kernels.
constexpr uint64_t ATROUS_SCRATCH_PIXELS = 1ull << 21;  // 32 MiB: 1920x1080 fits
__device__ __align__(256) float4 scratch[ATROUS_SCRATCH_PIXELS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_lock;
float4* scratch_at[MAX_DEVICES] = {};

hipError_t scratch_of_current_device(float4** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(scratch_lock);
    if (!scratch_at[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_at[dev] = (float4*)ptr;
    }
    *out = scratch_at[dev];
    return hipSuccess;
}
}  // namespace

extern "C" pfxh_status pfxh_atrous_noalloc(uint32_t width, uint32_t height, const float* color, const float* albedo,
                                           const float* normal, uint32_t samples, float* out, void* stream,
                                           char* err, size_t err_len) {
    if (!color || !albedo || !normal || !out) return PFXH_INVALID;
    uint64_t len = (uint64_t)width * height;
    if (len > ATROUS_SCRATCH_PIXELS) {
        put_err(err, err_len, "image larger than the 2097152 pixels of the allocation-free a-trous filter");
        return PFXH_INVALID;
    }
    if (samples == 0) {
        put_err(err, err_len, "invalid denoise inputs");
        return PFXH_INVALID;
    }
    if ((const float*)out == color || (const float*)out == albedo || (const float*)out == normal) {
        put_err(err, err_len, "the output must not alias an input");
        return PFXH_INVALID;
    }
    if (len == 0) return PFXH_OK;
    float4* mid = nullptr;
    hipError_t e = scratch_of_current_device(&mid);
    if (e == hipSuccess) {
        hipStream_t s = (hipStream_t)stream;
        uint32_t blocks = (uint32_t)((len + 255u) / 256u);
        float scale = 2.8f / sqrtf((float)samples);
        const float4* a = (const float4*)albedo;
        const float4* n = (const float4*)normal;
        float4* o = (float4*)out;
        // colour -> out (step 1), out -> scratch (step 2), scratch -> out (step 4)
        hipLaunchKernelGGL(k_atrous, dim3(blocks), dim3(256), 0, s, (const float4*)color, o, a, n, width, height, 1,
                           scale);
        hipLaunchKernelGGL(k_atrous, dim3(blocks), dim3(256), 0, s, (const float4*)o, mid, a, n, width, height, 2,
                           scale);
        hipLaunchKernelGGL(k_atrous, dim3(blocks), dim3(256), 0, s, (const float4*)mid, o, a, n, width, height, 4,
                           scale);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        put_err(err, err_len, std::string("a-trous filter: ") + hipGetErrorString(e));
        return PFXH_DEVICE;
    }
    return PFXH_OK;
}

extern "C" pfxh_status pfxh_atrous(uint32_t width, uint32_t height, const float* color, const float* albedo,
                                   const float* normal, uint32_t samples, float* out, char* err, size_t err_len) {
    if (!color || !albedo || !normal || !out) return PFXH_INVALID;
    uint64_t len = (uint64_t)width * height;
    if (len > 0xffffffffull) {
        put_err(err, err_len, "image dimensions overflow");
        return PFXH_INVALID;
    }
    if (samples == 0) {
        put_err(err, err_len, "invalid denoise inputs");
        return PFXH_INVALID;
    }
    if (len == 0) return PFXH_OK;
    size_t bytes = (size_t)len * 16;
    float4* buf = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&buf, bytes * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(buf, color, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(buf + 2 * len, albedo, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(buf + 3 * len, normal, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = pfxh_atrous_device(s, buf, buf + len, buf + 2 * len, buf + 3 * len, width, height, samples);
    if (e == hipSuccess) e = hipMemcpyAsync(out, buf, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    pfxh_status st = PFXH_OK;
    if (e != hipSuccess) {
        (void)hipGetLastError();
        put_err(err, err_len, std::string("a-trous filter: ") + hipGetErrorString(e));
        st = e == hipErrorOutOfMemory ? PFXH_REFUSED : PFXH_DEVICE;
    }
    if (buf) (void)hipFree(buf);
    if (s) (void)hipStreamDestroy(s);
    return st;
}
