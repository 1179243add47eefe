// CPU restatement of the reference's a-trous denoiser -- TEST INFRASTRUCTURE
// ONLY (checker for the HIP a-trous filter).
//
//   filter     crates/trace/src/denoise.rs:1-62: three passes, steps 1, 2, 4;
//              5x5 kernel [0.375, 0.25, 0.0625]; per tap
//              k(dx) k(dy) exp(-8 |albedo gap|_1 - 4 |normal gap|_1)
//                          exp(-|luma gap| / max(tol, 1e-5)),
//              tol = 2.8 / sqrt(spp) * (luma + 0.05), luma Rec. 709
//              (denoise.rs:64-66); taps outside the image skipped; sum / max(w, 1e-8);
//              alpha kept from the colour.
// exp is taken as exp2(x log2 e) with the shared deterministic exp2, as the
// HIP build does, so both give the same bits; against Rust's f32::exp this
// differs by a few ulp.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <utility>
#include <vector>

#include "../pfx_amd/csrc/fmath.h"

namespace {
float luminance(const float* c) { return c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f; }
float exp_of(float x) { return pfx::fm_exp2(x * 1.44269504088896341f); }
}  // namespace

extern "C" int or_atrous(uint32_t width, uint32_t height, const float* color, const float* albedo, const float* normal,
                         uint32_t samples, float* out) {
    size_t len = (size_t)width * height;
    if (samples == 0) return 1;
    std::vector<float> source(color, color + len * 4), target(source);
    const float kernel[3] = {0.375f, 0.25f, 0.0625f};
    const float scale = 2.8f / sqrtf((float)samples);
    for (int step : {1, 2, 4}) {
        for (int y = 0; y < (int)height; ++y)
            for (int x = 0; x < (int)width; ++x) {
                size_t index = (size_t)y * width + x;
                const float* center = &source[index * 4];
                const float* base = albedo + index * 4;
                const float* n = normal + index * 4;
                float luma = luminance(center);
                float tolerance = scale * (luma + 0.05f);
                float sum[3] = {0.0f, 0.0f, 0.0f}, weight_sum = 0.0f;
                for (int dy = -2; dy <= 2; ++dy)
                    for (int dx = -2; dx <= 2; ++dx) {
                        int xx = x + dx * step, yy = y + dy * step;
                        if (xx < 0 || yy < 0 || xx >= (int)width || yy >= (int)height) continue;
                        size_t other = (size_t)yy * width + xx;
                        const float* b = albedo + other * 4;
                        const float* m = normal + other * 4;
                        float base_gap = (fabsf(base[0] - b[0]) + fabsf(base[1] - b[1])) + fabsf(base[2] - b[2]);
                        float normal_gap = (fabsf(n[0] - m[0]) + fabsf(n[1] - m[1])) + fabsf(n[2] - m[2]);
                        float light_gap = fabsf(luminance(&source[other * 4]) - luma);
                        float guide = exp_of(-base_gap * 8.0f - normal_gap * 4.0f);
                        float light = exp_of(-light_gap / fmaxf(tolerance, 1e-5f));
                        float weight = kernel[dx < 0 ? -dx : dx] * kernel[dy < 0 ? -dy : dy] * guide * light;
                        for (int c = 0; c < 3; ++c) sum[c] += source[other * 4 + c] * weight;
                        weight_sum += weight;
                    }
                for (int c = 0; c < 3; ++c) target[index * 4 + c] = sum[c] / fmaxf(weight_sum, 1e-8f);
            }
        std::swap(source, target);
    }
    memcpy(out, source.data(), len * 16);
    return 0;
}
