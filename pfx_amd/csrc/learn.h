// The learned denoiser's kernels (learn.hip) as the host code launches them.
// Tensors are NHWC FP32, channel counts padded to multiples of 16, a batch
// of images one after another.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfxl {

// Weight gradient of one 3x3 convolution: dW[o][pc][tap] over the stacked
// input src0 (upsampled 2x when `up`) ++ src1.
struct DwArgs {
    const float* dpre;   // [batch][h][w][np] gradient before the activation
    const float* src0;   // [batch][h or h/2][w or w/2][c0]
    const float* src1;   // [batch][h][w][c1], or null when c1 == 0
    float* part;         // chunks * 9 * np * (c0 + c1) partial sums
    float* part_b;       // chunks * np partial bias sums
    uint32_t c0, c1, np, h, w, batch, up;
    uint32_t ct;         // 16-channel input tiles per block (set from the plan)
};
// Launch shape of the weight gradient for one layer and frame.
struct DwPlan {
    uint32_t ct, groups, chunks;
    size_t lds;    // FP32-operand kernel
    size_t lds16;  // BF16-operand kernel
};
DwPlan dw_plan(uint32_t np, uint32_t cin, uint32_t batch, uint32_t h, uint32_t w);

void pack(hipStream_t s, const float* master, uint32_t np, uint32_t cin, void* fwd, void* transposed, bool bf16);
void pool_fwd(hipStream_t s, const float* r, uint32_t b, uint32_t h, uint32_t w, uint32_t c, float* p);
void pool_bwd(hipStream_t s, const float* dp, const float* r, uint32_t b, uint32_t h, uint32_t w, uint32_t c,
              float* dr);
void relu_bwd(hipStream_t s, const float* dy, const float* y, size_t n, float* out);
void scatter(hipStream_t s, const float* dx, uint32_t b, uint32_t h, uint32_t w, uint32_t c0, uint32_t c1, bool up,
             float* d0, float* d1);
void weight_grads(hipStream_t s, const DwArgs& a, const DwPlan& p, bool bf16, float* dw, float* db);
// the network input of `mirrors` frames (1, 2 or 4): frame b the image
// mirrored by b (bit 0 in x, bit 1 in y)
void learn_in(hipStream_t s, const float4* color, const float4* albedo, const float4* normal, const float* depth,
              uint32_t depth_stride, const float* deviation, uint32_t w, uint32_t h, uint32_t w16, uint32_t h16,
              uint32_t mirrors, float* x);
uint32_t loss(hipStream_t s, const float* y, const float4* target, size_t n, uint32_t mode, float* dy, float* sums);
void loss_total(hipStream_t s, const float* sums, uint32_t nb, size_t n, float* out);
// the mean radiance over the `mirrors` frames, each mirrored back
void learn_out(hipStream_t s, const float* y, uint32_t w16, uint32_t h16, uint32_t mirrors, const float4* color,
               uint32_t w, uint32_t h, float4* out);
void gather(hipStream_t s, const float* inputs, const float4* targets, size_t crop_px, uint32_t side,
            const uint32_t* picks, uint32_t batch, float* x, float4* t);
void tile_residual(hipStream_t s, const float4* noisy, const float4* den, uint32_t w, uint32_t h, uint32_t tile,
                   float* r);
void tile_merge(hipStream_t s, float4* out, const float4* repo, const uint8_t* mask, uint32_t w, uint32_t h,
                uint32_t tile);
void adam(hipStream_t s, float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
          float eps, int t);

}  // namespace pfxl
