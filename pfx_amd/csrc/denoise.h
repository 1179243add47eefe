// The repo denoiser (OIDN 2.5.1's RT filter, as crates/denoise runs it) on
// gfx950: layer descriptions shared by the host setup and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfxd {

// One 3x3 convolution of the U-Net. Tensors are NHWC FP32 with channel
// counts padded to multiples of 16 (zeros in the padding).
struct ConvDev {
    const float* src0;   // level `level`, or level + 1 when upsampled
    const float* src1;   // skip connection at level `level` (concat), or null
    float* dst;          // level `level`, or level + 1 when pooled
    const void* w;       // packed: [chunk][tap][n][16 channels of the chunk];
                         // float, or bfloat16 bits when `bf16`
    const float* bias;   // n_pad floats
    uint32_t c0, c1;     // padded channels of src0 and src1
    uint32_t n_pad;      // padded output channels (16 * n-tiles)
    uint32_t h, w_;      // conv resolution (before pooling)
    uint32_t upsample, pool;
    uint32_t relu;       // 1: max(0, x) after the bias
    uint32_t bf16;       // 1: bfloat16 operands (inputs rounded to nearest even
                         // as they are staged), FP32 sums
    uint32_t batch;      // images, each tensor holding them one after another
};

// One tile of the image (tile.rs Tile, [y, x] pairs).
struct TileDev {
    int32_t src[2], dst[2], size[2], out_src[2], out_dst[2], out_size[2];
};

void launch_conv(hipStream_t s, const ConvDev& c);
void launch_exposure(hipStream_t s, const float4* color, uint32_t w, uint32_t h, float* bins, float* scale);
void launch_input(hipStream_t s, const float4* color, const float4* albedo, const float4* normal, const float* scale,
                  uint32_t image_w, const TileDev& t, uint32_t tile_w, uint32_t tile_h, float norm, float* tensor);
void launch_output(hipStream_t s, const float* tensor, uint32_t tile_w, const float* scale, const float4* color,
                   uint32_t image_w, const TileDev& t, float rcp_norm, float4* out);

}  // namespace pfxd
