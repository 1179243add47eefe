// Kernels of the repo denoiser (OIDN 2.5.1 RT filter, crates/denoise) for
// gfx950.
//
//   k_conv       3x3 convolution as an implicit GEMM on FP32 matrix cores
//                (v_mfma_f32_16x16x4f32): a 256-thread block computes 4 rows
//                of 16 output pixels for every output channel; each wave owns
//                one row (M = 16 pixels) and all N tiles of 16 channels. The
//                K loop walks 16-channel chunks of the input (the upsampled
//                coarser tensor first, then the skip tensor) staged with a
//                1-pixel halo in LDS, and the nine taps. Bias, ReLU and the
//                2x2 max pool are fused into the epilogue; nearest
//                upsampling and concatenation are fused into the staging.
//   k_conv_bf16x32  the same convolution with BF16 operands for the
//                learned denoiser on gfx950's v_mfma_f32_16x16x32_bf16
//                (FP32 sums): inputs rounded to BF16 once as staged, two rows
//                per wave, the next channel pair prefetched while weights and
//                patch are read from LDS, wide layers split over groups of
//                up to four output tiles.
//   k_bins / k_scale   autoexposure (process.wgsl:104-157)
//   k_input / k_output PU transfer in and out of the network
//                      (process.wgsl:33-102)
#include <hip/hip_runtime.h>

#include "denoise.h"

namespace pfxd {

typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

constexpr int TW = 16, TH = 4, PW = TW + 2, PH = TH + 2;

__device__ __forceinline__ uint2 bf16x4_bits(float4 v) {
    return make_uint2((uint32_t)(uint16_t)bf16_bits(v.x) | (uint32_t)(uint16_t)bf16_bits(v.y) << 16,
                      (uint32_t)(uint16_t)bf16_bits(v.z) | (uint32_t)(uint16_t)bf16_bits(v.w) << 16);
}

template <int NT>
__global__ __launch_bounds__(256) void k_conv(ConvDev c) {
    __shared__ float4 patch[PH * PW * 4];    // [row][col][16 channels]
    __shared__ float pooled[2][NT][64][2];   // odd rows' x-pooled values
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, kk = lane >> 4;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const int h = (int)c.h, w = (int)c.w_;
    const int sw = c.upsample ? w / 2 : w;
    {
        // image blockIdx.z of the batch
        const size_t z = blockIdx.z, px = (size_t)h * w;
        const size_t spx = c.upsample ? (size_t)(h / 2) * sw : px;
        c.src0 += z * spx * c.c0;
        if (c.src1) c.src1 += z * px * c.c1;
        c.dst += z * (c.pool ? px / 4 : px) * c.n_pad;
    }
    v4f acc[NT];
    for (int t = 0; t < NT; ++t) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t chunks = (c.c0 + c.c1) / 16u;
    for (uint32_t ch = 0; ch < chunks; ++ch) {
        __syncthreads();
        for (int e = tid; e < PH * PW * 4; e += 256) {
            int q = e & 3, p = e >> 2;
            int py = p / PW, px = p - py * PW;
            int y = y0 + py - 1, x = x0 + px - 1;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (y >= 0 && y < h && x >= 0 && x < w) {
                uint32_t cc = ch * 16u + (uint32_t)q * 4u;
                if (cc < c.c0) {
                    int yy = c.upsample ? y >> 1 : y, xx = c.upsample ? x >> 1 : x;
                    v = *reinterpret_cast<const float4*>(c.src0 + ((size_t)yy * sw + xx) * c.c0 + cc);
                } else {
                    v = *reinterpret_cast<const float4*>(c.src1 + ((size_t)y * w + x) * c.c1 + (cc - c.c0));
                }
            }
            patch[e] = v;
        }
        __syncthreads();
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap - ky * 3;
            const float4 a = patch[((wave + ky) * PW + i + kx) * 4 + kk];
            const float* wt = static_cast<const float*>(c.w) + ((size_t)ch * 9u + tap) * c.n_pad * 16u;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const float4 b = *reinterpret_cast<const float4*>(wt + (size_t)(16 * t + i) * 16u + 4 * kk);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.z, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.w, acc[t], 0, 0, 0);
            }
        }
    }
    // D[pixel 4*kk + v][channel i] of each N tile; bias, ReLU
    const int y = y0 + wave;
    if (!c.pool) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int n = 16 * t + i;
            const float bias = c.bias ? c.bias[n] : 0.0f;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int x = x0 + 4 * kk + v;
                const float r = acc[t][v] + bias;
                if (y < h && x < w) c.dst[((size_t)y * w + x) * c.n_pad + n] = c.relu ? fmaxf(r, 0.0f) : r;
            }
        }
        return;
    }
    // 2x2 max pool: x pairs sit in one lane, row pairs in waves (0,1), (2,3)
    float m[NT][2];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const float bias = c.bias ? c.bias[16 * t + i] : 0.0f;
#pragma unroll
        for (int p = 0; p < 2; ++p)
            m[t][p] = fmaxf(fmaxf(acc[t][2 * p] + bias, acc[t][2 * p + 1] + bias), 0.0f);
    }
    if (wave & 1)
        for (int t = 0; t < NT; ++t)
            for (int p = 0; p < 2; ++p) pooled[wave >> 1][t][lane][p] = m[t][p];
    __syncthreads();
    if (wave & 1) return;
    const int oy = (y0 + wave) / 2, ho = h / 2, wo = w / 2;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int ox = x0 / 2 + 2 * kk + p;
            const float v = fmaxf(m[t][p], pooled[wave >> 1][t][lane][p]);
            if (oy < ho && ox < wo) c.dst[((size_t)oy * wo + ox) * c.n_pad + 16 * t + i] = v;
        }
}

// The epilogue of the BF16 convolutions: bias, ReLU, and the fused 2x2 max
// pool when the layer pools. acc holds D[pixel 4*kk + v][channel i] of each
// N tile for rows y0 + ROWS * wave + r.
template <int NT, int ROWS>
__device__ __forceinline__ void conv_store(const ConvDev& c, v4f (&acc)[ROWS][NT], float (&pooled)[2][NT][64][2],
                                           int x0, int y0, int tiles = NT) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, kk = lane >> 4;
    const int h = (int)c.h, w = (int)c.w_;
    if (!c.pool) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (t >= tiles) break;
            const int n = 16 * t + i;
            const float bias = c.bias ? c.bias[n] : 0.0f;
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                const int y = y0 + ROWS * wave + r;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int x = x0 + 4 * kk + v;
                    const float o = acc[r][t][v] + bias;
                    if (y < h && x < w) c.dst[((size_t)y * w + x) * c.n_pad + n] = c.relu ? fmaxf(o, 0.0f) : o;
                }
            }
        }
        return;
    }
    // 2x2 max pool after the ReLU: x pairs (v 0-1, 2-3) sit in one lane; the
    // row pairs (2j, 2j + 1) sit in one wave (even ROWS) or in waves (0,1),
    // (2,3) (ROWS = 1)
    constexpr int PAIRS = ROWS == 1 ? 1 : ROWS / 2;
    float m[PAIRS][NT][2];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const float bias = c.bias && t < tiles ? c.bias[16 * t + i] : 0.0f;
#pragma unroll
        for (int j = 0; j < PAIRS; ++j)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                float v = fmaxf(acc[2 * j][t][2 * p] + bias, acc[2 * j][t][2 * p + 1] + bias);
                if (ROWS > 1)
                    v = fmaxf(v, fmaxf(acc[2 * j + 1][t][2 * p] + bias, acc[2 * j + 1][t][2 * p + 1] + bias));
                m[j][t][p] = fmaxf(v, 0.0f);
            }
    }
    int oy;
    if (ROWS == 1) {
        if (wave & 1)
            for (int t = 0; t < NT; ++t)
                for (int p = 0; p < 2; ++p) pooled[wave >> 1][t][lane][p] = m[0][t][p];
        __syncthreads();
        if (wave & 1) return;
        for (int t = 0; t < NT; ++t)
            for (int p = 0; p < 2; ++p) m[0][t][p] = fmaxf(m[0][t][p], pooled[wave >> 1][t][lane][p]);
        oy = (y0 + wave) / 2;
    } else {
        oy = (y0 + ROWS * wave) / 2;
    }
    const int ho = h / 2, wo = w / 2;
#pragma unroll
    for (int j = 0; j < PAIRS; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int ox = x0 / 2 + 2 * kk + p;
                if (t < tiles && oy + j < ho && ox < wo)
                    c.dst[((size_t)(oy + j) * wo + ox) * c.n_pad + 16 * t + i] = m[j][t][p];
            }
}


typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
__device__ __forceinline__ v8bf as_v8bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }


// The BF16 convolution on gfx950's K = 32 matrix-core op
// (v_mfma_f32_16x16x32_bf16, FP32 sums): the K loop takes the input
// channels 32 at a time (two 16-channel chunks of the packed weights; an
// odd last chunk pairs with zeros), each lane holding 8 consecutive channels
// of one pixel (A) and of one output's weights (B). A block computes 4 *
// ROWS rows x 16 pixels, each wave ROWS rows, so with ROWS = 2 every weight
// fragment feeds two matrix-core ops and the 2x2 pool sits inside a wave.
// Software-pipelined: the input patch of the next channel pair is loaded
// into registers while the current pair is computed, and the current
// pair's weights are staged in LDS beside its patch, so the matrix cores
// read both operands from LDS and no vector-memory wait (vmcnt, which
// retires loads in order) ties a weight fragment to the prefetch in
// flight. Per pair: a barrier, the registers rounded to BF16 into the patch
// and the weights copied (from L2: every block reads the same ones) into
// LDS, a barrier, the next pair's loads issued, then the nine taps. Used for
// layers of up to 4 N tiles: there it runs 17-24% faster than the same
// kernel reading weight fragments from memory per op, which was 9-16%
// faster than K = 16; the wide layers (one row per wave, few blocks at the
// deep levels) keep K = 16.
template <int NT, int ROWS>
__global__ __launch_bounds__(256) void k_conv_bf16x32(ConvDev c) {
    constexpr int TR = 4 * ROWS, PR = TR + 2;
    constexpr int ITEMS = PR * PW * 4, PER = (ITEMS + 255) / 256;  // patch items: 8 channels of one pixel
    constexpr int WUNITS = 9 * NT * 16 * 4;                       // weight uint4s of a pair: [tap][n][4 x 8 ch]
    __shared__ uint4 patch[ITEMS];
    __shared__ uint4 wl[WUNITS];
    __shared__ float pooled[2][NT][64][2];  // ROWS = 1: odd rows' x-pooled values
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, kk = lane >> 4;
    // blockIdx.x: the 16-pixel column tile and, for layers wider than NT
    // tiles, the group of up to NT output tiles this block computes (the
    // last group may hold fewer: `tiles`)
    const int nt = (int)(c.n_pad / 16u), groups = (nt + NT - 1) / NT, grp = (int)blockIdx.x % groups;
    const int tiles = min(NT, nt - grp * NT);
    const int x0 = ((int)blockIdx.x / groups) * TW, y0 = blockIdx.y * TR;
    const int h = (int)c.h, w = (int)c.w_;
    const int sw = c.upsample ? w / 2 : w;
    {
        const size_t z = blockIdx.z, px = (size_t)h * w;
        const size_t spx = c.upsample ? (size_t)(h / 2) * sw : px;
        c.src0 += z * spx * c.c0;
        if (c.src1) c.src1 += z * px * c.c1;
        c.dst += z * (c.pool ? px / 4 : px) * c.n_pad + grp * NT * 16;
        if (c.bias) c.bias += grp * NT * 16;
    }
    v4f acc[ROWS][NT];
    for (int r = 0; r < ROWS; ++r)
        for (int t = 0; t < NT; ++t) acc[r][t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t cin = c.c0 + c.c1, chunks = cin / 16u, pairs = (chunks + 1u) / 2u;
    float4 pre[PER][2];
    auto fetch = [&](uint32_t g) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = tid + 256 * k;
            float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v1 = v0;
            const int q = e & 3, p = e >> 2;
            const int py = p / PW, px = p - py * PW;
            const int y = y0 + py - 1, x = x0 + px - 1;
            const uint32_t cc = g * 32u + (uint32_t)q * 8u;
            if (e < ITEMS && y >= 0 && y < h && x >= 0 && x < w && cc < cin) {
                const float* at;
                if (cc < c.c0) {
                    const int yy = c.upsample ? y >> 1 : y, xx = c.upsample ? x >> 1 : x;
                    at = c.src0 + ((size_t)yy * sw + xx) * c.c0 + cc;
                } else {
                    at = c.src1 + ((size_t)y * w + x) * c.c1 + (cc - c.c0);
                }
                v0 = *reinterpret_cast<const float4*>(at);
                v1 = *reinterpret_cast<const float4*>(at + 4);
            }
            pre[k][0] = v0;
            pre[k][1] = v1;
        }
    };
    const uint16_t* wg = static_cast<const uint16_t*>(c.w);
    fetch(0);
    for (uint32_t g = 0; g < pairs; ++g) {
        __syncthreads();  // the block is done with pair g - 1's patch and weights
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = tid + 256 * k;
            if (e < ITEMS) {
                const uint2 lo = bf16x4_bits(pre[k][0]), hi = bf16x4_bits(pre[k][1]);
                patch[e] = make_uint4(lo.x, lo.y, hi.x, hi.y);
            }
        }
        // weights of chunks 2g and 2g + 1 (zero past the last chunk) as
        // [tap][n][32 channels]
        for (int u = tid; u < WUNITS; u += 256) {
            const int q = u & 3, n = (u >> 2) % (NT * 16), tap = (u >> 2) / (NT * 16);
            const uint32_t chunk = 2u * g + (uint32_t)(q >> 1);
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (chunk < chunks && n < tiles * 16)
                v = *reinterpret_cast<const uint4*>(
                    wg + ((size_t)(chunk * 9u + (uint32_t)tap) * c.n_pad + (uint32_t)(grp * NT * 16 + n)) * 16u +
                    (size_t)(q & 1) * 8u);
            wl[u] = v;
        }
        __syncthreads();
        if (g + 1u < pairs) fetch(g + 1u);
#pragma unroll 3
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap - ky * 3;
            v8bf a[ROWS];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) a[r] = as_v8bf(patch[((ROWS * wave + r + ky) * PW + i + kx) * 4 + kk]);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (t >= tiles) break;  // block-uniform
                const v8bf b = as_v8bf(wl[(tap * NT * 16 + 16 * t + i) * 4 + kk]);
#pragma unroll
                for (int r = 0; r < ROWS; ++r)
                    acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b, acc[r][t], 0, 0, 0);
            }
        }
    }
    conv_store<NT, ROWS>(c, acc, pooled, x0, y0, tiles);
}

static void launch_conv_fp32(hipStream_t s, const ConvDev& c, dim3 grid) {
    switch (c.n_pad / 16u) {
#define PFXD_CONV(n) \
    case n: hipLaunchKernelGGL((k_conv<n>), grid, dim3(256), 0, s, c); break;
        PFXD_CONV(1) PFXD_CONV(2) PFXD_CONV(3) PFXD_CONV(4) PFXD_CONV(5) PFXD_CONV(6)
        PFXD_CONV(7) PFXD_CONV(8) PFXD_CONV(9) PFXD_CONV(10) PFXD_CONV(11) PFXD_CONV(12)
        PFXD_CONV(13) PFXD_CONV(14)
#undef PFXD_CONV
        default: break;  // the host refuses other widths
    }
}

// Output tiles per block of k_conv_bf16x32 for a layer of nt tiles of 16
// channels: all of them up to four, else as even a split into groups of at
// most four as the count allows (6 = 3 + 3, 7 = 4 + 3, 14 = 4 + 4 + 4 + 2),
// so a block's weights stay within 36 KiB of LDS.
constexpr uint32_t x32_tiles(uint32_t nt) { return (nt + (nt + 3u) / 4u - 1u) / ((nt + 3u) / 4u); }

static void launch_conv_bf16(hipStream_t s, const ConvDev& c, uint32_t z) {
    const uint32_t nt = c.n_pad / 16u, tiles_x = (c.w_ + TW - 1) / TW, bt = x32_tiles(nt);
    // two rows per wave, 8 rows per block
    const dim3 grid(tiles_x * ((nt + bt - 1) / bt), (c.h + 7) / 8, z);
    switch (bt) {
        case 1: hipLaunchKernelGGL((k_conv_bf16x32<1, 2>), grid, dim3(256), 0, s, c); break;
        case 2: hipLaunchKernelGGL((k_conv_bf16x32<2, 2>), grid, dim3(256), 0, s, c); break;
        case 3: hipLaunchKernelGGL((k_conv_bf16x32<3, 2>), grid, dim3(256), 0, s, c); break;
        default: hipLaunchKernelGGL((k_conv_bf16x32<4, 2>), grid, dim3(256), 0, s, c); break;
    }
}

void launch_conv(hipStream_t s, const ConvDev& c) {
    const uint32_t z = c.batch ? c.batch : 1u;
    if (c.bf16) {
        launch_conv_bf16(s, c, z);
    } else {
        launch_conv_fp32(s, c, dim3((c.w_ + TW - 1) / TW, (c.h + TH - 1) / TH, z));
    }
}

// ---- autoexposure (process.wgsl:104-157) --------------------------------------
constexpr float FLT_MAX_ = 3.4028234663852886e38f;
__device__ __forceinline__ float nan_to_zero(float v) { return v != v ? 0.0f : v; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

__global__ __launch_bounds__(256) void k_bins(const float4* color, uint32_t w, uint32_t h, float* bins) {
    __shared__ float sums[256];
    const uint32_t nx = gridDim.x, ny = gridDim.y;
    const uint32_t h0 = blockIdx.y * h / ny, h1 = (blockIdx.y + 1) * h / ny;
    const uint32_t w0 = blockIdx.x * w / nx, w1 = (blockIdx.x + 1) * w / nx;
    const uint32_t yy = h0 + threadIdx.x / 16u, xx = w0 + threadIdx.x % 16u;
    float l = 0.0f;
    if (yy < h1 && xx < w1) {
        float4 c = color[(size_t)yy * w + xx];
        float r = clampf(nan_to_zero(c.x), 0.0f, FLT_MAX_), g = clampf(nan_to_zero(c.y), 0.0f, FLT_MAX_),
              b = clampf(nan_to_zero(c.z), 0.0f, FLT_MAX_);
        l = 0.212671f * r + 0.715160f * g + 0.072169f * b;
    }
    sums[threadIdx.x] = l;
    for (uint32_t k = 128; k > 0; k >>= 1) {
        __syncthreads();
        if (threadIdx.x < k) sums[threadIdx.x] += sums[threadIdx.x + k];
    }
    if (threadIdx.x == 0) bins[blockIdx.y * nx + blockIdx.x] = sums[0] / (float)((h1 - h0) * (w1 - w0));
}

__global__ __launch_bounds__(256) void k_scale(const float* bins, uint32_t count, float* scale) {
    __shared__ float sums[256];
    __shared__ uint32_t counts[256];
    float sum = 0.0f;
    uint32_t n = 0;
    for (uint32_t i = threadIdx.x; i < count; i += 256u) {
        float l = bins[i];
        if (l > 1e-8f) {
            sum += log2f(l);
            n += 1;
        }
    }
    sums[threadIdx.x] = sum;
    counts[threadIdx.x] = n;
    for (uint32_t k = 128; k > 0; k >>= 1) {
        __syncthreads();
        if (threadIdx.x < k) {
            sums[threadIdx.x] += sums[threadIdx.x + k];
            counts[threadIdx.x] += counts[threadIdx.x + k];
        }
    }
    if (threadIdx.x == 0) scale[0] = counts[0] > 0 ? 0.18f / exp2f(sums[0] / (float)counts[0]) : 1.0f;
}

void launch_exposure(hipStream_t s, const float4* color, uint32_t w, uint32_t h, float* bins, float* scale) {
    dim3 grid((w + 15) / 16, (h + 15) / 16);
    hipLaunchKernelGGL(k_bins, grid, dim3(256), 0, s, color, w, h, bins);
    hipLaunchKernelGGL(k_scale, dim3(1), dim3(256), 0, s, bins, grid.x * grid.y, scale);
}

// ---- PU transfer in and out (process.wgsl:33-102) ------------------------------
__device__ __forceinline__ float pu_forward(float y) {
    if (y <= 1.57945760e-06f) return 1.41283765e+03f * y;
    if (y <= 3.22087631e-02f) return 1.64593172e+00f * powf(y, 4.31384981e-01f) + -2.94139609e-03f;
    return 1.92653254e-01f * logf(y + 6.26026094e-03f) + 9.98620152e-01f;
}
__device__ __forceinline__ float pu_inverse(float x) {
    if (x <= 2.23151711e-03f) return x / 1.41283765e+03f;
    if (x <= 3.70974749e-01f) return powf((x - -2.94139609e-03f) / 1.64593172e+00f, 1.0f / 4.31384981e-01f);
    return expf((x - 9.98620152e-01f) / 1.92653254e-01f) - 6.26026094e-03f;
}

// the network input: colour (PU, scaled), albedo, normal (to [0, 1]); 16 channels, 9 used
__global__ __launch_bounds__(256) void k_input(const float4* color, const float4* albedo, const float4* normal,
                                               const float* scale, uint32_t image_w, TileDev t, uint32_t tile_w,
                                               uint32_t tile_h, float norm, float* tensor) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= tile_w * tile_h) return;
    int y = (int)(p / tile_w), x = (int)(p % tile_w);
    int sx = x - t.dst[1], sy = y - t.dst[0];
    float v[16];
    for (int k = 0; k < 16; ++k) v[k] = 0.0f;
    if (sx >= 0 && sx < t.size[1] && sy >= 0 && sy < t.size[0]) {
        size_t at = (size_t)(sy + t.src[0]) * image_w + (uint32_t)(sx + t.src[1]);
        float s = scale[0];
        float4 c = color[at], a = albedo[at], n = normal[at];
        float cv[3] = {c.x, c.y, c.z}, av[3] = {a.x, a.y, a.z}, nv[3] = {n.x, n.y, n.z};
        for (int k = 0; k < 3; ++k) {
            v[k] = pu_forward(clampf(nan_to_zero(cv[k] * s), 0.0f, FLT_MAX_)) * norm;
            v[3 + k] = clampf(nan_to_zero(av[k]), 0.0f, 1.0f);
            v[6 + k] = clampf(nan_to_zero(nv[k]), -1.0f, 1.0f) * 0.5f + 0.5f;
        }
    }
    float4* out = reinterpret_cast<float4*>(tensor + (size_t)p * 16u);
    for (int q = 0; q < 4; ++q) out[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

__global__ __launch_bounds__(256) void k_output(const float* tensor, uint32_t tile_w, const float* scale,
                                                const float4* color, uint32_t image_w, TileDev t, float rcp_norm,
                                                float4* out) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    uint32_t ow = (uint32_t)t.out_size[1];
    if (p >= ow * (uint32_t)t.out_size[0]) return;
    uint32_t y = p / ow, x = p % ow;
    const float* v = tensor + ((size_t)(y + t.out_src[0]) * tile_w + x + t.out_src[1]) * 16u;
    float s = scale[0];
    float inv = s != 0.0f ? 1.0f / s : 0.0f;
    float r[3];
    for (int k = 0; k < 3; ++k) r[k] = pu_inverse(clampf(nan_to_zero(v[k]), 0.0f, FLT_MAX_) * rcp_norm) * inv;
    size_t at = (size_t)(y + t.out_dst[0]) * image_w + x + t.out_dst[1];
    out[at] = make_float4(r[0], r[1], r[2], color[at].w);
}

void launch_input(hipStream_t s, const float4* color, const float4* albedo, const float4* normal, const float* scale,
                  uint32_t image_w, const TileDev& t, uint32_t tile_w, uint32_t tile_h, float norm, float* tensor) {
    uint32_t n = tile_w * tile_h;
    hipLaunchKernelGGL(k_input, dim3((n + 255) / 256), dim3(256), 0, s, color, albedo, normal, scale, image_w, t,
                       tile_w, tile_h, norm, tensor);
}

void launch_output(hipStream_t s, const float* tensor, uint32_t tile_w, const float* scale, const float4* color,
                   uint32_t image_w, const TileDev& t, float rcp_norm, float4* out) {
    uint32_t n = (uint32_t)t.out_size[0] * (uint32_t)t.out_size[1];
    if (!n) return;
    hipLaunchKernelGGL(k_output, dim3((n + 255) / 256), dim3(256), 0, s, tensor, tile_w, scale, color, image_w, t,
                       rcp_norm, out);
}

}  // namespace pfxd
