// Kernels of the learned denoiser: training and inference of a U-Net on the
// tracer's own renders (noisy radiance + albedo, normal and depth guides in,
// log radiance out). The convolutions are k_conv (denoise.hip); here are the
// pieces around them:
//
//   k_pack_fwd / k_pack_t   master weights [o][stacked input channel][tap]
//                           into k_conv's layout, as is (forward) or
//                           transposed with the taps mirrored (input gradient)
//   k_pool_fwd / k_pool_bwd 2x2 max pool and its gradient (to the first max)
//   k_relu_bwd              gradient through ReLU (the output is kept)
//   k_scatter               the stacked input gradient back to the upsampled
//                           coarser tensor (2x2 sums) and the skip tensor
//   k_dw32 / k_dw           weight and bias gradients: for each tap a GEMM
//                           over pixels on the matrix cores (BF16 operands
//                           on the K = 32 op, or FP32 operands), partial
//                           sums per pixel chunk, then k_dw_sum / k_db_sum
//   k_learn_in / k_learn_out / k_loss   input encoding, decoding, the loss (L1 on
//                           log radiance, or relative L2 of linear radiance)
//                           and its gradient
//   k_adam                  Adam (Kingma & Ba 2015) with bias correction
#include <hip/hip_runtime.h>

#include <algorithm>

#include "learn.h"

namespace pfxl {

typedef float v4f __attribute__((ext_vector_type(4)));

static inline uint32_t blocks(size_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

// ---- weight layouts ---------------------------------------------------------------
// master M[o][pc][tap], o < np (padded outputs), pc < cin (padded stacked
// inputs), into k_conv's [chunk][tap][n][16] as float or bfloat16 bits
__device__ __forceinline__ uint16_t bf16_round(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
template <bool BF16>
__global__ void k_pack_fwd(const float* M, uint32_t np, uint32_t cin, void* W) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)np * cin * 9;
    if (e >= total) return;
    uint32_t q = e % 16, rest = (uint32_t)(e / 16);
    uint32_t n = rest % np, rest2 = rest / np;
    uint32_t tap = rest2 % 9, chunk = rest2 / 9;
    uint32_t pc = chunk * 16 + q;
    float v = M[((size_t)n * cin + pc) * 9 + tap];
    if (BF16)
        static_cast<uint16_t*>(W)[e] = bf16_round(v);
    else
        static_cast<float*>(W)[e] = v;
}
// transposed: inputs are the np output channels, outputs the cin channels
template <bool BF16>
__global__ void k_pack_t(const float* M, uint32_t np, uint32_t cin, void* W) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)np * cin * 9;
    if (e >= total) return;
    uint32_t q = e % 16, rest = (uint32_t)(e / 16);
    uint32_t pc = rest % cin, rest2 = rest / cin;
    uint32_t tap = rest2 % 9, chunk = rest2 / 9;
    uint32_t o = chunk * 16 + q;
    float v = M[((size_t)o * cin + pc) * 9 + (8 - tap)];
    if (BF16)
        static_cast<uint16_t*>(W)[e] = bf16_round(v);
    else
        static_cast<float*>(W)[e] = v;
}
void pack(hipStream_t s, const float* M, uint32_t np, uint32_t cin, void* fwd, void* t, bool bf16) {
    size_t n = (size_t)np * cin * 9;
    if (bf16) {
        hipLaunchKernelGGL(k_pack_fwd<true>, dim3(blocks(n, 256)), dim3(256), 0, s, M, np, cin, fwd);
        hipLaunchKernelGGL(k_pack_t<true>, dim3(blocks(n, 256)), dim3(256), 0, s, M, np, cin, t);
    } else {
        hipLaunchKernelGGL(k_pack_fwd<false>, dim3(blocks(n, 256)), dim3(256), 0, s, M, np, cin, fwd);
        hipLaunchKernelGGL(k_pack_t<false>, dim3(blocks(n, 256)), dim3(256), 0, s, M, np, cin, t);
    }
}

// ---- pooling and activations ------------------------------------------------------
__global__ void k_pool_fwd(const float* R, uint32_t h, uint32_t w, uint32_t c, size_t n, float* P) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    uint32_t ch = e % c;
    size_t px = e / c;
    uint32_t wo = w / 2, ho = h / 2;
    uint32_t ox = px % wo, oy = (px / wo) % ho;
    size_t b = px / ((size_t)wo * ho);
    const float* r = R + ((b * h + 2 * oy) * w + 2 * ox) * c + ch;
    P[e] = fmaxf(fmaxf(r[0], r[c]), fmaxf(r[(size_t)w * c], r[(size_t)w * c + c]));
}
// dR at the first maximum of each window, times the ReLU mask of R
__global__ void k_pool_bwd(const float* dP, const float* R, uint32_t h, uint32_t w, uint32_t c, size_t n, float* dR) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    uint32_t ch = e % c;
    size_t px = e / c;
    uint32_t wo = w / 2, ho = h / 2;
    uint32_t ox = px % wo, oy = (px / wo) % ho;
    size_t b = px / ((size_t)wo * ho);
    size_t at[4] = {((b * h + 2 * oy) * w + 2 * ox) * c + ch, ((b * h + 2 * oy) * w + 2 * ox + 1) * c + ch,
                    ((b * h + 2 * oy + 1) * w + 2 * ox) * c + ch, ((b * h + 2 * oy + 1) * w + 2 * ox + 1) * c + ch};
    int best = 0;
    for (int k = 1; k < 4; ++k)
        if (R[at[k]] > R[at[best]]) best = k;
    float g = R[at[best]] > 0.0f ? dP[e] : 0.0f;
    for (int k = 0; k < 4; ++k) dR[at[k]] = k == best ? g : 0.0f;
}
__global__ void k_relu_bwd(const float* dY, const float* Y, size_t n, float* out) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) out[e] = Y[e] > 0.0f ? dY[e] : 0.0f;
}
void pool_fwd(hipStream_t s, const float* R, uint32_t b, uint32_t h, uint32_t w, uint32_t c, float* P) {
    size_t n = (size_t)b * (h / 2) * (w / 2) * c;
    hipLaunchKernelGGL(k_pool_fwd, dim3(blocks(n, 256)), dim3(256), 0, s, R, h, w, c, n, P);
}
void pool_bwd(hipStream_t s, const float* dP, const float* R, uint32_t b, uint32_t h, uint32_t w, uint32_t c,
              float* dR) {
    size_t n = (size_t)b * (h / 2) * (w / 2) * c;
    hipLaunchKernelGGL(k_pool_bwd, dim3(blocks(n, 256)), dim3(256), 0, s, dP, R, h, w, c, n, dR);
}
void relu_bwd(hipStream_t s, const float* dY, const float* Y, size_t n, float* out) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(n, 256)), dim3(256), 0, s, dY, Y, n, out);
}

// ---- the stacked input gradient back to its sources --------------------------------
// dX [b][h][w][c0 + c1]; d0 += its first c0 channels (2x2 sums when src0 was
// upsampled, d0 then being [b][h/2][w/2][c0]); d1 += the rest
__global__ void k_scatter0(const float* dX, uint32_t h, uint32_t w, uint32_t c0, uint32_t c1, uint32_t up, size_t n,
                           float* d0) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    uint32_t ch = e % c0;
    size_t px = e / c0;
    uint32_t cs = c0 + c1;
    if (!up) {
        d0[e] += dX[px * cs + ch];
        return;
    }
    uint32_t wo = w / 2, ho = h / 2;
    uint32_t ox = px % wo, oy = (px / wo) % ho;
    size_t b = px / ((size_t)wo * ho);
    size_t r0 = ((b * h + 2 * oy) * w + 2 * ox) * cs + ch, r1 = r0 + (size_t)w * cs;
    d0[e] += (dX[r0] + dX[r0 + cs]) + (dX[r1] + dX[r1 + cs]);
}
__global__ void k_scatter1(const float* dX, uint32_t c0, uint32_t c1, size_t n, float* d1) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    uint32_t ch = e % c1;
    size_t px = e / c1;
    d1[e] += dX[px * (c0 + c1) + c0 + ch];
}
void scatter(hipStream_t s, const float* dX, uint32_t b, uint32_t h, uint32_t w, uint32_t c0, uint32_t c1, bool up,
             float* d0, float* d1) {
    if (d0) {
        size_t n = (size_t)b * (up ? (h / 2) * (w / 2) : h * w) * c0;
        hipLaunchKernelGGL(k_scatter0, dim3(blocks(n, 256)), dim3(256), 0, s, dX, h, w, c0, c1, up ? 1u : 0u, n, d0);
    }
    if (d1 && c1) {
        size_t n = (size_t)b * h * w * c1;
        hipLaunchKernelGGL(k_scatter1, dim3(blocks(n, 256)), dim3(256), 0, s, dX, c0, c1, n, d1);
    }
}

// ---- weight and bias gradients ------------------------------------------------------
// dW_tap[o][c] = sum over pixels p of dPre[p][o] * X[p + offset(tap)][c]: per
// tap a GEMM with M = outputs, N = stacked inputs and K = pixels, split
// over K into chunks. A block (4 waves) takes a segment of 64 pixels of one
// row at a time: dPre [64 px][np] and the input rows y-1..y+1 over 66 px
// for the block's `ct` 16-channel input tiles are staged in LDS with float4
// loads, then each wave adds the segment to its (16 outputs, 16 inputs)
// tiles for the nine taps with FP32 operands -- four v_mfma_f32_16x16x4_f32
// per 4 pixels (A[o][k] = dPre[k][o], B[k][c] = X[k][c]). Blocks of the
// first input group also sum dPre over the segment for the bias gradient.
// Partial sums per chunk; k_dw_sum and k_db_sum add them.
constexpr int SEG = 64;

__global__ __launch_bounds__(256) void k_dw(DwArgs a) {
    extern __shared__ float lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, kk = lane >> 4;
    const uint32_t np = a.np, cin = a.c0 + a.c1, ot_n = np / 16, ct_all = cin / 16;
    const uint32_t ct0 = blockIdx.x * a.ct, ctn = min(a.ct, ct_all - ct0), pairs = ot_n * ctn;
    const uint32_t dstride = np + 4, xw = a.ct * 16 + 4;
    float* D = lds;
    float* X = lds + SEG * dstride;
    const uint32_t h = a.h, w = a.w, segw = (w + SEG - 1) / SEG;
    const size_t segs = (size_t)a.batch * h * segw;
    const size_t s0 = segs * blockIdx.y / gridDim.y, s1 = segs * (blockIdx.y + 1) / gridDim.y;
    const uint32_t sh = a.up ? h / 2 : h, sw = a.up ? w / 2 : w;
    const bool bias_block = blockIdx.x == 0 && (uint32_t)tid < np;
    float bsum = 0.0f;
    v4f acc[2][9];
    for (int p = 0; p < 2; ++p)
        for (int t = 0; t < 9; ++t) acc[p][t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t q4 = ctn * 4;  // float4s per staged input pixel
    for (size_t sg = s0; sg < s1; ++sg) {
        const uint32_t xs = (uint32_t)(sg % segw) * SEG;
        const uint32_t y = (uint32_t)((sg / segw) % h);
        const size_t b = sg / ((size_t)segw * h);
        __syncthreads();
        for (uint32_t e = tid; e < SEG * np / 4; e += 256) {
            uint32_t px = e / (np / 4), q = e % (np / 4), x = xs + px;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (x < w) v = *reinterpret_cast<const float4*>(a.dpre + ((b * h + y) * w + x) * np + 4 * q);
            float* d = D + px * dstride + 4 * q;
            d[0] = v.x, d[1] = v.y, d[2] = v.z, d[3] = v.w;
        }
        for (uint32_t e = tid; e < 3 * (SEG + 2) * q4; e += 256) {
            uint32_t q = e % q4, rest = e / q4, px = rest % (SEG + 2), r = rest / (SEG + 2);
            int yy = (int)y + (int)r - 1, xx = (int)xs + (int)px - 1;
            uint32_t cc = ct0 * 16 + 4 * q;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (yy >= 0 && yy < (int)h && xx >= 0 && xx < (int)w) {
                if (cc < a.c0) {
                    int sy = a.up ? yy >> 1 : yy, sx = a.up ? xx >> 1 : xx;
                    v = *reinterpret_cast<const float4*>(a.src0 + ((b * sh + (uint32_t)sy) * sw + (uint32_t)sx) * a.c0 + cc);
                } else {
                    v = *reinterpret_cast<const float4*>(a.src1 + ((b * h + (uint32_t)yy) * w + (uint32_t)xx) * a.c1 +
                                                         (cc - a.c0));
                }
            }
            float* d = X + (r * (SEG + 2) + px) * xw + 4 * q;
            d[0] = v.x, d[1] = v.y, d[2] = v.z, d[3] = v.w;
        }
        __syncthreads();
        if (bias_block)
            for (int px = 0; px < SEG; ++px) bsum += D[px * dstride + tid];
        for (int s = 0; s < SEG / 16; ++s) {
            if (xs + 16u * s >= w) break;
#pragma unroll
            for (int pi = 0; pi < 2; ++pi) {
                const uint32_t p = (uint32_t)wave + 4u * pi;
                if (p >= pairs) continue;
                const uint32_t ot = p % ot_n, ctl = p / ot_n;
                const float* dcol = D + ot * 16 + i;
                const float* xcol = X + ctl * 16 + i;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int k = 16 * s + 4 * g + kk;
                    const float av = dcol[k * dstride];
#pragma unroll
                    for (int tap = 0; tap < 9; ++tap) {
                        const int ky = tap / 3, kx = tap - ky * 3;
                        const float bv = xcol[(ky * (SEG + 2) + k + kx) * xw];
                        acc[pi][tap] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[pi][tap], 0, 0, 0);
                    }
                }
            }
        }
    }
    // lane: input column c = i, output rows 4 kk + v
#pragma unroll
    for (int pi = 0; pi < 2; ++pi) {
        const uint32_t p = (uint32_t)wave + 4u * pi;
        if (p >= pairs) continue;
        const uint32_t ot = p % ot_n, ctl = p / ot_n;
        for (int tap = 0; tap < 9; ++tap)
            for (int v = 0; v < 4; ++v)
                a.part[(((size_t)blockIdx.y * 9 + tap) * np + ot * 16 + 4 * kk + v) * cin + (ct0 + ctl) * 16 + i] =
                    acc[pi][tap][v];
    }
    if (bias_block) a.part_b[(size_t)blockIdx.y * np + tid] = bsum;
}
// BF16 operands on gfx950's K = 32 matrix-core op (v_mfma_f32_16x16x32_bf16):
// the same GEMMs with K = 32 pixels per instruction. A segment's dPre and
// input rows are rounded to BF16 once as they are staged, pixel-contiguous
// per channel, so each lane takes its 8 pixels of an operand with one
// 16-byte LDS read: A = Dt[o][p0 + 8g ..], and for the three taps of an
// input row one 20-byte read gives B at pixel offsets 0, 1 and 2 (the odd
// offset by a 2-byte funnel shift). The bias sums stay FP32: each staging
// item sums its 8 pixels, and the bias threads add the items in pixel order.
constexpr uint32_t XQ = SEG + 8;  // staged input pixels per row: SEG + 2 halo, padded to 16 bytes
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
__device__ __forceinline__ v8bf as_v8bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }
__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
    return (uint32_t)bf16_round(lo) | ((uint32_t)bf16_round(hi) << 16);
}

__global__ __launch_bounds__(256) void k_dw32(DwArgs a) {
    extern __shared__ uint4 lds16[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, g = lane >> 4;
    const uint32_t np = a.np, cin = a.c0 + a.c1, ot_n = np / 16, ct_all = cin / 16;
    const uint32_t ct0 = blockIdx.x * a.ct, ctn = min(a.ct, ct_all - ct0), pairs = ot_n * ctn;
    uint16_t* Dt = reinterpret_cast<uint16_t*>(lds16);  // [np][SEG]
    uint16_t* Xt = Dt + (size_t)np * SEG;                // [3 rows][a.ct * 16 channels][XQ]
    const uint32_t xrows = a.ct * 16;
    float* bpart = reinterpret_cast<float*>(Xt + (size_t)3 * xrows * XQ);  // [SEG / 8][np]
    const uint32_t h = a.h, w = a.w, segw = (w + SEG - 1) / SEG;
    const size_t segs = (size_t)a.batch * h * segw;
    const size_t s0 = segs * blockIdx.y / gridDim.y, s1 = segs * (blockIdx.y + 1) / gridDim.y;
    const uint32_t sh = a.up ? h / 2 : h, sw = a.up ? w / 2 : w;
    const bool bias_block = blockIdx.x == 0 && (uint32_t)tid < np;
    float bsum = 0.0f;
    v4f acc[2][9];
    for (int p = 0; p < 2; ++p)
        for (int t = 0; t < 9; ++t) acc[p][t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    for (size_t sg = s0; sg < s1; ++sg) {
        const uint32_t xs = (uint32_t)(sg % segw) * SEG;
        const uint32_t y = (uint32_t)((sg / segw) % h);
        const size_t b = sg / ((size_t)segw * h);
        const float* drow = a.dpre + (b * h + y) * w * np;
        __syncthreads();
        // dPre: items of 4 outputs x 8 pixels
        for (uint32_t e = tid; e < (np / 4) * (SEG / 8); e += 256) {
            const uint32_t oq = e % (np / 4), po = e / (np / 4);
            float4 v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t x = xs + 8 * po + j;
                v[j] = x < w ? *reinterpret_cast<const float4*>(drow + (size_t)x * np + 4 * oq)
                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            float4 bs = v[0];
#pragma unroll
            for (int j = 1; j < 8; ++j) bs = make_float4(bs.x + v[j].x, bs.y + v[j].y, bs.z + v[j].z, bs.w + v[j].w);
            *reinterpret_cast<float4*>(bpart + (size_t)po * np + 4 * oq) = bs;
            uint16_t* d = Dt + (size_t)(4 * oq) * SEG + 8 * po;
            *reinterpret_cast<uint4*>(d) = make_uint4(bf16_pair(v[0].x, v[1].x), bf16_pair(v[2].x, v[3].x),
                                                      bf16_pair(v[4].x, v[5].x), bf16_pair(v[6].x, v[7].x));
            *reinterpret_cast<uint4*>(d + SEG) = make_uint4(bf16_pair(v[0].y, v[1].y), bf16_pair(v[2].y, v[3].y),
                                                            bf16_pair(v[4].y, v[5].y), bf16_pair(v[6].y, v[7].y));
            *reinterpret_cast<uint4*>(d + 2 * SEG) = make_uint4(bf16_pair(v[0].z, v[1].z), bf16_pair(v[2].z, v[3].z),
                                                                bf16_pair(v[4].z, v[5].z), bf16_pair(v[6].z, v[7].z));
            *reinterpret_cast<uint4*>(d + 3 * SEG) = make_uint4(bf16_pair(v[0].w, v[1].w), bf16_pair(v[2].w, v[3].w),
                                                                bf16_pair(v[4].w, v[5].w), bf16_pair(v[6].w, v[7].w));
        }
        // input rows y - 1 .. y + 1, pixels xs - 1 + q: items of 4 channels x 8 pixels
        for (uint32_t e = tid; e < 3 * (ctn * 4) * (XQ / 8); e += 256) {
            const uint32_t cq = e % (ctn * 4), rest = e / (ctn * 4), qo = rest % (XQ / 8), r = rest / (XQ / 8);
            const int yy = (int)y + (int)r - 1;
            const uint32_t cc = ct0 * 16 + 4 * cq;
            float4 v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t q = 8 * qo + j;
                const int xx = (int)xs + (int)q - 1;
                v[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (q < SEG + 2 && yy >= 0 && yy < (int)h && xx >= 0 && xx < (int)w) {
                    if (cc < a.c0) {
                        const int sy = a.up ? yy >> 1 : yy, sx = a.up ? xx >> 1 : xx;
                        v[j] = *reinterpret_cast<const float4*>(
                            a.src0 + ((b * sh + (uint32_t)sy) * sw + (uint32_t)sx) * a.c0 + cc);
                    } else {
                        v[j] = *reinterpret_cast<const float4*>(
                            a.src1 + ((b * h + (uint32_t)yy) * w + (uint32_t)xx) * a.c1 + (cc - a.c0));
                    }
                }
            }
            uint16_t* d = Xt + ((size_t)r * xrows + 4 * cq) * XQ + 8 * qo;
            *reinterpret_cast<uint4*>(d) = make_uint4(bf16_pair(v[0].x, v[1].x), bf16_pair(v[2].x, v[3].x),
                                                      bf16_pair(v[4].x, v[5].x), bf16_pair(v[6].x, v[7].x));
            *reinterpret_cast<uint4*>(d + XQ) = make_uint4(bf16_pair(v[0].y, v[1].y), bf16_pair(v[2].y, v[3].y),
                                                           bf16_pair(v[4].y, v[5].y), bf16_pair(v[6].y, v[7].y));
            *reinterpret_cast<uint4*>(d + 2 * XQ) = make_uint4(bf16_pair(v[0].z, v[1].z), bf16_pair(v[2].z, v[3].z),
                                                               bf16_pair(v[4].z, v[5].z), bf16_pair(v[6].z, v[7].z));
            *reinterpret_cast<uint4*>(d + 3 * XQ) = make_uint4(bf16_pair(v[0].w, v[1].w), bf16_pair(v[2].w, v[3].w),
                                                               bf16_pair(v[4].w, v[5].w), bf16_pair(v[6].w, v[7].w));
        }
        __syncthreads();
        if (bias_block)
            for (uint32_t po = 0; po < SEG / 8; ++po) bsum += bpart[(size_t)po * np + tid];
        for (uint32_t p0 = 0; p0 < SEG && xs + p0 < w; p0 += 32) {
#pragma unroll
            for (int pi = 0; pi < 2; ++pi) {
                const uint32_t p = (uint32_t)wave + 4u * pi;
                if (p >= pairs) continue;
                const uint32_t ot = p % ot_n, ctl = p / ot_n;
                const v8bf av = as_v8bf(*reinterpret_cast<const uint4*>(Dt + (size_t)(ot * 16 + i) * SEG + p0 + 8 * g));
#pragma unroll
                for (int ky = 0; ky < 3; ++ky) {
                    const uint16_t* xr = Xt + ((size_t)ky * xrows + ctl * 16 + i) * XQ + p0 + 8 * g;
                    const uint4 w0 = *reinterpret_cast<const uint4*>(xr);
                    const uint32_t w4 = *reinterpret_cast<const uint32_t*>(xr + 8);
                    const uint4 w1 = make_uint4(__builtin_amdgcn_alignbyte(w0.y, w0.x, 2u),
                                                __builtin_amdgcn_alignbyte(w0.z, w0.y, 2u),
                                                __builtin_amdgcn_alignbyte(w0.w, w0.z, 2u),
                                                __builtin_amdgcn_alignbyte(w4, w0.w, 2u));
                    const uint4 w2 = make_uint4(w0.y, w0.z, w0.w, w4);
                    acc[pi][3 * ky + 0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, as_v8bf(w0), acc[pi][3 * ky + 0], 0, 0, 0);
                    acc[pi][3 * ky + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, as_v8bf(w1), acc[pi][3 * ky + 1], 0, 0, 0);
                    acc[pi][3 * ky + 2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, as_v8bf(w2), acc[pi][3 * ky + 2], 0, 0, 0);
                }
            }
        }
    }
    // lane: input column c = i, output rows 4 g + v
#pragma unroll
    for (int pi = 0; pi < 2; ++pi) {
        const uint32_t p = (uint32_t)wave + 4u * pi;
        if (p >= pairs) continue;
        const uint32_t ot = p % ot_n, ctl = p / ot_n;
        for (int tap = 0; tap < 9; ++tap)
            for (int v = 0; v < 4; ++v)
                a.part[(((size_t)blockIdx.y * 9 + tap) * np + ot * 16 + 4 * g + v) * cin + (ct0 + ctl) * 16 + i] =
                    acc[pi][tap][v];
    }
    if (bias_block) a.part_b[(size_t)blockIdx.y * np + tid] = bsum;
}

// dW[o][pc][tap] = sum over chunks: a block takes 16 consecutive partial
// entries, its 16 slices of threads the chunks g = slice, slice + 16, ... in
// order, and the slices' sums are added in slice order (a fixed order, so
// the sum is the same every run)
constexpr uint32_t SUM_SLICES = 16;
__global__ __launch_bounds__(256) void k_dw_sum(const float* part, uint32_t chunks, uint32_t np, uint32_t cin, float* dW) {
    __shared__ float red[SUM_SLICES][16];
    const uint32_t el = threadIdx.x & 15u, slice = threadIdx.x >> 4;
    const size_t n = (size_t)np * cin * 9;
    const size_t e = (size_t)blockIdx.x * 16 + el;
    float s = 0.0f;
    if (e < n)
        for (uint32_t g = slice; g < chunks; g += SUM_SLICES) s += part[(size_t)g * n + e];
    red[slice][el] = s;
    __syncthreads();
    if (slice != 0 || e >= n) return;
    float t = 0.0f;
    for (uint32_t k = 0; k < SUM_SLICES; ++k) t += red[k][el];
    const uint32_t pc = (uint32_t)(e % cin), rest = (uint32_t)(e / cin);
    const uint32_t o = rest % np, tap = rest / np;
    dW[((size_t)o * cin + pc) * 9 + tap] = t;
}
// one block per output channel
__global__ void k_db_sum(const float* part, uint32_t chunks, uint32_t np, float* db) {
    __shared__ float red[256];
    const uint32_t o = blockIdx.x;
    float s = 0.0f;
    for (uint32_t g = threadIdx.x; g < chunks; g += 256) s += part[(size_t)g * np + o];
    red[threadIdx.x] = s;
    for (int k = 128; k > 0; k >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    }
    if (threadIdx.x == 0) db[o] = red[0];
}
DwPlan dw_plan(uint32_t np, uint32_t cin, uint32_t batch, uint32_t h, uint32_t w) {
    DwPlan p;
    uint32_t ot_n = np / 16, ct_all = cin / 16;
    p.ct = std::max(1u, std::min(ct_all, 4u / std::max(ot_n, 1u)));
    p.groups = (ct_all + p.ct - 1) / p.ct;
    size_t segs = (size_t)batch * h * ((w + SEG - 1) / SEG);
    p.chunks = (uint32_t)std::max<size_t>(1, std::min<size_t>(segs, (1024 + p.groups - 1) / p.groups));
    p.lds = (size_t)(SEG * (np + 4) + 3 * (SEG + 2) * (p.ct * 16 + 4)) * 4;
    p.lds16 = ((size_t)np * SEG + 3 * (size_t)p.ct * 16 * XQ) * 2 + (size_t)(SEG / 8) * np * 4;
    // each of the 4 waves holds at most two (output, input) tile pairs
    if (ot_n * p.ct > 8 || p.lds > 64 * 1024 || p.lds16 > 64 * 1024) p.chunks = 0;
    return p;
}
void weight_grads(hipStream_t s, const DwArgs& a, const DwPlan& p, bool bf16, float* dW, float* db) {
    uint32_t cin = a.c0 + a.c1;
    if (p.chunks == 0) return;  // the host refuses such layers
    DwArgs b = a;
    b.ct = p.ct;
    dim3 grid(p.groups, p.chunks);
    if (bf16)
        hipLaunchKernelGGL(k_dw32, grid, dim3(256), p.lds16, s, b);
    else
        hipLaunchKernelGGL(k_dw, grid, dim3(256), p.lds, s, b);
    size_t n = (size_t)a.np * cin * 9;
    hipLaunchKernelGGL(k_dw_sum, dim3(blocks(n, 16)), dim3(256), 0, s, a.part, p.chunks, a.np, cin, dW);
    hipLaunchKernelGGL(k_db_sum, dim3(a.np), dim3(256), 0, s, a.part_b, p.chunks, a.np, db);
}

// ---- encoding, loss, decoding, batches ---------------------------------------------
// channels: log(1 + colour) x3, albedo x3, normal * 0.5 + 0.5 x3, log(1 + depth) / 8,
// log(1 + the standard error of the pixel's mean luminance);
// the network's frame is the image padded to multiples of 16 with zeros
__device__ __forceinline__ float finite_or0(float v) { return v == v && fabsf(v) < 1e30f ? v : 0.0f; }
// Frame b of a batch of `mirrors` holds the image mirrored by b (bit 0 in
// x, bit 1 in y) within its w x h pixels; the padding stays zero.
__global__ void k_learn_in(const float4* color, const float4* albedo, const float4* normal, const float* depth,
                           uint32_t depth_stride, const float* deviation, uint32_t w, uint32_t h, uint32_t w16,
                           uint32_t h16, uint32_t mirrors, float* x) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x, frame = (size_t)w16 * h16;
    if (e >= frame * mirrors) return;
    const uint32_t b = (uint32_t)(e / frame);
    const size_t q = e % frame;
    uint32_t px = q % w16, py = (uint32_t)(q / w16);
    float v[16] = {};
    if (px < w && py < h) {
        size_t p = (size_t)(b & 2u ? h - 1 - py : py) * w + (b & 1u ? w - 1 - px : px);
        float4 c = color[p], a = albedo[p], m = normal[p];
        float d = depth ? finite_or0(depth[p * depth_stride]) : 0.0f;
        v[0] = log1pf(fmaxf(finite_or0(c.x), 0.0f));
        v[1] = log1pf(fmaxf(finite_or0(c.y), 0.0f));
        v[2] = log1pf(fmaxf(finite_or0(c.z), 0.0f));
        v[3] = finite_or0(a.x);
        v[4] = finite_or0(a.y);
        v[5] = finite_or0(a.z);
        v[6] = finite_or0(m.x) * 0.5f + 0.5f;
        v[7] = finite_or0(m.y) * 0.5f + 0.5f;
        v[8] = finite_or0(m.z) * 0.5f + 0.5f;
        v[9] = log1pf(fmaxf(d, 0.0f)) * 0.125f;
        v[10] = deviation ? log1pf(fmaxf(finite_or0(deviation[p]), 0.0f)) : 0.0f;
    }
    float4* o = reinterpret_cast<float4*>(x + e * 16);
    for (int k = 0; k < 4; ++k) o[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
}
// The loss between the network's log radiance (channels 0..2 of y, 16 per
// pixel) and the target; writes the gradient of the mean and per-block
// sums. Mode 0: L1 between y and log(1 + target). Mode 1: the relative L2
// of linear radiance, (x - target)^2 / (x^2 + 0.01) with x = expm1(y) and
// the denominator held fixed (Lehtinen et al., Noise2Noise, 2018), whose
// minimiser over noisy targets is their mean; y is capped at 40 so x^2
// stays finite in FP32.
__global__ void k_loss(const float* y, const float4* target, size_t n, float scale, uint32_t mode, float* dy,
                       float* sums) {
    __shared__ float red[256];
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (p < n) {
        float4 t = target[p];
        float tl[3] = {fmaxf(t.x, 0.0f), fmaxf(t.y, 0.0f), fmaxf(t.z, 0.0f)};
        float g[16] = {};
        for (int k = 0; k < 3; ++k) {
            if (mode == 1u) {
                float v = fminf(y[p * 16 + k], 40.0f);
                float x = expm1f(v);
                float den = x * x + 0.01f;
                float dlt = x - tl[k];
                l += dlt * dlt / den;
                g[k] = y[p * 16 + k] > 40.0f ? 0.0f : scale * (2.0f * dlt / den) * (x + 1.0f);
            } else {
                float dlt = y[p * 16 + k] - log1pf(tl[k]);
                l += fabsf(dlt);
                g[k] = dlt > 0.0f ? scale : (dlt < 0.0f ? -scale : 0.0f);
            }
        }
        float4* o = reinterpret_cast<float4*>(dy + p * 16);
        for (int k = 0; k < 4; ++k) o[k] = make_float4(g[4 * k], g[4 * k + 1], g[4 * k + 2], g[4 * k + 3]);
    }
    red[threadIdx.x] = l;
    for (int k = 128; k > 0; k >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}
// Radiance of each pixel: the mean, over the `mirrors` frames of k_learn_in,
// of the network's output at the pixel's mirrored position, decoded
// (expm1) before averaging; frames added in order, so one frame is exact.
__global__ void k_learn_out(const float* y, uint32_t w16, uint32_t h16, uint32_t mirrors, const float4* color,
                            uint32_t w, uint32_t h, float4* out) {
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)w * h) return;
    uint32_t px = p % w, py = (uint32_t)(p / w);
    float r = 0.0f, g = 0.0f, bl = 0.0f;
    for (uint32_t b = 0; b < mirrors; ++b) {
        const size_t mx = b & 1u ? w - 1 - px : px, my = b & 2u ? h - 1 - py : py;
        const float* v = y + ((size_t)b * h16 * w16 + my * w16 + mx) * 16;
        r += expm1f(fmaxf(v[0], 0.0f));
        g += expm1f(fmaxf(v[1], 0.0f));
        bl += expm1f(fmaxf(v[2], 0.0f));
    }
    const float inv = 1.0f / (float)mirrors;
    out[p] = make_float4(r * inv, g * inv, bl * inv, color ? color[p].w : 1.0f);
}
// batch item b = crop picks[b] >> 3 under the square's symmetry picks[b] & 7
// (bit 0 mirrors x, bit 1 mirrors y, bit 2 transposes)
__global__ void k_gather(const float* inputs, const float4* targets, size_t crop_px, uint32_t side,
                         const uint32_t* picks, uint32_t batch, float* x, float4* t) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= crop_px * batch) return;
    uint32_t b = (uint32_t)(e / crop_px);
    uint32_t q = (uint32_t)(e % crop_px), ox = q % side, oy = q / side;
    uint32_t pick = picks[b], sym = pick & 7u;
    size_t crop = pick >> 3;
    uint32_t sx = sym & 1u ? side - 1 - ox : ox, sy = sym & 2u ? side - 1 - oy : oy;
    if (sym & 4u) {
        uint32_t k = sx;
        sx = sy;
        sy = k;
    }
    size_t src = crop * crop_px + (size_t)sy * side + sx;
    const float4* i4 = reinterpret_cast<const float4*>(inputs + src * 16);
    float4* o4 = reinterpret_cast<float4*>(x + e * 16);
    for (int k = 0; k < 4; ++k) o4[k] = i4[k];
    t[e] = targets[src];
}
void learn_in(hipStream_t s, const float4* c, const float4* a, const float4* nm, const float* d, uint32_t ds,
              const float* dev, uint32_t w, uint32_t h, uint32_t w16, uint32_t h16, uint32_t mirrors, float* x) {
    hipLaunchKernelGGL(k_learn_in, dim3(blocks((size_t)w16 * h16 * mirrors, 256)), dim3(256), 0, s, c, a, nm, d, ds,
                       dev, w, h, w16, h16, mirrors, x);
}
uint32_t loss(hipStream_t s, const float* y, const float4* t, size_t n, uint32_t mode, float* dy, float* sums) {
    uint32_t nb = blocks(n, 256);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, y, t, n, 1.0f / (3.0f * (float)n), mode, dy, sums);
    return nb;
}
void learn_out(hipStream_t s, const float* y, uint32_t w16, uint32_t h16, uint32_t mirrors, const float4* color,
               uint32_t w, uint32_t h, float4* out) {
    hipLaunchKernelGGL(k_learn_out, dim3(blocks((size_t)w * h, 256)), dim3(256), 0, s, y, w16, h16, mirrors, color, w,
                       h, out);
}
void gather(hipStream_t s, const float* inputs, const float4* targets, size_t crop_px, uint32_t side,
            const uint32_t* picks, uint32_t batch, float* x, float4* t) {
    hipLaunchKernelGGL(k_gather, dim3(blocks(crop_px * batch, 256)), dim3(256), 0, s, inputs, targets, crop_px, side,
                       picks, batch, x, t);
}

// the step's mean loss from the per-block sums of k_loss
__global__ void k_loss_total(const float* sums, uint32_t nb, float scale, float* out) {
    __shared__ float red[256];
    float l = 0.0f;
    for (uint32_t k = threadIdx.x; k < nb; k += 256) l += sums[k];
    red[threadIdx.x] = l;
    for (int k = 128; k > 0; k >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}
void loss_total(hipStream_t s, const float* sums, uint32_t nb, size_t n, float* out) {
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(256), 0, s, sums, nb, 1.0f / (3.0f * (float)n), out);
}

// ---- the per-tile check that hands tiles to the repo denoiser --------------------
// r[tile] = mean over the tile's pixels and colour channels of
// (log(1 + noisy) - log(1 + denoised))^2: the variance the network removed,
// an estimate of the tile's noise.
__global__ __launch_bounds__(256) void k_tile_residual(const float4* noisy, const float4* den, uint32_t w, uint32_t h,
                                                       uint32_t tile, uint32_t tiles_x, float* r) {
    __shared__ float red[256];
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const uint32_t x0 = tx * tile, y0 = ty * tile;
    float acc = 0.0f;
    for (uint32_t k = threadIdx.x; k < tile * tile; k += 256) {
        uint32_t x = x0 + k % tile, y = y0 + k / tile;
        if (x >= w || y >= h) continue;
        size_t p = (size_t)y * w + x;
        float4 a = noisy[p], b = den[p];
        float d0 = log1pf(fmaxf(finite_or0(a.x), 0.0f)) - log1pf(fmaxf(finite_or0(b.x), 0.0f));
        float d1 = log1pf(fmaxf(finite_or0(a.y), 0.0f)) - log1pf(fmaxf(finite_or0(b.y), 0.0f));
        float d2 = log1pf(fmaxf(finite_or0(a.z), 0.0f)) - log1pf(fmaxf(finite_or0(b.z), 0.0f));
        acc += d0 * d0 + d1 * d1 + d2 * d2;
    }
    red[threadIdx.x] = acc;
    for (int k = 128; k > 0; k >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    }
    if (threadIdx.x == 0) {
        uint32_t nx = min(tile, w - x0), ny = min(tile, h - y0);
        r[blockIdx.x] = red[0] / (3.0f * (float)(nx * ny));
    }
}
// out = repo where the pixel's tile is marked
__global__ void k_tile_merge(float4* out, const float4* repo, const uint8_t* mask, uint32_t w, uint32_t h,
                             uint32_t tile, uint32_t tiles_x) {
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)w * h) return;
    uint32_t x = p % w, y = (uint32_t)(p / w);
    if (mask[(y / tile) * tiles_x + x / tile]) out[p] = repo[p];
}
void tile_residual(hipStream_t s, const float4* noisy, const float4* den, uint32_t w, uint32_t h, uint32_t tile,
                   float* r) {
    uint32_t tx = (w + tile - 1) / tile, ty = (h + tile - 1) / tile;
    hipLaunchKernelGGL(k_tile_residual, dim3(tx * ty), dim3(256), 0, s, noisy, den, w, h, tile, tx, r);
}
void tile_merge(hipStream_t s, float4* out, const float4* repo, const uint8_t* mask, uint32_t w, uint32_t h,
                uint32_t tile) {
    uint32_t tx = (w + tile - 1) / tile;
    hipLaunchKernelGGL(k_tile_merge, dim3(blocks((size_t)w * h, 256)), dim3(256), 0, s, out, repo, mask, w, h, tile,
                       tx);
}

// ---- Adam ---------------------------------------------------------------------------------
__global__ void k_adam(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    float gr = g[e];
    float mm = b1 * m[e] + (1.0f - b1) * gr;
    float vv = b2 * v[e] + (1.0f - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}
void adam(hipStream_t s, float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
          float eps, int t) {
    float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

}  // namespace pfxl
