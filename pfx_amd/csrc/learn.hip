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
//   k_dw                    weight gradient: for each tap a GEMM over pixels
//                           on v_mfma_f32_16x16x4f32, partial sums per pixel
//                           chunk, then k_dw_sum
//   k_db                    bias gradient
//   k_learn_in / k_learn_out / k_loss   input encoding, decoding, L1 loss on
//                           log radiance and its gradient
//   k_adam                  Adam (Kingma & Ba 2015) with bias correction
#include <hip/hip_runtime.h>

#include "learn.h"

namespace pfxl {

typedef float v4f __attribute__((ext_vector_type(4)));

static inline uint32_t blocks(size_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

// ---- weight layouts ---------------------------------------------------------------
// master M[o][pc][tap], o < np (padded outputs), pc < cin (padded stacked inputs)
__global__ void k_pack_fwd(const float* M, uint32_t np, uint32_t cin, float* W) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)np * cin * 9;
    if (e >= total) return;
    uint32_t q = e % 16, rest = (uint32_t)(e / 16);
    uint32_t n = rest % np, rest2 = rest / np;
    uint32_t tap = rest2 % 9, chunk = rest2 / 9;
    uint32_t pc = chunk * 16 + q;
    W[e] = M[((size_t)n * cin + pc) * 9 + tap];
}
// transposed: inputs are the np output channels, outputs the cin channels
__global__ void k_pack_t(const float* M, uint32_t np, uint32_t cin, float* W) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)np * cin * 9;
    if (e >= total) return;
    uint32_t q = e % 16, rest = (uint32_t)(e / 16);
    uint32_t pc = rest % cin, rest2 = rest / cin;
    uint32_t tap = rest2 % 9, chunk = rest2 / 9;
    uint32_t o = chunk * 16 + q;
    W[e] = M[((size_t)o * cin + pc) * 9 + (8 - tap)];
}
void pack(hipStream_t s, const float* M, uint32_t np, uint32_t cin, float* fwd, float* t) {
    size_t n = (size_t)np * cin * 9;
    hipLaunchKernelGGL(k_pack_fwd, dim3(blocks(n, 256)), dim3(256), 0, s, M, np, cin, fwd);
    hipLaunchKernelGGL(k_pack_t, dim3(blocks(n, 256)), dim3(256), 0, s, M, np, cin, t);
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
// One wave per (16 outputs, 16 inputs, pixel chunk). A segment is 16 pixels
// of one row: dPre [16 px][16 o] and the input rows y-1..y+1 over 18 px [16 c]
// go to LDS, then for each tap four MFMAs add the segment's 16 pixels:
// A[o][k] = dPre[4g + k][o], B[k][c] = X[ky][4g + k + kx][c].
__global__ __launch_bounds__(64) void k_dw(DwArgs a) {
    __shared__ float D[16][17];
    __shared__ float X[3][18][17];
    const int lane = threadIdx.x, i = lane & 15, kk = lane >> 4;
    const uint32_t otiles = a.np / 16;
    const uint32_t ot = blockIdx.x % otiles, ct = blockIdx.x / otiles;
    const uint32_t o0 = ot * 16, cc0 = ct * 16;
    const bool from1 = cc0 >= a.c0;
    const uint32_t h = a.h, w = a.w, segw = (w + 15) / 16;
    const size_t segs = (size_t)a.batch * h * segw;
    const size_t s0 = segs * blockIdx.y / gridDim.y, s1 = segs * (blockIdx.y + 1) / gridDim.y;
    v4f acc[9];
    for (int t = 0; t < 9; ++t) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t sh = a.up ? h / 2 : h, sw = a.up ? w / 2 : w;
    for (size_t sg = s0; sg < s1; ++sg) {
        const uint32_t xs = (uint32_t)(sg % segw) * 16;
        const uint32_t y = (uint32_t)((sg / segw) % h);
        const size_t b = sg / ((size_t)segw * h);
        __syncthreads();
        for (int e = lane; e < 256; e += 64) {
            int px = e >> 4, o = e & 15;
            uint32_t x = xs + (uint32_t)px;
            D[px][o] = x < w ? a.dpre[((b * h + y) * w + x) * a.np + o0 + o] : 0.0f;
        }
        for (int e = lane; e < 3 * 18 * 16; e += 64) {
            int cc = e & 15, rest = e >> 4, px = rest % 18, r = rest / 18;
            int yy = (int)y + r - 1, xx = (int)xs + px - 1;
            float v = 0.0f;
            if (yy >= 0 && yy < (int)h && xx >= 0 && xx < (int)w) {
                if (!from1) {
                    int sy = a.up ? yy >> 1 : yy, sx = a.up ? xx >> 1 : xx;
                    v = a.src0[((b * sh + (uint32_t)sy) * sw + (uint32_t)sx) * a.c0 + cc0 + cc];
                } else {
                    v = a.src1[((b * h + (uint32_t)yy) * w + (uint32_t)xx) * a.c1 + (cc0 - a.c0) + cc];
                }
            }
            X[r][px][cc] = v;
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap - ky * 3;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float av = D[4 * g + kk][i];
                const float bv = X[ky][4 * g + kk + kx][i];
                acc[tap] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[tap], 0, 0, 0);
            }
        }
    }
    // lane: column c = i, rows o = 4 kk + v
    const uint32_t cin = a.c0 + a.c1;
    for (int tap = 0; tap < 9; ++tap)
        for (int v = 0; v < 4; ++v)
            a.part[(((size_t)blockIdx.y * 9 + tap) * a.np + o0 + 4 * kk + v) * cin + cc0 + i] = acc[tap][v];
}
// dW[o][pc][tap] = sum over chunks
__global__ void k_dw_sum(const float* part, uint32_t chunks, uint32_t np, uint32_t cin, float* dW) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t n = (size_t)np * cin * 9;
    if (e >= n) return;
    uint32_t tap = e % 9, rest = (uint32_t)(e / 9);
    uint32_t pc = rest % cin, o = rest / cin;
    float s = 0.0f;
    for (uint32_t g = 0; g < chunks; ++g) s += part[(((size_t)g * 9 + tap) * np + o) * cin + pc];
    dW[e] = s;
}
// bias: each block sums a pixel chunk for every output channel
__global__ void k_db(const float* dpre, size_t pixels, uint32_t np, uint32_t chunks, float* part) {
    uint32_t o = threadIdx.x;
    if (o >= np) return;
    size_t p0 = pixels * blockIdx.x / chunks, p1 = pixels * (blockIdx.x + 1) / chunks;
    float s = 0.0f;
    for (size_t p = p0; p < p1; ++p) s += dpre[p * np + o];
    part[(size_t)blockIdx.x * np + o] = s;
}
__global__ void k_db_sum(const float* part, uint32_t chunks, uint32_t np, float* db) {
    uint32_t o = threadIdx.x;
    if (o >= np) return;
    float s = 0.0f;
    for (uint32_t g = 0; g < chunks; ++g) s += part[(size_t)g * np + o];
    db[o] = s;
}
void weight_grads(hipStream_t s, const DwArgs& a, uint32_t chunks, float* dW, float* db) {
    uint32_t cin = a.c0 + a.c1;
    hipLaunchKernelGGL(k_dw, dim3((a.np / 16) * (cin / 16), chunks), dim3(64), 0, s, a);
    size_t n = (size_t)a.np * cin * 9;
    hipLaunchKernelGGL(k_dw_sum, dim3(blocks(n, 256)), dim3(256), 0, s, a.part, chunks, a.np, cin, dW);
    size_t pixels = (size_t)a.batch * a.h * a.w;
    hipLaunchKernelGGL(k_db, dim3(chunks), dim3(128), 0, s, a.dpre, pixels, a.np, chunks, a.part);
    hipLaunchKernelGGL(k_db_sum, dim3(1), dim3(128), 0, s, a.part, chunks, a.np, db);
}

// ---- encoding, loss, decoding, batches ---------------------------------------------
// channels: log(1 + colour) x3, albedo x3, normal * 0.5 + 0.5 x3, log(1 + depth) / 8;
// the network's frame is the image padded to multiples of 16 with zeros
__device__ __forceinline__ float finite_or0(float v) { return v == v && fabsf(v) < 1e30f ? v : 0.0f; }
__global__ void k_learn_in(const float4* color, const float4* albedo, const float4* normal, const float* depth,
                           uint32_t depth_stride, uint32_t w, uint32_t h, uint32_t w16, uint32_t h16, float* x) {
    size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= (size_t)w16 * h16) return;
    uint32_t px = q % w16, py = (uint32_t)(q / w16);
    float v[16] = {};
    if (px < w && py < h) {
        size_t p = (size_t)py * w + px;
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
    }
    float4* o = reinterpret_cast<float4*>(x + q * 16);
    for (int k = 0; k < 4; ++k) o[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
}
// L1 between the network's log radiance (channels 0..2 of y, 16 per pixel)
// and log(1 + target); writes the gradient of the mean and per-block sums
__global__ void k_loss(const float* y, const float4* target, size_t n, float scale, float* dy, float* sums) {
    __shared__ float red[256];
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (p < n) {
        float4 t = target[p];
        float tv[3] = {log1pf(fmaxf(t.x, 0.0f)), log1pf(fmaxf(t.y, 0.0f)), log1pf(fmaxf(t.z, 0.0f))};
        float g[16] = {};
        for (int k = 0; k < 3; ++k) {
            float dlt = y[p * 16 + k] - tv[k];
            l += fabsf(dlt);
            g[k] = dlt > 0.0f ? scale : (dlt < 0.0f ? -scale : 0.0f);
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
__global__ void k_learn_out(const float* y, uint32_t w16, const float4* color, uint32_t w, uint32_t h, float4* out) {
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)w * h) return;
    uint32_t px = p % w, py = (uint32_t)(p / w);
    const float* v = y + ((size_t)py * w16 + px) * 16;
    out[p] = make_float4(expm1f(fmaxf(v[0], 0.0f)), expm1f(fmaxf(v[1], 0.0f)), expm1f(fmaxf(v[2], 0.0f)),
                         color ? color[p].w : 1.0f);
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
              uint32_t w, uint32_t h, uint32_t w16, uint32_t h16, float* x) {
    hipLaunchKernelGGL(k_learn_in, dim3(blocks((size_t)w16 * h16, 256)), dim3(256), 0, s, c, a, nm, d, ds, w, h, w16,
                       h16, x);
}
uint32_t loss(hipStream_t s, const float* y, const float4* t, size_t n, float* dy, float* sums) {
    uint32_t nb = blocks(n, 256);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, y, t, n, 1.0f / (3.0f * (float)n), dy, sums);
    return nb;
}
void learn_out(hipStream_t s, const float* y, uint32_t w16, const float4* color, uint32_t w, uint32_t h, float4* out) {
    hipLaunchKernelGGL(k_learn_out, dim3(blocks((size_t)w * h, 256)), dim3(256), 0, s, y, w16, color, w, h, out);
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
