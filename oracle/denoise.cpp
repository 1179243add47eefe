// CPU restatement of the repo denoiser, the OIDN 2.5.1 RT filter port of
// crates/denoise -- TEST INFRASTRUCTURE ONLY (checker for the HIP denoiser).
//
//   tza_parse      crates/denoise/src/tza.rs:50-111
//   build_net      crates/denoise/src/net.rs:137-232 (layer graph)
//   plan           crates/denoise/src/tile.rs:31-101
//   exposure       crates/denoise/src/process.wgsl:104-157
//   input/output   process.wgsl:53-102 (PU transfer :33-51)
//   conv           crates/denoise/src/wgsl.rs:24-225: 3x3 "same" conv with
//                  zero padding at the tile edge, bias first, then for each
//                  4-channel input block, each tap (ky, kx) and each channel
//                  of the block one FMA; ReLU on every layer, 2x2 max pool
//                  after it, nearest upsampling and [upsampled, skip]
//                  concatenation for the decoder.
// All FP32 (glibc powf/logf/expf/exp2f/log2f for the transfer functions).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

namespace {

struct Tensor {
    std::vector<uint32_t> dims;
    std::vector<float> data;
};

float half_to_float(uint16_t bits) {  // tza.rs:113-128
    uint32_t sign = (uint32_t)(bits & 0x8000u) << 16;
    uint32_t e = (bits >> 10) & 0x1fu, m = bits & 0x3ffu, v;
    if (e == 0 && m == 0) {
        v = sign;
    } else if (e == 0) {
        uint32_t shift = (uint32_t)__builtin_clz(m) - 21u;
        v = sign | ((113u - shift) << 23) | (((m << shift) & 0x3ffu) << 13);
    } else if (e == 31) {
        v = sign | 0x7f800000u | (m ? 0x400000u | (m << 13) : 0u);
    } else {
        v = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &v, 4);
    return f;
}

bool tza_parse(const uint8_t* b, size_t len, std::map<std::string, Tensor>& out) {
    auto rd = [&](size_t at, size_t n, uint64_t* v) {
        if (at > len || n > len - at) return false;
        uint64_t x = 0;
        for (size_t i = 0; i < n; ++i) x |= (uint64_t)b[at + i] << (8 * i);
        *v = x;
        return true;
    };
    uint64_t magic, major, table, count;
    if (!rd(0, 2, &magic) || magic != 0x41D7 || !rd(2, 1, &major) || major != 2 || !rd(4, 8, &table)) return false;
    size_t at = table;
    if (!rd(at, 4, &count)) return false;
    at += 4;
    for (uint64_t t = 0; t < count; ++t) {
        uint64_t nl, rank, kind, off;
        if (!rd(at, 2, &nl) || at + 2 + nl > len) return false;
        std::string name((const char*)b + at + 2, nl);
        at += 2 + nl;
        if (!rd(at, 1, &rank)) return false;
        at += 1;
        Tensor ten;
        uint64_t elements = 1;
        for (uint64_t r = 0; r < rank; ++r) {
            uint64_t d;
            if (!rd(at, 4, &d) || d == 0 || d > (1u << 30)) return false;
            ten.dims.push_back((uint32_t)d);
            elements *= d;
            at += 4;
        }
        at += rank;  // layout string ("x" or "oihw")
        if (!rd(at, 1, &kind)) return false;
        at += 1;
        if (!rd(at, 8, &off)) return false;
        at += 8;
        size_t size = kind == 'f' ? 4 : (kind == 'h' ? 2 : 0);
        if (!size || off > len || elements * size > len - off) return false;
        ten.data.resize(elements);
        for (uint64_t i = 0; i < elements; ++i) {
            if (size == 4) {
                memcpy(&ten.data[i], b + off + 4 * i, 4);
            } else {
                uint16_t h;
                memcpy(&h, b + off + 2 * i, 2);
                ten.data[i] = half_to_float(h);
            }
        }
        out[name] = std::move(ten);
    }
    return true;
}

struct Layer {
    int src0, src1;  // -1 = network input, else a layer index; src1 -2 = none
    bool pool, upsample;
    uint32_t level, in0, in1, out;
    std::vector<float> w, bias;  // OIHW, as stored
};

uint32_t out_level(const std::vector<Layer>& ls, int src) { return src < 0 ? 0 : ls[src].level + (ls[src].pool ? 1 : 0); }
uint32_t channels(const std::vector<Layer>& ls, int src, uint32_t input) { return src < 0 ? input : ls[src].out; }

bool build_net(std::map<std::string, Tensor>& t, std::vector<Layer>& ls) {
    auto push = [&](const char* name, int s0, int s1, bool pool) -> int {
        auto wi = t.find(std::string(name) + ".weight"), bi = t.find(std::string(name) + ".bias");
        if (wi == t.end() || bi == t.end()) return -99;
        Layer l;
        l.src0 = s0;
        l.src1 = s1;
        l.pool = pool;
        l.upsample = s1 != -2;
        l.in0 = channels(ls, s0, 9);
        l.in1 = s1 != -2 ? channels(ls, s1, 9) : 0;
        l.level = s1 != -2 ? out_level(ls, s1) : out_level(ls, s0);
        l.out = wi->second.dims[0];
        if (wi->second.dims.size() != 4 || wi->second.dims[1] != l.in0 + l.in1 || bi->second.dims[0] != l.out)
            return -99;
        l.w = wi->second.data;
        l.bias = bi->second.data;
        ls.push_back(l);
        return (int)ls.size() - 1;
    };
    const int IN = -1, NONE = -2;
    int x = push("enc_conv0", IN, NONE, false);
    int p1 = push("enc_conv1", x, NONE, true);
    int p2 = push("enc_conv2", p1, NONE, true);
    int p3 = push("enc_conv3", p2, NONE, true);
    int p4 = push("enc_conv4", p3, NONE, true);
    x = push("enc_conv5a", p4, NONE, false);
    x = push("enc_conv5b", x, NONE, false);
    x = push("dec_conv4a", x, p3, false);
    x = push("dec_conv4b", x, NONE, false);
    x = push("dec_conv3a", x, p2, false);
    x = push("dec_conv3b", x, NONE, false);
    x = push("dec_conv2a", x, p1, false);
    x = push("dec_conv2b", x, NONE, false);
    x = push("dec_conv1a", x, IN, false);
    x = push("dec_conv1b", x, NONE, false);
    x = push("dec_conv0", x, NONE, false);
    return x >= 0 && ls.back().out == 3;
}

// CHW planes, channels padded to blocks of 4 (zeros), as the reference keeps them
struct Planes {
    uint32_t c, h, w;  // c: padded channel count
    std::vector<float> v;
    float at(uint32_t ch, int y, int x) const { return v[((size_t)ch * h + (uint32_t)y) * w + (uint32_t)x]; }
};

uint32_t pad4(uint32_t c) { return (c + 3u) / 4u * 4u; }

Planes conv(const Layer& l, const Planes& s0, const Planes* s1, uint32_t h, uint32_t w) {
    uint32_t in0p = pad4(l.in0), in1p = pad4(l.in1), inputs = l.in0 + l.in1;
    uint32_t ho = l.pool ? h / 2 : h, wo = l.pool ? w / 2 : w;
    Planes d{pad4(l.out), ho, wo, std::vector<float>((size_t)pad4(l.out) * ho * wo, 0.0f)};
    std::vector<float> full((size_t)h * w);
    for (uint32_t o = 0; o < l.out; ++o) {
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                float acc = l.bias[o];
                for (uint32_t pc = 0; pc < in0p + in1p; pc += 4)
                    for (int ky = 0; ky < 3; ++ky)
                        for (int kx = 0; kx < 3; ++kx)
                            for (uint32_t lane = 0; lane < 4; ++lane) {
                                uint32_t padded = pc + lane;
                                int ch;  // channel of the weight tensor, or -1 for padding
                                if (padded < in0p)
                                    ch = padded < l.in0 ? (int)padded : -1;
                                else
                                    ch = padded - in0p < l.in1 ? (int)(l.in0 + padded - in0p) : -1;
                                if (ch < 0) continue;  // zero weight times zero input
                                int yy = (int)y + ky - 1, xx = (int)x + kx - 1;
                                float v = 0.0f;
                                if (yy >= 0 && yy < (int)h && xx >= 0 && xx < (int)w) {
                                    if (padded < in0p)
                                        v = l.upsample ? s0.at(padded, yy >> 1, xx >> 1) : s0.at(padded, yy, xx);
                                    else
                                        v = s1->at(padded - in0p, yy, xx);
                                }
                                float wt = l.w[((size_t)o * inputs + (uint32_t)ch) * 9 + (uint32_t)(ky * 3 + kx)];
                                acc = fmaf(wt, v, acc);
                            }
                full[(size_t)y * w + x] = acc;
            }
        for (uint32_t y = 0; y < ho; ++y)
            for (uint32_t x = 0; x < wo; ++x) {
                float r;
                if (l.pool) {
                    float a = full[(size_t)(2 * y) * w + 2 * x], b = full[(size_t)(2 * y) * w + 2 * x + 1];
                    float c = full[(size_t)(2 * y + 1) * w + 2 * x], e = full[(size_t)(2 * y + 1) * w + 2 * x + 1];
                    r = std::max(std::max(std::max(a, b), std::max(c, e)), 0.0f);
                } else {
                    r = std::max(full[(size_t)y * w + x], 0.0f);
                }
                d.v[((size_t)o * ho + y) * wo + x] = r;
            }
    }
    return d;
}

// PU transfer (process.wgsl:33-51)
float pu_forward(float y) {
    if (y <= 1.57945760e-06f) return 1.41283765e+03f * y;
    if (y <= 3.22087631e-02f) return 1.64593172e+00f * powf(y, 4.31384981e-01f) + -2.94139609e-03f;
    return 1.92653254e-01f * logf(y + 6.26026094e-03f) + 9.98620152e-01f;
}
float pu_inverse(float x) {
    if (x <= 2.23151711e-03f) return x / 1.41283765e+03f;
    if (x <= 3.70974749e-01f) return powf((x - -2.94139609e-03f) / 1.64593172e+00f, 1.0f / 4.31384981e-01f);
    return expf((x - 9.98620152e-01f) / 1.92653254e-01f) - 6.26026094e-03f;
}
float nan_to_zero(float v) { return v != v ? 0.0f : v; }
float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }
const float FLT_MAX_ = 3.4028234663852886e38f;

// autoexposure (process.wgsl:104-157): 16x16-ish bins of mean luminance,
// then KEY / exp2(mean log2 of the bins above EPS)
float exposure(const float* color, uint32_t w, uint32_t h) {
    uint32_t nx = (w + 15) / 16, ny = (h + 15) / 16;
    double sum = 0.0;
    uint32_t count = 0;
    for (uint32_t gy = 0; gy < ny; ++gy)
        for (uint32_t gx = 0; gx < nx; ++gx) {
            uint32_t h0 = gy * h / ny, h1 = (gy + 1) * h / ny, w0 = gx * w / nx, w1 = (gx + 1) * w / nx;
            float s = 0.0f;
            for (uint32_t y = h0; y < h1; ++y)
                for (uint32_t x = w0; x < w1; ++x) {
                    const float* c = color + ((size_t)y * w + x) * 4;
                    float r = clampf(nan_to_zero(c[0]), 0.0f, FLT_MAX_), g = clampf(nan_to_zero(c[1]), 0.0f, FLT_MAX_),
                          b = clampf(nan_to_zero(c[2]), 0.0f, FLT_MAX_);
                    s += 0.212671f * r + 0.715160f * g + 0.072169f * b;
                }
            float l = s / (float)((h1 - h0) * (w1 - w0));
            if (l > 1e-8f) {
                sum += log2f(l);
                count += 1;
            }
        }
    return count ? 0.18f / exp2f((float)(sum / count)) : 1.0f;
}

struct Tile {
    uint32_t src[2], dst[2], size[2], out_src[2], out_dst[2], out_size[2];
};

uint32_t round_up(uint32_t a, uint32_t b) { return (a + b - 1) / b * b; }
uint32_t round_up_rem(uint32_t a, uint32_t b, uint32_t c) { return (a + b - c - 1) / b * b + c; }

// tile.rs:31-101 (RECEPTIVE_FIELD 174, MIN_TILE_ALIGNMENT 16)
void plan(uint32_t h, uint32_t w, uint64_t max_pixels, uint32_t* tile_h_out, uint32_t* tile_w_out,
          std::vector<Tile>& tiles) {
    const uint32_t align = 16, overlap = round_up(174 / 2, 16);
    uint32_t tile_h = round_up(h, 16), tile_w = round_up(w, 16);
    uint32_t pad_h = tile_h % align, pad_w = tile_w % align;
    uint32_t count_h = 1, count_w = 1;
    uint32_t min_dim = std::max(4 * overlap, 768u);
    uint32_t min_tile_h = round_up_rem(min_dim, align, pad_h), min_tile_w = round_up_rem(min_dim, align, pad_w);
    uint32_t span_h = 2 * overlap + pad_h, span_w = 2 * overlap + pad_w;
    auto sat = [](uint32_t a, uint32_t b) { return a > b ? a - b : 0u; };
    while ((uint64_t)tile_h * tile_w > max_pixels) {
        if (tile_h > min_tile_h && tile_h > tile_w) {
            uint32_t next = (h + span_h * count_h + count_h) / (count_h + 1);
            tile_h = std::min(std::max(round_up_rem(next, align, pad_h), min_tile_h), tile_h - align);
            count_h = std::max((sat(h, span_h) + (tile_h - span_h) - 1) / (tile_h - span_h), 1u);
        } else if (tile_w > min_tile_w) {
            uint32_t next = (w + span_w * count_w + count_w) / (count_w + 1);
            tile_w = std::min(std::max(round_up_rem(next, align, pad_w), min_tile_w), tile_w - align);
            count_w = std::max((sat(w, span_w) + (tile_w - span_w) - 1) / (tile_w - span_w), 1u);
        } else {
            break;
        }
    }
    tiles.clear();
    for (uint32_t i = 0; i < count_h; ++i) {
        uint32_t y = i * sat(tile_h, span_h);
        uint32_t bh = i > 0 ? overlap : 0, eh = i < count_h - 1 ? overlap + pad_h : 0;
        uint32_t th1 = std::min(h - y, tile_h), th2 = th1 - bh - eh, ah = tile_h - round_up(th1, 16);
        for (uint32_t j = 0; j < count_w; ++j) {
            uint32_t x = j * sat(tile_w, span_w);
            uint32_t bw = j > 0 ? overlap : 0, ew = j < count_w - 1 ? overlap + pad_w : 0;
            uint32_t tw1 = std::min(w - x, tile_w), tw2 = tw1 - bw - ew, aw = tile_w - round_up(tw1, 16);
            tiles.push_back(Tile{{y, x}, {ah, aw}, {th1, tw1}, {ah + bh, aw + bw}, {y + bh, x + bw}, {th2, tw2}});
        }
    }
    *tile_h_out = tile_h;
    *tile_w_out = tile_w;
}

}  // namespace

extern "C" {

// Number of tensors in a weights blob, or -1 when it does not parse.
int or_tza_count(const uint8_t* bytes, size_t len) {
    std::map<std::string, Tensor> t;
    return tza_parse(bytes, len, t) ? (int)t.size() : -1;
}

int or_denoise_plan(uint32_t h, uint32_t w, uint64_t max_pixels, uint32_t* tile_h, uint32_t* tile_w, uint32_t* rows,
                    uint32_t cap) {
    std::vector<Tile> tiles;
    plan(h, w, max_pixels, tile_h, tile_w, tiles);
    for (size_t i = 0; i < tiles.size() && i < cap; ++i) memcpy(rows + 12 * i, &tiles[i], sizeof(Tile));
    return (int)tiles.size();
}

// Denoise RGBA32F images (row-major, top row first) as the reference
// Denoiser::denoise does, with tiles of at most max_pixels pixels.
int or_denoise(const uint8_t* tza, size_t len, uint32_t w, uint32_t h, const float* color, const float* albedo,
               const float* normal, float* out, uint64_t max_pixels) {
    std::map<std::string, Tensor> t;
    std::vector<Layer> ls;
    if (!tza_parse(tza, len, t) || !build_net(t, ls)) return 1;
    float scale = exposure(color, w, h);
    float maxv = 1.92653254e-1f * logf(65504.0f + 6.260261e-3f) + 9.9862015e-1f;  // lib.rs:70-73
    float norm = (float)(1.0 / (double)maxv), rcp_norm = maxv;
    uint32_t tile_h, tile_w;
    std::vector<Tile> tiles;
    plan(h, w, max_pixels, &tile_h, &tile_w, tiles);
    for (const Tile& tile : tiles) {
        Planes in{12, tile_h, tile_w, std::vector<float>((size_t)12 * tile_h * tile_w, 0.0f)};
        for (uint32_t y = 0; y < tile_h; ++y)
            for (uint32_t x = 0; x < tile_w; ++x) {
                int sx = (int)x - (int)tile.dst[1], sy = (int)y - (int)tile.dst[0];
                float v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (sx >= 0 && sx < (int)tile.size[1] && sy >= 0 && sy < (int)tile.size[0]) {
                    size_t at = ((size_t)(sy + tile.src[0]) * w + (uint32_t)(sx + tile.src[1])) * 4;
                    for (int k = 0; k < 3; ++k) {
                        float c = clampf(nan_to_zero(color[at + k] * scale), 0.0f, FLT_MAX_);
                        v[k] = pu_forward(c) * norm;
                        v[3 + k] = clampf(nan_to_zero(albedo[at + k]), 0.0f, 1.0f);
                        v[6 + k] = clampf(nan_to_zero(normal[at + k]), -1.0f, 1.0f) * 0.5f + 0.5f;
                    }
                }
                for (int k = 0; k < 9; ++k) in.v[((size_t)k * tile_h + y) * tile_w + x] = v[k];
            }
        std::vector<Planes> outs;
        for (const Layer& l : ls) {
            uint32_t lh = tile_h >> l.level, lw = tile_w >> l.level;
            const Planes& s0 = l.src0 < 0 ? in : outs[l.src0];
            const Planes* s1 = l.src1 == -2 ? nullptr : (l.src1 < 0 ? &in : &outs[l.src1]);
            outs.push_back(conv(l, s0, s1, lh, lw));
        }
        const Planes& r = outs.back();
        for (uint32_t y = 0; y < tile.out_size[0]; ++y)
            for (uint32_t x = 0; x < tile.out_size[1]; ++x) {
                uint32_t sy = y + tile.out_src[0], sx = x + tile.out_src[1];
                size_t at = ((size_t)(y + tile.out_dst[0]) * w + x + tile.out_dst[1]) * 4;
                for (int k = 0; k < 3; ++k) {
                    float v = clampf(nan_to_zero(r.at((uint32_t)k, (int)sy, (int)sx)), 0.0f, FLT_MAX_) * rcp_norm;
                    out[at + k] = pu_inverse(v) * (scale != 0.0f ? 1.0f / scale : 0.0f);
                }
                out[at + 3] = color[at + 3];
            }
    }
    return 0;
}

}  // extern "C"
