// Host side of the repo denoiser (OIDN 2.5.1's RT filter as crates/denoise
// runs it): the TZA weights reader (crates/denoise/src/tza.rs:50-111), the
// U-Net layer graph (net.rs:137-232), OIDN's tiling (tile.rs:31-101), the
// per-size setup and the C ABI.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pfxhip.h"
#include "denoise.h"

using pfxd::ConvDev;
using pfxd::TileDev;

namespace {

struct Failure {
    pfxh_status status;
    std::string message;
};

void put_err(char* err, size_t n, const std::string& s) {
    if (!err || n == 0) return;
    size_t k = std::min(s.size(), n - 1);
    memcpy(err, s.data(), k);
    err[k] = 0;
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Failure{PFXH_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
    }
}

// ---- weights (tza.rs) -----------------------------------------------------------
struct Tensor {
    std::vector<uint32_t> dims;
    std::vector<float> data;
};

const char* CORRUPT = "invalid or corrupted weights blob";

float half_to_f32(uint16_t bits) {
    uint32_t sign = (uint32_t)(bits & 0x8000u) << 16;
    uint32_t e = (bits >> 10) & 0x1fu, m = bits & 0x3ffu, v;
    if (e == 0 && m == 0)
        v = sign;
    else if (e == 0) {
        uint32_t shift = (uint32_t)__builtin_clz(m) - 21u;
        v = sign | ((113u - shift) << 23) | (((m << shift) & 0x3ffu) << 13);
    } else if (e == 31)
        v = sign | 0x7f800000u | (m ? (0x400000u | (m << 13)) : 0u);
    else
        v = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    memcpy(&f, &v, 4);
    return f;
}

struct Bytes {
    const uint8_t* p;
    size_t len;
    uint64_t get(size_t at, size_t n) const {
        if (at > len || n > len - at) throw Failure{PFXH_SETUP, CORRUPT};
        uint64_t v = 0;
        for (size_t i = 0; i < n; ++i) v |= (uint64_t)p[at + i] << (8 * i);
        return v;
    }
};

std::map<std::string, Tensor> parse_tza(const uint8_t* p, size_t len) {
    Bytes b{p, len};
    if (b.get(0, 2) != 0x41D7) throw Failure{PFXH_SETUP, CORRUPT};
    if (b.get(2, 1) != 2) throw Failure{PFXH_SETUP, "unsupported weights blob version"};
    size_t at = (size_t)b.get(4, 8);
    uint64_t count = b.get(at, 4);
    at += 4;
    std::map<std::string, Tensor> out;
    for (uint64_t t = 0; t < count; ++t) {
        size_t nl = (size_t)b.get(at, 2);
        (void)b.get(at + 2, nl);
        std::string name((const char*)p + at + 2, nl);
        at += 2 + nl;
        size_t rank = (size_t)b.get(at, 1);
        at += 1;
        Tensor ten;
        uint64_t elements = 1;
        for (size_t r = 0; r < rank; ++r) {
            uint64_t d = b.get(at, 4);
            if (d == 0 || d > (1u << 30)) throw Failure{PFXH_SETUP, "invalid tensor dimension"};
            ten.dims.push_back((uint32_t)d);
            elements *= d;
            at += 4;
        }
        (void)b.get(at, rank);
        std::string layout((const char*)p + at, rank);
        if (layout != "x" && layout != "oihw") throw Failure{PFXH_SETUP, "invalid tensor layout"};
        at += rank;
        uint64_t kind = b.get(at, 1);
        size_t size = kind == 'f' ? 4 : (kind == 'h' ? 2 : 0);
        if (!size) throw Failure{PFXH_SETUP, "invalid tensor data type"};
        size_t off = (size_t)b.get(at + 1, 8);
        at += 9;
        if (elements > (uint64_t)len || off > len || elements * size > len - off) throw Failure{PFXH_SETUP, CORRUPT};
        ten.data.resize(elements);
        for (uint64_t i = 0; i < elements; ++i) {
            if (size == 4) {
                memcpy(&ten.data[i], p + off + 4 * i, 4);
            } else {
                uint16_t hb;
                memcpy(&hb, p + off + 2 * i, 2);
                ten.data[i] = half_to_f32(hb);
            }
        }
        out[name] = std::move(ten);
    }
    return out;
}

// ---- the network (net.rs) ---------------------------------------------------------
uint32_t pad16(uint32_t c) { return (c + 15u) / 16u * 16u; }

struct Layer {
    std::string name;
    int src0, src1;  // -1 input, else layer index; src1 -2 none
    bool pool, upsample;
    uint32_t level, in0, in1, out;
    std::vector<float> packed, bias;  // MFMA layout (ConvDev::w), padded bias
};

const int INPUT = -1, NONE = -2;

uint32_t out_level(const std::vector<Layer>& ls, int s) { return s < 0 ? 0 : ls[s].level + (ls[s].pool ? 1u : 0u); }
uint32_t chans(const std::vector<Layer>& ls, int s) { return s < 0 ? 9u : ls[s].out; }

void push(std::vector<Layer>& ls, std::map<std::string, Tensor>& t, const char* name, int s0, int s1, bool pool) {
    auto wi = t.find(std::string(name) + ".weight"), bi = t.find(std::string(name) + ".bias");
    if (wi == t.end() || bi == t.end()) throw Failure{PFXH_SETUP, CORRUPT};
    Layer l;
    l.name = name;
    l.src0 = s0;
    l.src1 = s1;
    l.pool = pool;
    l.upsample = s1 != NONE;
    l.in0 = chans(ls, s0);
    l.in1 = s1 != NONE ? chans(ls, s1) : 0;
    l.level = s1 != NONE ? out_level(ls, s1) : out_level(ls, s0);
    if (l.upsample && out_level(ls, s0) != l.level + 1)
        throw Failure{PFXH_SETUP, std::string(name) + ": the upsampled source is not one level down"};
    const Tensor& w = wi->second;
    const Tensor& b = bi->second;
    if (w.dims.size() != 4 || b.dims.size() != 1) throw Failure{PFXH_SETUP, "invalid convolution weight/bias"};
    l.out = w.dims[0];
    if (w.dims[1] != l.in0 + l.in1 || w.dims[2] != 3 || w.dims[3] != 3 || b.dims[0] != l.out)
        throw Failure{PFXH_SETUP, std::string(name) + ": the weights do not match the network"};
    if (pad16(l.out) > 112) throw Failure{PFXH_SETUP, std::string(name) + ": more output channels than supported"};
    // packed[chunk][tap][n][16]: input channel `chunk * 16 + q` of the padded
    // [src0 | src1] stack for output channel n
    uint32_t c0 = pad16(l.in0), c1 = pad16(l.in1), np = pad16(l.out), inputs = l.in0 + l.in1;
    uint32_t chunks = (c0 + c1) / 16u;
    l.packed.assign((size_t)chunks * 9 * np * 16, 0.0f);
    for (uint32_t ch = 0; ch < chunks; ++ch)
        for (uint32_t tap = 0; tap < 9; ++tap)
            for (uint32_t n = 0; n < l.out; ++n)
                for (uint32_t q = 0; q < 16; ++q) {
                    uint32_t pc = ch * 16 + q;
                    int src;
                    if (pc < c0)
                        src = pc < l.in0 ? (int)pc : -1;
                    else
                        src = pc - c0 < l.in1 ? (int)(l.in0 + pc - c0) : -1;
                    if (src < 0) continue;
                    l.packed[(((size_t)ch * 9 + tap) * np + n) * 16 + q] = w.data[((size_t)n * inputs + src) * 9 + tap];
                }
    l.bias.assign(np, 0.0f);
    std::copy(b.data.begin(), b.data.end(), l.bias.begin());
    ls.push_back(std::move(l));
}

std::vector<Layer> build_net(std::map<std::string, Tensor>& t) {
    std::vector<Layer> ls;
    auto last = [&] { return (int)ls.size() - 1; };
    push(ls, t, "enc_conv0", INPUT, NONE, false);
    push(ls, t, "enc_conv1", last(), NONE, true);
    int p1 = last();
    push(ls, t, "enc_conv2", p1, NONE, true);
    int p2 = last();
    push(ls, t, "enc_conv3", p2, NONE, true);
    int p3 = last();
    push(ls, t, "enc_conv4", p3, NONE, true);
    push(ls, t, "enc_conv5a", last(), NONE, false);
    push(ls, t, "enc_conv5b", last(), NONE, false);
    push(ls, t, "dec_conv4a", last(), p3, false);
    push(ls, t, "dec_conv4b", last(), NONE, false);
    push(ls, t, "dec_conv3a", last(), p2, false);
    push(ls, t, "dec_conv3b", last(), NONE, false);
    push(ls, t, "dec_conv2a", last(), p1, false);
    push(ls, t, "dec_conv2b", last(), NONE, false);
    push(ls, t, "dec_conv1a", last(), INPUT, false);
    push(ls, t, "dec_conv1b", last(), NONE, false);
    push(ls, t, "dec_conv0", last(), NONE, false);
    if (ls.back().out != 3) throw Failure{PFXH_SETUP, "the network does not end in three channels"};
    return ls;
}

// ---- tiling (tile.rs:31-101) ------------------------------------------------------
uint32_t round_up(uint32_t a, uint32_t b) { return (a + b - 1) / b * b; }
uint32_t round_up_rem(uint32_t a, uint32_t b, uint32_t c) { return (a + b - c - 1) / b * b + c; }
uint32_t sat(uint32_t a, uint32_t b) { return a > b ? a - b : 0u; }

void plan(uint32_t h, uint32_t w, uint64_t max_pixels, uint32_t* th_out, uint32_t* tw_out,
          std::vector<pfxh_denoise_tile>& tiles) {
    const uint32_t align = 16, overlap = round_up(174 / 2, 16);  // RECEPTIVE_FIELD 174
    uint32_t tile_h = round_up(h, 16), tile_w = round_up(w, 16);
    uint32_t pad_h = tile_h % align, pad_w = tile_w % align, count_h = 1, count_w = 1;
    uint32_t min_dim = std::max(4 * overlap, 768u);
    uint32_t min_th = round_up_rem(min_dim, align, pad_h), min_tw = round_up_rem(min_dim, align, pad_w);
    uint32_t span_h = 2 * overlap + pad_h, span_w = 2 * overlap + pad_w;
    while (max_pixels && (uint64_t)tile_h * tile_w > max_pixels) {
        if (tile_h > min_th && tile_h > tile_w) {
            uint32_t next = (h + span_h * count_h + count_h) / (count_h + 1);
            tile_h = std::min(std::max(round_up_rem(next, align, pad_h), min_th), tile_h - align);
            count_h = std::max((sat(h, span_h) + (tile_h - span_h) - 1) / (tile_h - span_h), 1u);
        } else if (tile_w > min_tw) {
            uint32_t next = (w + span_w * count_w + count_w) / (count_w + 1);
            tile_w = std::min(std::max(round_up_rem(next, align, pad_w), min_tw), tile_w - align);
            count_w = std::max((sat(w, span_w) + (tile_w - span_w) - 1) / (tile_w - span_w), 1u);
        } else {
            break;
        }
    }
    tiles.clear();
    for (uint32_t i = 0; i < count_h; ++i) {
        uint32_t y = i * sat(tile_h, span_h);
        uint32_t bh = i > 0 ? overlap : 0, eh = i < count_h - 1 ? overlap + pad_h : 0;
        uint32_t th1 = std::min(h - y, tile_h), ah = tile_h - round_up(th1, 16);
        for (uint32_t j = 0; j < count_w; ++j) {
            uint32_t x = j * sat(tile_w, span_w);
            uint32_t bw = j > 0 ? overlap : 0, ew = j < count_w - 1 ? overlap + pad_w : 0;
            uint32_t tw1 = std::min(w - x, tile_w), aw = tile_w - round_up(tw1, 16);
            tiles.push_back(pfxh_denoise_tile{{y, x}, {ah, aw}, {th1, tw1}, {ah + bh, aw + bw}, {y + bh, x + bw},
                                              {th1 - bh - eh, tw1 - bw - ew}});
        }
    }
    *th_out = tile_h;
    *tw_out = tile_w;
}

}  // namespace

struct pfxh_denoiser {
    std::vector<Layer> layers;
    std::vector<float*> weights, biases;  // device, per layer
    hipStream_t stream = nullptr;
    uint64_t max_tile_pixels = 0;  // 0: one tile
    // per-size setup
    uint32_t width = 0, height = 0, tile_h = 0, tile_w = 0;
    std::vector<pfxh_denoise_tile> tiles;
    std::vector<float*> tensors;  // input, then one per layer
    float4* images = nullptr;     // color, albedo, normal, output
    float* bins = nullptr;
    float* scale = nullptr;
    std::vector<void*> setup_mem;

    void free_setup() {
        for (void* p : setup_mem) (void)hipFree(p);
        setup_mem.clear();
        tensors.clear();
        images = nullptr;
        bins = scale = nullptr;
        width = height = 0;
    }
    void* alloc(size_t bytes) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Failure{PFXH_REFUSED, "the denoiser's buffers need " + std::to_string(bytes) + " bytes on the GPU: " +
                                            hipGetErrorString(e)};
        }
        setup_mem.push_back(p);
        return p;
    }
    void prepare(uint32_t w, uint32_t h) {
        if (w == width && h == height) return;
        free_setup();
        plan(h, w, max_tile_pixels, &tile_h, &tile_w, tiles);
        tensors.push_back((float*)alloc((size_t)tile_h * tile_w * 16 * 4));
        for (const Layer& l : layers) {
            uint32_t lv = l.level + (l.pool ? 1u : 0u);
            size_t px = (size_t)(tile_h >> lv) * (tile_w >> lv);
            tensors.push_back((float*)alloc(px * pad16(l.out) * 4));
        }
        images = (float4*)alloc((size_t)w * h * 16 * 4);
        bins = (float*)alloc((size_t)((w + 15) / 16) * ((h + 15) / 16) * 4);
        scale = (float*)alloc(16);
        width = w;
        height = h;
    }
    // color/albedo/normal/out are device images of width x height
    void run(const float4* color, const float4* albedo, const float4* normal, float4* out) {
        float maxv = 1.92653254e-1f * logf(65504.0f + 6.260261e-3f) + 9.9862015e-1f;  // lib.rs:70-73
        float norm = (float)(1.0 / (double)maxv), rcp_norm = maxv;
        pfxd::launch_exposure(stream, color, width, height, bins, scale);
        for (const pfxh_denoise_tile& pt : tiles) {
            TileDev t;
            for (int k = 0; k < 2; ++k) {
                t.src[k] = (int32_t)pt.src[k];
                t.dst[k] = (int32_t)pt.dst[k];
                t.size[k] = (int32_t)pt.size[k];
                t.out_src[k] = (int32_t)pt.out_src[k];
                t.out_dst[k] = (int32_t)pt.out_dst[k];
                t.out_size[k] = (int32_t)pt.out_size[k];
            }
            pfxd::launch_input(stream, color, albedo, normal, scale, width, t, tile_w, tile_h, norm, tensors[0]);
            for (size_t i = 0; i < layers.size(); ++i) {
                const Layer& l = layers[i];
                ConvDev c;
                c.src0 = tensors[(size_t)(l.src0 + 1)];
                c.src1 = l.src1 == NONE ? nullptr : tensors[(size_t)(l.src1 + 1)];
                c.dst = tensors[i + 1];
                c.w = weights[i];
                c.bias = biases[i];
                c.c0 = pad16(l.in0);
                c.c1 = pad16(l.in1);
                c.n_pad = pad16(l.out);
                c.h = tile_h >> l.level;
                c.w_ = tile_w >> l.level;
                c.upsample = l.upsample ? 1u : 0u;
                c.pool = l.pool ? 1u : 0u;
                c.relu = 1u;  // OIDN's RT filter applies ReLU after every layer
                c.bf16 = 0u;
                c.batch = 1u;
                pfxd::launch_conv(stream, c);
            }
            pfxd::launch_output(stream, tensors.back(), tile_w, scale, color, width, t, rcp_norm, out);
        }
        hip_check(hipGetLastError(), "denoise launch");
    }
    ~pfxh_denoiser() {
        free_setup();
        for (float* p : weights) (void)hipFree(p);
        for (float* p : biases) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
template <class F>
pfxh_status guard(char* err, size_t err_len, F&& f) {
    try {
        f();
        return PFXH_OK;
    } catch (const Failure& e) {
        put_err(err, err_len, e.message);
        return e.status;
    } catch (const std::bad_alloc&) {
        put_err(err, err_len, "host memory is exhausted");
        return PFXH_REFUSED;
    }
}
}  // namespace

// Internal: denoise device images (width x height RGBA32F) written on
// `after`, and copy the result to the host buffer `out`; with `out` null the
// result stays on the device at *device_out (valid until the next call).
pfxh_status pfxh_denoise_device(pfxh_denoiser* d, uint32_t width, uint32_t height, const void* color,
                                const void* albedo, const void* normal, float* out, hipStream_t after, char* err,
                                size_t err_len, const void** device_out) {
    return guard(err, err_len, [&] {
        if (width == 0 || height == 0) throw Failure{PFXH_INVALID, "image dimensions must be nonzero"};
        d->prepare(width, height);
        if (after) hip_check(hipStreamSynchronize(after), "denoise input");
        float4* o = d->images + 3 * (size_t)width * height;
        d->run((const float4*)color, (const float4*)albedo, (const float4*)normal, o);
        if (device_out) *device_out = o;
        if (out)
            hip_check(hipMemcpyAsync(out, o, (size_t)width * height * 16, hipMemcpyDeviceToHost, d->stream),
                      "denoise readback");
        hip_check(hipStreamSynchronize(d->stream), "denoise");
    });
}

extern "C" {

pfxh_status pfxh_denoiser_new(const uint8_t* weights, size_t len, pfxh_denoiser** out, char* err, size_t err_len) {
    if (!weights || !out) return PFXH_INVALID;
    *out = nullptr;
    pfxh_denoiser* d = nullptr;
    pfxh_status st = guard(err, err_len, [&] {
        std::map<std::string, Tensor> t = parse_tza(weights, len);
        std::vector<Layer> ls = build_net(t);
        d = new pfxh_denoiser();
        d->layers = std::move(ls);
        hip_check(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking), "denoise stream");
        for (const Layer& l : d->layers) {
            float* w = nullptr;
            float* b = nullptr;
            hip_check(hipMalloc(&w, l.packed.size() * 4), "denoise weights");
            d->weights.push_back(w);
            hip_check(hipMalloc(&b, l.bias.size() * 4), "denoise biases");
            d->biases.push_back(b);
            hip_check(hipMemcpy(w, l.packed.data(), l.packed.size() * 4, hipMemcpyHostToDevice), "denoise weights");
            hip_check(hipMemcpy(b, l.bias.data(), l.bias.size() * 4, hipMemcpyHostToDevice), "denoise biases");
        }
    });
    if (st != PFXH_OK) {
        delete d;
        return st;
    }
    *out = d;
    return PFXH_OK;
}

pfxh_status pfxh_denoise(pfxh_denoiser* d, uint32_t width, uint32_t height, const float* color, const float* albedo,
                         const float* normal, float* out, char* err, size_t err_len) {
    if (!d || !color || !albedo || !normal || !out) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (width == 0 || height == 0) throw Failure{PFXH_INVALID, "image dimensions must be nonzero"};
        d->prepare(width, height);
        size_t bytes = (size_t)width * height * 16;
        float4* im = d->images;
        size_t px = (size_t)width * height;
        hip_check(hipMemcpyAsync(im, color, bytes, hipMemcpyHostToDevice, d->stream), "denoise upload");
        hip_check(hipMemcpyAsync(im + px, albedo, bytes, hipMemcpyHostToDevice, d->stream), "denoise upload");
        hip_check(hipMemcpyAsync(im + 2 * px, normal, bytes, hipMemcpyHostToDevice, d->stream), "denoise upload");
        d->run(im, im + px, im + 2 * px, im + 3 * px);
        hip_check(hipMemcpyAsync(out, im + 3 * px, bytes, hipMemcpyDeviceToHost, d->stream), "denoise readback");
        hip_check(hipStreamSynchronize(d->stream), "denoise");
    });
}

void pfxh_denoiser_set_max_tile(pfxh_denoiser* d, uint64_t pixels) {
    if (!d) return;
    d->max_tile_pixels = pixels;
    d->free_setup();
}

pfxh_status pfxh_denoiser_layers(const uint8_t* weights, size_t len, uint32_t* layers, uint64_t* params, char* err,
                                 size_t err_len) {
    if (!weights || !layers || !params) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        std::map<std::string, Tensor> t = parse_tza(weights, len);
        std::vector<Layer> ls = build_net(t);
        uint64_t n = 0;
        for (auto& kv : t) n += kv.second.data.size();
        *layers = (uint32_t)ls.size();
        *params = n;
    });
}

pfxh_status pfxh_denoise_plan(uint32_t height, uint32_t width, uint64_t max_tile_pixels, uint32_t* tile_h,
                              uint32_t* tile_w, pfxh_denoise_tile* tiles, uint32_t cap, uint32_t* count) {
    if (!tile_h || !tile_w || !count || height == 0 || width == 0) return PFXH_INVALID;
    std::vector<pfxh_denoise_tile> ts;
    plan(height, width, max_tile_pixels, tile_h, tile_w, ts);
    *count = (uint32_t)ts.size();
    if (tiles)
        for (uint32_t i = 0; i < *count && i < cap; ++i) tiles[i] = ts[i];
    return PFXH_OK;
}

void pfxh_denoiser_free(pfxh_denoiser* d) { delete d; }

}  // extern "C"
