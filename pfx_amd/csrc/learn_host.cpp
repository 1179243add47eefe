// Host side of the learned denoiser: the network table, parameter layouts,
// the device-resident training set, the training step (forward, L1 loss on
// log radiance, backward, Adam) and inference, behind the C ABI.
//
// The network is a 5-level encoder-decoder of 16 3x3 convolutions with skip
// connections, shaped like the U-Net of OIDN's RT filter that the repo
// denoiser runs (crates/denoise/src/net.rs:202-232) but 32->128 channels
// wide: 1,239,651 parameters. The reference has no trained denoiser; this is
// the part its still path lacks.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "../../include/pfxhip.h"
#include "denoise.h"
#include "learn.h"

using pfxd::ConvDev;

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

uint32_t pad16(uint32_t c) { return (c + 15u) / 16u * 16u; }

// ---- the network --------------------------------------------------------------------
// Tensor 0 is the encoded input (10 channels); layer k writes tensor k + 1.
// `level` is the convolution's resolution (full >> level); a pooled layer's
// output sits one level down, an upsampled input one level up.
struct Def {
    const char* name;
    int in0, in1;  // tensor ids; in1 < 0: no skip input
    uint32_t out, level;
    bool up, pool, relu;
};
constexpr uint32_t INPUT_CHANNELS = 11;
constexpr Def NET[PFXH_LEARN_LAYERS] = {
    {"enc_conv0", 0, -1, 32, 0, false, false, true},    {"enc_conv1", 1, -1, 32, 0, false, true, true},
    {"enc_conv2", 2, -1, 64, 1, false, true, true},     {"enc_conv3", 3, -1, 96, 2, false, true, true},
    {"enc_conv4", 4, -1, 128, 3, false, true, true},    {"enc_conv5a", 5, -1, 128, 4, false, false, true},
    {"enc_conv5b", 6, -1, 128, 4, false, false, true},  {"dec_conv4a", 7, 4, 112, 3, true, false, true},
    {"dec_conv4b", 8, -1, 112, 3, false, false, true},  {"dec_conv3a", 9, 3, 96, 2, true, false, true},
    {"dec_conv3b", 10, -1, 96, 2, false, false, true},  {"dec_conv2a", 11, 2, 64, 1, true, false, true},
    {"dec_conv2b", 12, -1, 64, 1, false, false, true},  {"dec_conv1a", 13, 0, 64, 0, true, false, true},
    {"dec_conv1b", 14, -1, 32, 0, false, false, true},  {"dec_conv0", 15, -1, 3, 0, false, false, false},
};
constexpr uint32_t TENSORS = PFXH_LEARN_LAYERS + 1;

uint32_t channels(int t) { return t == 0 ? INPUT_CHANNELS : NET[t - 1].out; }
uint32_t tensor_level(int t) { return t == 0 ? 0 : NET[t - 1].level + (NET[t - 1].pool ? 1u : 0u); }

struct Shape {
    uint32_t c0, c1, np, cin;  // padded
    uint32_t r0, r1;           // real input channels
};
Shape shape(uint32_t k) {
    const Def& d = NET[k];
    Shape s;
    s.r0 = channels(d.in0);
    s.r1 = d.in1 >= 0 ? channels(d.in1) : 0;
    s.c0 = pad16(s.r0);
    s.c1 = d.in1 >= 0 ? pad16(s.r1) : 0;
    s.np = pad16(d.out);
    s.cin = s.c0 + s.c1;
    return s;
}
// canonical parameters: per layer weights [out][r0 + r1][3][3], then bias [out]
uint64_t canonical_count() {
    uint64_t n = 0;
    for (uint32_t k = 0; k < PFXH_LEARN_LAYERS; ++k) {
        Shape s = shape(k);
        n += (uint64_t)NET[k].out * (s.r0 + s.r1) * 9 + NET[k].out;
    }
    return n;
}

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

// ---- device state ----------------------------------------------------------------------
struct pfxh_learner {
    hipStream_t stream = nullptr;
    bool bf16 = true;
    // padded master parameters, per layer [np][cin][9] then [np]; gradients and
    // Adam moments in the same layout
    std::vector<size_t> w_off, b_off;
    size_t total = 0;
    float *master = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    std::vector<float*> wf, wt;  // packed for k_conv: forward, transposed
    int adam_t = 0;
    uint32_t loss_mode = 0;  // PFXH_LEARN_LOSS_*
    uint32_t passes = 1;     // inference: mirrored copies of the frame averaged (1, 2 or 4)
    // training set: encoded inputs [crop][side][side][16], targets [crop][side][side]
    uint32_t side = 0, capacity = 0, crops = 0;
    float* inputs = nullptr;
    float4* targets = nullptr;
    float4* staging = nullptr;  // one crop of raw images on the way in
    std::vector<void*> mem;

    // activations of one frame size: batch x h x w (multiples of 16)
    struct Frame {
        uint32_t batch = 0, h = 0, w = 0;
        bool train = false;
        float* t[TENSORS] = {};
        float* r[PFXH_LEARN_LAYERS] = {};  // pre-pool outputs (training)
        float* dt[TENSORS] = {};
        float *dpre = nullptr, *dx = nullptr, *part = nullptr, *part_b = nullptr, *sums = nullptr;
        float4* target = nullptr;
        float4* images = nullptr;  // inference: colour, albedo, normal, output
        float* depth = nullptr;
        float* deviation = nullptr;  // inference: the noise channel's input
        uint32_t* picks = nullptr;
        float* losses = nullptr;
        uint32_t pick_cap = 0, loss_cap = 0;
        std::vector<pfxl::DwPlan> plans;
        std::vector<void*> mem;
    };
    Frame train_frame, infer_frame;

    void* dalloc(std::vector<void*>& owner, size_t bytes, const char* what) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Failure{PFXH_REFUSED, std::string("the learned denoiser's ") + what + " need " +
                                            std::to_string(bytes) + " bytes on the GPU: " + hipGetErrorString(e)};
        }
        owner.push_back(p);
        return p;
    }
    static void release(std::vector<void*>& owner) {
        for (void* p : owner) (void)hipFree(p);
        owner.clear();
    }

    void init() {
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "learner stream");
        for (uint32_t k = 0; k < PFXH_LEARN_LAYERS; ++k) {
            Shape s = shape(k);
            w_off.push_back(total);
            total += (size_t)s.np * s.cin * 9;
            b_off.push_back(total);
            total += s.np;
        }
        master = (float*)dalloc(mem, total * 4, "parameters");
        grad = (float*)dalloc(mem, total * 4, "gradients");
        m1 = (float*)dalloc(mem, total * 4, "moments");
        m2 = (float*)dalloc(mem, total * 4, "moments");
        for (float* p : {master, grad, m1, m2}) hip_check(hipMemsetAsync(p, 0, total * 4, stream), "learner clear");
        for (uint32_t k = 0; k < PFXH_LEARN_LAYERS; ++k) {
            Shape s = shape(k);
            wf.push_back((float*)dalloc(mem, (size_t)s.np * s.cin * 9 * 4, "packed weights"));
            wt.push_back((float*)dalloc(mem, (size_t)s.np * s.cin * 9 * 4, "packed weights"));
        }
        repack();
        hip_check(hipStreamSynchronize(stream), "learner init");
    }
    void repack() {
        for (uint32_t k = 0; k < PFXH_LEARN_LAYERS; ++k) {
            Shape s = shape(k);
            pfxl::pack(stream, master + w_off[k], s.np, s.cin, wf[k], wt[k], bf16);
        }
    }

    // canonical index -> padded master index, for layer k
    template <class F>
    void each_param(F&& f) const {
        uint64_t c = 0;
        for (uint32_t k = 0; k < PFXH_LEARN_LAYERS; ++k) {
            Shape s = shape(k);
            uint32_t real_in = s.r0 + s.r1;
            for (uint32_t o = 0; o < NET[k].out; ++o)
                for (uint32_t ci = 0; ci < real_in; ++ci) {
                    uint32_t pc = ci < s.r0 ? ci : s.c0 + (ci - s.r0);
                    for (uint32_t tap = 0; tap < 9; ++tap) f(c++, w_off[k] + ((size_t)o * s.cin + pc) * 9 + tap);
                }
            for (uint32_t o = 0; o < NET[k].out; ++o) f(c++, b_off[k] + o);
        }
    }
    void download(const float* dev, float* out) {
        std::vector<float> host(total);
        hip_check(hipMemcpyAsync(host.data(), dev, total * 4, hipMemcpyDeviceToHost, stream), "learner readback");
        hip_check(hipStreamSynchronize(stream), "learner readback");
        each_param([&](uint64_t c, size_t i) { out[c] = host[i]; });
    }
    void upload(const float* in) {
        std::vector<float> host(total, 0.0f);
        each_param([&](uint64_t c, size_t i) { host[i] = in[c]; });
        hip_check(hipMemcpyAsync(master, host.data(), total * 4, hipMemcpyHostToDevice, stream), "learner upload");
        repack();
        hip_check(hipStreamSynchronize(stream), "learner upload");
    }

    void frame(Frame& f, uint32_t batch, uint32_t h, uint32_t w, bool train) {
        if (f.batch == batch && f.h == h && f.w == w && f.train == train) return;
        release(f.mem);
        f = Frame();
        const size_t px = (size_t)batch * h * w;
        for (uint32_t t = 0; t < TENSORS; ++t) {
            size_t n = (px >> (2 * tensor_level((int)t))) * pad16(channels((int)t));
            f.t[t] = (float*)dalloc(f.mem, n * 4, "activations");
            if (train && t > 0) f.dt[t] = (float*)dalloc(f.mem, n * 4, "activation gradients");
        }
        size_t scratch_pre = 0, scratch_x = 0, part = 0, part_b = 0;
        for (uint32_t k = 0; k < PFXH_LEARN_LAYERS; ++k) {
            Shape s = shape(k);
            size_t lp = px >> (2 * NET[k].level);
            if (train && NET[k].pool) f.r[k] = (float*)dalloc(f.mem, lp * s.np * 4, "activations");
            scratch_pre = std::max(scratch_pre, lp * s.np);
            scratch_x = std::max(scratch_x, lp * s.cin);
            pfxl::DwPlan pl = pfxl::dw_plan(s.np, s.cin, batch, h >> NET[k].level, w >> NET[k].level);
            if (pl.chunks == 0) throw Failure{PFXH_SETUP, std::string(NET[k].name) + ": no weight-gradient launch fits"};
            f.plans.push_back(pl);
            part = std::max(part, (size_t)pl.chunks * 9 * s.np * s.cin);
            part_b = std::max(part_b, (size_t)pl.chunks * s.np);
        }
        if (train) {
            f.dpre = (float*)dalloc(f.mem, scratch_pre * 4, "gradient scratch");
            f.dx = (float*)dalloc(f.mem, scratch_x * 4, "gradient scratch");
            f.part = (float*)dalloc(f.mem, part * 4, "gradient partial sums");
            f.part_b = (float*)dalloc(f.mem, part_b * 4, "gradient partial sums");
            f.sums = (float*)dalloc(f.mem, ((px + 255) / 256) * 4, "loss sums");
            f.target = (float4*)dalloc(f.mem, px * 16, "targets");
        } else {
            f.images = (float4*)dalloc(f.mem, px * 16 * 4, "images");
            f.depth = (float*)dalloc(f.mem, px * 4, "depth");
            f.deviation = (float*)dalloc(f.mem, px * 4, "deviation");
        }
        f.batch = batch;
        f.h = h;
        f.w = w;
        f.train = train;
    }

    ConvDev conv(const Frame& f, uint32_t k, bool fused_pool) const {
        const Def& d = NET[k];
        Shape s = shape(k);
        ConvDev c;
        c.src0 = f.t[d.in0];
        c.src1 = d.in1 >= 0 ? f.t[d.in1] : nullptr;
        c.dst = d.pool && !fused_pool ? f.r[k] : f.t[k + 1];
        c.w = wf[k];
        c.bias = master + b_off[k];
        c.c0 = s.c0;
        c.c1 = s.c1;
        c.n_pad = s.np;
        c.h = f.h >> d.level;
        c.w_ = f.w >> d.level;
        c.upsample = d.up ? 1u : 0u;
        c.pool = d.pool && fused_pool ? 1u : 0u;
        c.relu = d.relu ? 1u : 0u;
        c.bf16 = bf16 ? 1u : 0u;
        c.batch = f.batch;
        return c;
    }
    void forward(Frame& f) {
        for (uint32_t k = 0; k < PFXH_LEARN_LAYERS; ++k) {
            pfxd::launch_conv(stream, conv(f, k, !f.train));
            if (f.train && NET[k].pool)
                pfxl::pool_fwd(stream, f.r[k], f.batch, f.h >> NET[k].level, f.w >> NET[k].level, shape(k).np,
                               f.t[k + 1]);
        }
    }
    void backward(Frame& f) {
        const size_t px = (size_t)f.batch * f.h * f.w;
        for (uint32_t t = 1; t + 1 < TENSORS; ++t) {
            size_t n = (px >> (2 * tensor_level((int)t))) * pad16(channels((int)t));
            hip_check(hipMemsetAsync(f.dt[t], 0, n * 4, stream), "gradient clear");
        }
        for (int k = PFXH_LEARN_LAYERS - 1; k >= 0; --k) {
            const Def& d = NET[k];
            Shape s = shape((uint32_t)k);
            uint32_t h = f.h >> d.level, w = f.w >> d.level;
            size_t n = (size_t)f.batch * h * w * s.np;
            const float* dpre = f.dt[k + 1];
            if (d.pool) {
                pfxl::pool_bwd(stream, f.dt[k + 1], f.r[k], f.batch, h, w, s.np, f.dpre);
                dpre = f.dpre;
            } else if (d.relu) {
                pfxl::relu_bwd(stream, f.dt[k + 1], f.t[k + 1], n, f.dpre);
                dpre = f.dpre;
            }
            pfxl::DwArgs a;
            a.dpre = dpre;
            a.src0 = f.t[d.in0];
            a.src1 = d.in1 >= 0 ? f.t[d.in1] : nullptr;
            a.part = f.part;
            a.part_b = f.part_b;
            a.c0 = s.c0;
            a.c1 = s.c1;
            a.np = s.np;
            a.h = h;
            a.w = w;
            a.batch = f.batch;
            a.up = d.up ? 1u : 0u;
            a.ct = 0;
            pfxl::weight_grads(stream, a, f.plans[(size_t)k], bf16, grad + w_off[(size_t)k], grad + b_off[(size_t)k]);
            float* d0 = d.in0 > 0 ? f.dt[d.in0] : nullptr;
            float* d1 = d.in1 > 0 ? f.dt[d.in1] : nullptr;
            if (!d0 && !d1) continue;
            ConvDev c;
            c.src0 = dpre;
            c.src1 = nullptr;
            c.dst = f.dx;
            c.w = wt[(size_t)k];
            c.bias = nullptr;
            c.c0 = s.np;
            c.c1 = 0;
            c.n_pad = s.cin;
            c.h = h;
            c.w_ = w;
            c.upsample = 0;
            c.pool = 0;
            c.relu = 0;
            c.bf16 = bf16 ? 1u : 0u;
            c.batch = f.batch;
            pfxd::launch_conv(stream, c);
            pfxl::scatter(stream, f.dx, f.batch, h, w, s.c0, s.c1, d.up, d0, d1);
        }
    }
    // forward + loss + backward on the picked crops; the mean loss lands in
    // f.losses[slot]
    void gradient(Frame& f, const uint32_t* picks_dev, uint32_t slot) {
        const size_t px = (size_t)f.batch * f.h * f.w;
        pfxl::gather(stream, inputs, targets, (size_t)side * side, side, picks_dev, f.batch, f.t[0], f.target);
        forward(f);
        uint32_t nb = pfxl::loss(stream, f.t[TENSORS - 1], f.target, px, loss_mode, f.dt[TENSORS - 1], f.sums);
        pfxl::loss_total(stream, f.sums, nb, px, f.losses + slot);
        backward(f);
        hip_check(hipGetLastError(), "training launch");
    }
    void ensure_io(Frame& f, uint32_t picks, uint32_t losses) {
        if (picks > f.pick_cap) {
            f.picks = (uint32_t*)dalloc(f.mem, (size_t)picks * 4, "batch picks");
            f.pick_cap = picks;
        }
        if (losses > f.loss_cap) {
            f.losses = (float*)dalloc(f.mem, (size_t)losses * 4, "losses");
            f.loss_cap = losses;
        }
    }

    // Size the inference frame for a width x height image and the current
    // pass count; pointers into it stay valid until the next call.
    Frame& prepare_inference(uint32_t width, uint32_t height) {
        frame(infer_frame, passes, pad16(height), pad16(width), false);
        return infer_frame;
    }
    void infer(uint32_t width, uint32_t height, const float4* color, const float4* albedo, const float4* normal,
               const float* depth, uint32_t depth_stride, const float* deviation, float4* out) {
        uint32_t w16 = pad16(width), h16 = pad16(height);
        // the caller sized the frame (prepare_inference) before taking
        // pointers into it: sizing it here could free them
        Frame& f = infer_frame;
        if (f.batch != passes || f.h != h16 || f.w != w16 || f.train)
            throw Failure{PFXH_SETUP, "inference frame not prepared for this image"};
        pfxl::learn_in(stream, color, albedo, normal, depth, depth_stride, deviation, width, height, w16, h16, passes,
                       f.t[0]);
        forward(f);
        pfxl::learn_out(stream, f.t[TENSORS - 1], w16, h16, passes, color, width, height, out);
        hip_check(hipGetLastError(), "learned denoise launch");
    }

    ~pfxh_learner() {
        if (stream) (void)hipStreamSynchronize(stream);
        release(train_frame.mem);
        release(infer_frame.mem);
        release(mem);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Internal: the learned denoiser on device images of `width` x `height`
// (depth with a stride of `depth_stride` floats; deviation, the standard
// error of each pixel's mean luminance, or null), written on `after`; the
// result is copied to the host buffer `out`, or with `out` null left on the
// device at *device_out (valid until the next call).
pfxh_status pfxh_learned_denoise_device(pfxh_learner* l, uint32_t width, uint32_t height, const void* color,
                                        const void* albedo, const void* normal, const float* depth,
                                        uint32_t depth_stride, const float* deviation, float* out,
                                        hipStream_t after, char* err, size_t err_len, void** device_out) {
    return guard(err, err_len, [&] {
        if (width == 0 || height == 0) throw Failure{PFXH_INVALID, "image dimensions must be nonzero"};
        if (after) hip_check(hipStreamSynchronize(after), "learned denoise input");
        uint32_t w16 = pad16(width), h16 = pad16(height);
        float4* o = l->prepare_inference(width, height).images + 3 * (size_t)w16 * h16;
        l->infer(width, height, (const float4*)color, (const float4*)albedo, (const float4*)normal, depth,
                 depth_stride, deviation, o);
        if (device_out) *device_out = o;
        if (out)
            hip_check(hipMemcpyAsync(out, o, (size_t)width * height * 16, hipMemcpyDeviceToHost, l->stream),
                      "learned denoise readback");
        hip_check(hipStreamSynchronize(l->stream), "learned denoise");
    });
}

extern "C" {

pfxh_status pfxh_learn_layer(uint32_t index, pfxh_learn_layer_info* out) {
    if (!out || index >= PFXH_LEARN_LAYERS) return PFXH_INVALID;
    const Def& d = NET[index];
    Shape s = shape(index);
    memset(out, 0, sizeof(*out));
    strncpy(out->name, d.name, sizeof(out->name) - 1);
    out->in0 = d.in0;
    out->in1 = d.in1;
    out->in_channels = s.r0 + s.r1;
    out->out_channels = d.out;
    out->level = d.level;
    out->upsample = d.up ? 1u : 0u;
    out->pool = d.pool ? 1u : 0u;
    out->relu = d.relu ? 1u : 0u;
    return PFXH_OK;
}

uint64_t pfxh_learn_param_count(void) { return canonical_count(); }

pfxh_status pfxh_learner_new(uint32_t flags, pfxh_learner** out, char* err, size_t err_len) {
    if (!out) return PFXH_INVALID;
    *out = nullptr;
    pfxh_learner* l = nullptr;
    pfxh_status st = guard(err, err_len, [&] {
        l = new pfxh_learner();
        l->bf16 = (flags & PFXH_LEARN_FP32) == 0;
        l->init();
    });
    if (st != PFXH_OK) {
        delete l;
        return st;
    }
    *out = l;
    return PFXH_OK;
}

pfxh_status pfxh_learner_set_params(pfxh_learner* l, const float* params, uint64_t count, char* err,
                                    size_t err_len) {
    if (!l || !params) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (count != canonical_count())
            throw Failure{PFXH_INVALID, "expected " + std::to_string(canonical_count()) + " parameters, got " +
                                            std::to_string(count)};
        for (uint64_t i = 0; i < count; ++i)
            if (!std::isfinite(params[i])) throw Failure{PFXH_INVALID, "parameters must be finite"};
        l->upload(params);
    });
}

pfxh_status pfxh_learner_params(pfxh_learner* l, float* out, uint64_t count, char* err, size_t err_len) {
    if (!l || !out) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (count != canonical_count()) throw Failure{PFXH_INVALID, "parameter buffer has the wrong length"};
        l->download(l->master, out);
    });
}

pfxh_status pfxh_learner_grads(pfxh_learner* l, float* out, uint64_t count, char* err, size_t err_len) {
    if (!l || !out) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (count != canonical_count()) throw Failure{PFXH_INVALID, "gradient buffer has the wrong length"};
        l->download(l->grad, out);
    });
}

pfxh_status pfxh_learner_set_loss(pfxh_learner* l, uint32_t mode) {
    if (!l || mode > PFXH_LEARN_LOSS_REL_L2) return PFXH_INVALID;
    l->loss_mode = mode;
    return PFXH_OK;
}

pfxh_status pfxh_learner_set_passes(pfxh_learner* l, uint32_t passes) {
    if (!l || !(passes == 1u || passes == 2u || passes == 4u)) return PFXH_INVALID;
    l->passes = passes;
    return PFXH_OK;
}

void pfxh_learner_reset_optimizer(pfxh_learner* l) {
    if (!l) return;
    l->adam_t = 0;
    (void)hipMemsetAsync(l->m1, 0, l->total * 4, l->stream);
    (void)hipMemsetAsync(l->m2, 0, l->total * 4, l->stream);
    (void)hipStreamSynchronize(l->stream);
}

pfxh_status pfxh_learner_reserve(pfxh_learner* l, uint32_t side, uint32_t capacity, char* err, size_t err_len) {
    if (!l) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (side == 0 || side % 16 != 0) throw Failure{PFXH_INVALID, "crop side must be a nonzero multiple of 16"};
        if (capacity == 0) throw Failure{PFXH_INVALID, "capacity must be nonzero"};
        hip_check(hipStreamSynchronize(l->stream), "learner reserve");
        for (void* p : {(void*)l->inputs, (void*)l->targets, (void*)l->staging}) {
            auto it = std::find(l->mem.begin(), l->mem.end(), p);
            if (p && it != l->mem.end()) {
                (void)hipFree(p);
                l->mem.erase(it);
            }
        }
        l->inputs = nullptr;
        l->targets = nullptr;
        l->staging = nullptr;
        l->crops = 0;
        size_t px = (size_t)side * side;
        l->inputs = (float*)l->dalloc(l->mem, px * capacity * 16 * 4, "training inputs");
        l->targets = (float4*)l->dalloc(l->mem, px * capacity * 16, "training targets");
        l->staging = (float4*)l->dalloc(l->mem, px * 16 * 3 + px * 8, "crop staging");
        l->side = side;
        l->capacity = capacity;
    });
}

pfxh_status pfxh_learner_add(pfxh_learner* l, const float* color, const float* albedo, const float* normal,
                             const float* depth, const float* deviation, const float* target, char* err,
                             size_t err_len) {
    if (!l || !color || !albedo || !normal || !target) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (!l->inputs) throw Failure{PFXH_INVALID, "no training set reserved"};
        if (l->crops >= l->capacity) throw Failure{PFXH_INVALID, "the training set is full"};
        size_t px = (size_t)l->side * l->side, bytes = px * 16;
        float4* st = l->staging;
        float* dep = depth ? (float*)(st + 3 * px) : nullptr;
        float* dev = deviation ? (float*)(st + 3 * px) + px : nullptr;
        hip_check(hipMemcpyAsync(st, color, bytes, hipMemcpyHostToDevice, l->stream), "crop upload");
        hip_check(hipMemcpyAsync(st + px, albedo, bytes, hipMemcpyHostToDevice, l->stream), "crop upload");
        hip_check(hipMemcpyAsync(st + 2 * px, normal, bytes, hipMemcpyHostToDevice, l->stream), "crop upload");
        if (dep) hip_check(hipMemcpyAsync(dep, depth, px * 4, hipMemcpyHostToDevice, l->stream), "crop upload");
        if (dev) hip_check(hipMemcpyAsync(dev, deviation, px * 4, hipMemcpyHostToDevice, l->stream), "crop upload");
        hip_check(hipMemcpyAsync(l->targets + px * l->crops, target, bytes, hipMemcpyHostToDevice, l->stream),
                  "crop upload");
        pfxl::learn_in(l->stream, st, st + px, st + 2 * px, dep, 1, dev, l->side, l->side, l->side, l->side, 1u,
                       l->inputs + px * 16 * l->crops);
        hip_check(hipGetLastError(), "crop encode");
        hip_check(hipStreamSynchronize(l->stream), "crop upload");
        l->crops++;
    });
}

uint32_t pfxh_learner_crops(const pfxh_learner* l) { return l ? l->crops : 0; }

pfxh_status pfxh_learner_gradient(pfxh_learner* l, const uint32_t* picks, uint32_t batch, float* loss, char* err,
                                  size_t err_len) {
    if (!l || !picks || !loss) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (batch == 0) throw Failure{PFXH_INVALID, "batch must be nonzero"};
        for (uint32_t b = 0; b < batch; ++b)
            if ((picks[b] >> 3) >= l->crops) throw Failure{PFXH_INVALID, "pick names a crop not in the training set"};
        pfxh_learner::Frame& f = l->train_frame;
        l->frame(f, batch, l->side, l->side, true);
        l->ensure_io(f, batch, 1);
        hip_check(hipMemcpyAsync(f.picks, picks, (size_t)batch * 4, hipMemcpyHostToDevice, l->stream), "picks");
        l->gradient(f, f.picks, 0);
        hip_check(hipMemcpyAsync(loss, f.losses, 4, hipMemcpyDeviceToHost, l->stream), "loss readback");
        hip_check(hipStreamSynchronize(l->stream), "gradient");
    });
}

pfxh_status pfxh_learner_train(pfxh_learner* l, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                               float* losses, char* err, size_t err_len) {
    if (!l || !losses) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (batch == 0) throw Failure{PFXH_INVALID, "batch must be nonzero"};
        if (l->crops == 0) throw Failure{PFXH_INVALID, "the training set is empty"};
        if (!(lr > 0.0f) || !std::isfinite(lr)) throw Failure{PFXH_INVALID, "learning rate must be positive"};
        if (steps == 0) return;
        pfxh_learner::Frame& f = l->train_frame;
        l->frame(f, batch, l->side, l->side, true);
        l->ensure_io(f, steps * batch, steps);
        std::vector<uint32_t> picks((size_t)steps * batch);
        uint64_t s = seed;
        for (uint32_t& p : picks) {
            uint64_t r = splitmix(s);
            p = (uint32_t)((r >> 3) % l->crops) << 3 | (uint32_t)(r & 7u);
        }
        hip_check(hipMemcpyAsync(f.picks, picks.data(), picks.size() * 4, hipMemcpyHostToDevice, l->stream), "picks");
        for (uint32_t k = 0; k < steps; ++k) {
            l->gradient(f, f.picks + (size_t)k * batch, k);
            l->adam_t++;
            pfxl::adam(l->stream, l->master, l->grad, l->m1, l->m2, l->total, lr, 0.9f, 0.999f, 1e-8f, l->adam_t);
            l->repack();
        }
        hip_check(hipGetLastError(), "training launch");
        hip_check(hipMemcpyAsync(losses, f.losses, (size_t)steps * 4, hipMemcpyDeviceToHost, l->stream),
                  "loss readback");
        hip_check(hipStreamSynchronize(l->stream), "training");
    });
}

pfxh_status pfxh_learned_denoise(pfxh_learner* l, uint32_t width, uint32_t height, const float* color,
                                 const float* albedo, const float* normal, const float* depth, const float* deviation,
                                 float* out, char* err, size_t err_len) {
    if (!l || !color || !albedo || !normal || !out) return PFXH_INVALID;
    return guard(err, err_len, [&] {
        if (width == 0 || height == 0) throw Failure{PFXH_INVALID, "image dimensions must be nonzero"};
        uint32_t w16 = pad16(width), h16 = pad16(height);
        pfxh_learner::Frame& f = l->prepare_inference(width, height);
        size_t px = (size_t)width * height, bytes = px * 16, cap = (size_t)w16 * h16;
        float4* im = f.images;
        hip_check(hipMemcpyAsync(im, color, bytes, hipMemcpyHostToDevice, l->stream), "learned denoise upload");
        hip_check(hipMemcpyAsync(im + cap, albedo, bytes, hipMemcpyHostToDevice, l->stream), "learned denoise upload");
        hip_check(hipMemcpyAsync(im + 2 * cap, normal, bytes, hipMemcpyHostToDevice, l->stream),
                  "learned denoise upload");
        if (depth) hip_check(hipMemcpyAsync(f.depth, depth, px * 4, hipMemcpyHostToDevice, l->stream), "upload");
        if (deviation)
            hip_check(hipMemcpyAsync(f.deviation, deviation, px * 4, hipMemcpyHostToDevice, l->stream), "upload");
        l->infer(width, height, im, im + cap, im + 2 * cap, depth ? f.depth : nullptr, 1,
                 deviation ? f.deviation : nullptr, im + 3 * cap);
        hip_check(hipMemcpyAsync(out, im + 3 * cap, bytes, hipMemcpyDeviceToHost, l->stream),
                  "learned denoise readback");
        hip_check(hipStreamSynchronize(l->stream), "learned denoise");
    });
}

void pfxh_learner_free(pfxh_learner* l) { delete l; }

}  // extern "C"
