// CPU restatement of the learned denoiser in FP64 -- TEST INFRASTRUCTURE
// ONLY (checker for the HIP training and inference kernels).
//
// The reference has no trained denoiser; what this restates is the network
// the HIP path defines (the U-Net shape of crates/denoise/src/net.rs:202-232
// at 32->128 channels) straight from its mathematics, with no channel
// padding and no kernel layouts, so layout mistakes in the HIP path show up:
//   input      log(1 + colour), albedo, normal * 0.5 + 0.5, log(1 + depth) / 8
//              (non-finite values read as 0), zero frame padding to a
//              multiple of 16
//   conv       3x3 cross-correlation with zero padding, inputs [src0
//              (nearest-upsampled 2x when the layer says so), src1]
//   relu, pool 2x2 max, the gradient going to the first maximum in
//              row-major order
//   loss       mean over pixels and the three colour channels of
//              |y - log(1 + target)| (mode 0), or of (x - target)^2 /
//              (x^2 + 0.01) with x = expm1(min(y, 40)) and the denominator
//              held fixed in the gradient (mode 1, relative L2)
//   backward   the chain rule through the above, by direct loops
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

struct L {
    int in0, in1;
    int out, level;
    bool up, pool, relu;
};
const L NET[16] = {
    {0, -1, 32, 0, false, false, true},  {1, -1, 32, 0, false, true, true},  {2, -1, 64, 1, false, true, true},
    {3, -1, 96, 2, false, true, true},   {4, -1, 128, 3, false, true, true}, {5, -1, 128, 4, false, false, true},
    {6, -1, 128, 4, false, false, true}, {7, 4, 112, 3, true, false, true},  {8, -1, 112, 3, false, false, true},
    {9, 3, 96, 2, true, false, true},    {10, -1, 96, 2, false, false, true}, {11, 2, 64, 1, true, false, true},
    {12, -1, 64, 1, false, false, true}, {13, 0, 64, 0, true, false, true},  {14, -1, 32, 0, false, false, true},
    {15, -1, 3, 0, false, false, false},
};
const int IN_CH = 11;

int chans(int t) { return t == 0 ? IN_CH : NET[t - 1].out; }
int cin_of(int k) { return chans(NET[k].in0) + (NET[k].in1 >= 0 ? chans(NET[k].in1) : 0); }

struct T {  // [b][h][w][c]
    int b = 0, h = 0, w = 0, c = 0;
    std::vector<double> v;
    void make(int b_, int h_, int w_, int c_) {
        b = b_, h = h_, w = w_, c = c_;
        v.assign((size_t)b * h * w * c, 0.0);
    }
    double& at(int n, int y, int x, int k) { return v[(((size_t)n * h + y) * w + x) * c + k]; }
    double at(int n, int y, int x, int k) const { return v[(((size_t)n * h + y) * w + x) * c + k]; }
};

double fin(float v) { return (v == v && fabsf(v) < 1e30f) ? (double)v : 0.0; }

// stacked input of layer k at its resolution
T stacked(const std::vector<T>& t, int k) {
    const L& d = NET[k];
    const T& a = t[d.in0];
    int h = d.up ? a.h * 2 : a.h, w = d.up ? a.w * 2 : a.w;
    int c0 = a.c, c1 = d.in1 >= 0 ? t[d.in1].c : 0;
    T s;
    s.make(a.b, h, w, c0 + c1);
    for (int n = 0; n < a.b; ++n)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                for (int c = 0; c < c0; ++c) s.at(n, y, x, c) = d.up ? a.at(n, y / 2, x / 2, c) : a.at(n, y, x, c);
                for (int c = 0; c < c1; ++c) s.at(n, y, x, c0 + c) = t[d.in1].at(n, y, x, c);
            }
    return s;
}

struct Net {
    const double* p;
    std::vector<size_t> w_off, b_off;
    explicit Net(const double* params) : p(params) {
        size_t o = 0;
        for (int k = 0; k < 16; ++k) {
            w_off.push_back(o);
            o += (size_t)NET[k].out * cin_of(k) * 9;
            b_off.push_back(o);
            o += NET[k].out;
        }
    }
    double W(int k, int o, int c, int ky, int kx) const {
        return p[w_off[k] + (((size_t)o * cin_of(k) + c) * 3 + ky) * 3 + kx];
    }
};

void conv(const Net& net, int k, const T& x, T& pre) {
    int co = NET[k].out;
    pre.make(x.b, x.h, x.w, co);
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < x.b; ++n)
        for (int y = 0; y < x.h; ++y)
            for (int xx = 0; xx < x.w; ++xx)
                for (int o = 0; o < co; ++o) {
                    double s = net.p[net.b_off[k] + o];
                    for (int ky = 0; ky < 3; ++ky) {
                        int sy = y + ky - 1;
                        if (sy < 0 || sy >= x.h) continue;
                        for (int kx = 0; kx < 3; ++kx) {
                            int sx = xx + kx - 1;
                            if (sx < 0 || sx >= x.w) continue;
                            for (int c = 0; c < x.c; ++c) s += x.at(n, sy, sx, c) * net.W(k, o, c, ky, kx);
                        }
                    }
                    pre.at(n, y, xx, o) = s;
                }
}

struct Pass {
    std::vector<T> t;     // tensors: 0 input, k + 1 output of layer k
    std::vector<T> x;     // stacked inputs
    std::vector<T> act;   // activation before pooling
};

void encode(int b, int h, int w, int h16, int w16, const float* color, const float* albedo, const float* normal,
            const float* depth, const float* deviation, T& x) {
    x.make(b, h16, w16, IN_CH);
    for (int n = 0; n < b; ++n)
        for (int y = 0; y < h; ++y)
            for (int xx = 0; xx < w; ++xx) {
                size_t p = ((size_t)n * h + y) * w + xx;
                const float* c = color + 4 * p;
                const float* a = albedo + 4 * p;
                const float* m = normal + 4 * p;
                double d = depth ? fin(depth[p]) : 0.0;
                for (int k = 0; k < 3; ++k) x.at(n, y, xx, k) = log1p(fmax(fin(c[k]), 0.0));
                for (int k = 0; k < 3; ++k) x.at(n, y, xx, 3 + k) = fin(a[k]);
                for (int k = 0; k < 3; ++k) x.at(n, y, xx, 6 + k) = fin(m[k]) * 0.5 + 0.5;
                x.at(n, y, xx, 9) = log1p(fmax(d, 0.0)) * 0.125;
                x.at(n, y, xx, 10) = deviation ? log1p(fmax(fin(deviation[p]), 0.0)) : 0.0;
            }
}

void forward(const Net& net, Pass& ps) {
    ps.t.resize(17);
    ps.x.resize(16);
    ps.act.resize(16);
    for (int k = 0; k < 16; ++k) {
        ps.x[k] = stacked(ps.t, k);
        T pre;
        conv(net, k, ps.x[k], pre);
        if (NET[k].relu)
            for (double& v : pre.v) v = v > 0.0 ? v : 0.0;
        if (!NET[k].pool) {
            ps.t[k + 1] = pre;
            continue;
        }
        ps.act[k] = pre;
        T& o = ps.t[k + 1];
        o.make(pre.b, pre.h / 2, pre.w / 2, pre.c);
        for (int n = 0; n < pre.b; ++n)
            for (int y = 0; y < o.h; ++y)
                for (int xx = 0; xx < o.w; ++xx)
                    for (int c = 0; c < pre.c; ++c)
                        o.at(n, y, xx, c) = fmax(fmax(pre.at(n, 2 * y, 2 * xx, c), pre.at(n, 2 * y, 2 * xx + 1, c)),
                                                 fmax(pre.at(n, 2 * y + 1, 2 * xx, c), pre.at(n, 2 * y + 1, 2 * xx + 1, c)));
    }
}

}  // namespace

extern "C" {

uint64_t or_learn_param_count(void) {
    uint64_t n = 0;
    for (int k = 0; k < 16; ++k) n += (uint64_t)NET[k].out * cin_of(k) * 9 + NET[k].out;
    return n;
}

void or_learn_layer(uint32_t k, int32_t* out /* in0, in1, cin, cout, level, up, pool, relu */) {
    const L& d = NET[k];
    int32_t v[8] = {d.in0, d.in1, cin_of((int)k), d.out, d.level, d.up, d.pool, d.relu};
    memcpy(out, v, sizeof(v));
}

// Loss (returned) and, when `grad` is not null, its gradient for `batch`
int g_loss_mode = 0;
void or_learn_set_loss(int mode) { g_loss_mode = mode; }

// crops of h x w pixels (multiples of 16) stored one after another.
double or_learn_grad(const double* params, uint32_t batch, uint32_t h, uint32_t w, const float* color,
                     const float* albedo, const float* normal, const float* depth, const float* deviation,
                     const float* target, double* grad) {
    Net net(params);
    Pass ps;
    ps.t.resize(17);
    encode((int)batch, (int)h, (int)w, (int)h, (int)w, color, albedo, normal, depth, deviation, ps.t[0]);
    forward(net, ps);
    const T& y = ps.t[16];
    size_t px = (size_t)batch * h * w;
    double loss = 0.0;
    std::vector<T> g(17);
    g[16].make(y.b, y.h, y.w, y.c);
    double scale = 1.0 / (3.0 * (double)px);
    for (size_t p = 0; p < px; ++p)
        for (int k = 0; k < 3; ++k) {
            double t = fmax((double)target[4 * p + k], 0.0);
            if (g_loss_mode == 1) {
                double v = fmin(y.v[p * 3 + k], 40.0);
                double x = expm1(v), den = x * x + 0.01, d = x - t;
                loss += d * d / den;
                g[16].v[p * 3 + k] = y.v[p * 3 + k] > 40.0 ? 0.0 : scale * (2.0 * d / den) * (x + 1.0);
            } else {
                double d = y.v[p * 3 + k] - log1p(t);
                loss += fabs(d);
                g[16].v[p * 3 + k] = d > 0.0 ? scale : (d < 0.0 ? -scale : 0.0);
            }
        }
    loss *= scale;
    if (!grad) return loss;
    memset(grad, 0, or_learn_param_count() * sizeof(double));
    for (int t = 1; t < 16; ++t) g[t].make(ps.t[t].b, ps.t[t].h, ps.t[t].w, ps.t[t].c);
    for (int k = 15; k >= 0; --k) {
        const L& d = NET[k];
        const T& x = ps.x[k];
        // gradient before the activation, at the conv's resolution
        T dpre;
        if (d.pool) {
            const T& a = ps.act[k];
            dpre.make(a.b, a.h, a.w, a.c);
            for (int n = 0; n < a.b; ++n)
                for (int yy = 0; yy < a.h / 2; ++yy)
                    for (int xx = 0; xx < a.w / 2; ++xx)
                        for (int c = 0; c < a.c; ++c) {
                            int by = 2 * yy, bx = 2 * xx;
                            for (int q = 1; q < 4; ++q) {
                                int qy = 2 * yy + q / 2, qx = 2 * xx + q % 2;
                                if (a.at(n, qy, qx, c) > a.at(n, by, bx, c)) by = qy, bx = qx;
                            }
                            if (a.at(n, by, bx, c) > 0.0) dpre.at(n, by, bx, c) = g[k + 1].at(n, yy, xx, c);
                        }
        } else {
            dpre = g[k + 1];
            if (d.relu)
                for (size_t i = 0; i < dpre.v.size(); ++i)
                    if (!(ps.t[k + 1].v[i] > 0.0)) dpre.v[i] = 0.0;
        }
        // parameters
        int co = d.out, ci = x.c;
#pragma omp parallel for collapse(2) schedule(static)
        for (int o = 0; o < co; ++o)
            for (int c = 0; c < ci; ++c)
                for (int ky = 0; ky < 3; ++ky)
                    for (int kx = 0; kx < 3; ++kx) {
                        double s = 0.0;
                        for (int n = 0; n < x.b; ++n)
                            for (int yy = 0; yy < x.h; ++yy) {
                                int sy = yy + ky - 1;
                                if (sy < 0 || sy >= x.h) continue;
                                for (int xx = 0; xx < x.w; ++xx) {
                                    int sx = xx + kx - 1;
                                    if (sx < 0 || sx >= x.w) continue;
                                    s += dpre.at(n, yy, xx, o) * x.at(n, sy, sx, c);
                                }
                            }
                        grad[net.w_off[k] + (((size_t)o * ci + c) * 3 + ky) * 3 + kx] = s;
                    }
        for (int o = 0; o < co; ++o) {
            double s = 0.0;
            for (size_t p = 0; p < dpre.v.size() / co; ++p) s += dpre.v[p * co + o];
            grad[net.b_off[k] + o] = s;
        }
        // inputs
        bool need0 = d.in0 > 0, need1 = d.in1 > 0;
        if (!need0 && !need1) continue;
        T dx;
        dx.make(x.b, x.h, x.w, ci);
#pragma omp parallel for collapse(2) schedule(static)
        for (int n = 0; n < x.b; ++n)
            for (int sy = 0; sy < x.h; ++sy)
                for (int sx = 0; sx < x.w; ++sx)
                    for (int c = 0; c < ci; ++c) {
                        double s = 0.0;
                        for (int ky = 0; ky < 3; ++ky) {
                            int yy = sy - ky + 1;
                            if (yy < 0 || yy >= x.h) continue;
                            for (int kx = 0; kx < 3; ++kx) {
                                int xx = sx - kx + 1;
                                if (xx < 0 || xx >= x.w) continue;
                                for (int o = 0; o < co; ++o) s += dpre.at(n, yy, xx, o) * net.W(k, o, c, ky, kx);
                            }
                        }
                        dx.at(n, sy, sx, c) = s;
                    }
        int c0 = chans(d.in0);
        for (int n = 0; n < x.b; ++n)
            for (int yy = 0; yy < x.h; ++yy)
                for (int xx = 0; xx < x.w; ++xx) {
                    if (need0)
                        for (int c = 0; c < c0; ++c) {
                            double v = dx.at(n, yy, xx, c);
                            if (d.up)
                                g[d.in0].at(n, yy / 2, xx / 2, c) += v;
                            else
                                g[d.in0].at(n, yy, xx, c) += v;
                        }
                    if (need1)
                        for (int c = 0; c < ci - c0; ++c) g[d.in1].at(n, yy, xx, c) += dx.at(n, yy, xx, c0 + c);
                }
    }
    return loss;
}

// Inference on one image of any size: RGBA out (alpha from the colour).
void or_learn_forward(const double* params, uint32_t h, uint32_t w, const float* color, const float* albedo,
                      const float* normal, const float* depth, const float* deviation, float* out) {
    Net net(params);
    Pass ps;
    ps.t.resize(17);
    int h16 = (int)((h + 15) / 16 * 16), w16 = (int)((w + 15) / 16 * 16);
    encode(1, (int)h, (int)w, h16, w16, color, albedo, normal, depth, deviation, ps.t[0]);
    forward(net, ps);
    const T& y = ps.t[16];
    for (uint32_t yy = 0; yy < h; ++yy)
        for (uint32_t xx = 0; xx < w; ++xx) {
            size_t p = (size_t)yy * w + xx;
            for (int k = 0; k < 3; ++k) out[4 * p + k] = (float)expm1(fmax(y.at(0, (int)yy, (int)xx, k), 0.0));
            out[4 * p + 3] = color[4 * p + 3];
        }
}

}  // extern "C"
