// Wavefront path tracer for gfx950 (MI355X).
//
// One wave of paths lives in HBM as structure-of-arrays state. Each bounce
// runs these kernels over compacted queues:
//   extend : closest-hit traversal of every live path over 128-byte 4-wide
//            nodes; persistent waves with an LDS short stack refill finished
//            lanes and alternate inner-node and leaf phases by lane majority
//   shade  : emission, next-event candidates, BSDF sampling, Russian
//            roulette; appends survivors to the next queue, paths with light
//            contributions to the shadow path queue and each contribution
//            that needs a visibility test to the shadow ray queue
//   shadow : any-hit traversal of every queued shadow ray, marking occluded
//            contributions
//   fold   : adds each path's unoccluded contributions in the reference's
//            slot order
// followed, per wave, by `finish`, which folds each pixel's samples into the
// accumulator in sample order.
//
// The per-vertex arithmetic restates crates/trace/src/trace.wgsl `radiance`
// (:2153-2502) and crates/trace/src/lights.wgsl with the same operation
// order; everything is built with -ffp-contract=off and IEEE div/sqrt.
#include <hip/hip_runtime.h>

#include "../../include/pfxhip.h"
#include "kernels.h"
#include "sobol.h"
#include "vec3.h"

namespace pfxh {

using pfx::basis;
using pfx::dot;
using pfx::fm_acos;
using pfx::fm_clamp;
using pfx::fm_cos;
using pfx::fm_pow;
using pfx::fm_sin;
using pfx::m3;
using pfx::mk;
using pfx::normalize;
using pfx::v3;

#define DEV __device__ __forceinline__

constexpr float PI = 3.141592653589793f;
constexpr uint32_t ONE_SIDED = 8388608u;
constexpr uint32_t NO_MATERIAL = 0xffffffffu;
constexpr uint32_t LAST_BIT = 1u << 24;

DEV v3 f3(float4 a) { return mk(a.x, a.y, a.z); }
DEV float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
DEV uint32_t fbits(float x) { return __float_as_uint(x); }
DEV float ubits(uint32_t x) { return __uint_as_float(x); }

// ---- wave-aggregated queue append ---------------------------------------------
DEV uint32_t wave_append(uint32_t* counter, bool pred, uint32_t amount_if = 1) {
    (void)amount_if;
    unsigned long long mask = __ballot(pred);
    uint32_t lane = __lane_id();
    unsigned long long below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
    uint32_t rank = (uint32_t)__popcll(below);
    uint32_t total = (uint32_t)__popcll(mask);
    if (total == 0) return 0;
    int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, total);
    base = (uint32_t)__shfl((int)base, leader);
    return base + rank;
}

// Inclusive prefix sum over the lanes of a wave.
DEV uint32_t wave_scan(uint32_t v) {
    uint32_t lane = __lane_id();
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t u = (uint32_t)__shfl_up((int)v, off);
        if ((int)lane >= off) v += u;
    }
    return v;
}

// Append a, b and c items per thread to the next-bounce, shadow path and
// shadow ray queues with one atomic per queue for the whole 256-thread
// block; returns each thread's first position in the three queues. Every
// thread of the block must call it.
DEV uint3 block_append3(uint32_t* counters, uint32_t a, uint32_t b, uint32_t c) {
    __shared__ uint32_t tot[3][4];
    __shared__ uint32_t base[3];
    uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t ia = wave_scan(a), ib = wave_scan(b), ic = wave_scan(c);
    if (lane == 63) {
        tot[0][wave] = ia;
        tot[1][wave] = ib;
        tot[2][wave] = ic;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t q = threadIdx.x;
        uint32_t sum = tot[q][0] + tot[q][1] + tot[q][2] + tot[q][3];
        const uint32_t slot[3] = {C_NEXT, C_SHADOW, C_SRAYS};
        base[q] = sum ? atomicAdd(counters + slot[q], sum) : 0u;
    }
    __syncthreads();
    uint3 r;
    r.x = base[0] + ia - a;
    r.y = base[1] + ib - b;
    r.z = base[2] + ic - c;
    for (uint32_t w = 0; w < wave; ++w) {
        r.x += tot[0][w];
        r.y += tot[1][w];
        r.z += tot[2][w];
    }
    return r;
}

// ---- traversal ------------------------------------------------------------------
// A ray in the space being traversed, with per-axis reciprocals.
struct Ray {
    v3 o, d, y;
    bool fast;  // no near-zero direction component, origin components 0 or >= 2^-72
};
DEV bool tame(float v) { return v == 0.0f || (fabsf(v) >= 0x1p-72f && fabsf(v) <= 0x1p60f); }
DEV Ray make_ray(v3 o, v3 d, bool bounds_tame) {
    Ray r;
    r.o = o;
    r.d = d;
    bool dx = fabsf(d.x) < 1e-12f, dy = fabsf(d.y) < 1e-12f, dz = fabsf(d.z) < 1e-12f;
    r.y = mk(dx ? 0.0f : 1.0f / d.x, dy ? 0.0f : 1.0f / d.y, dz ? 0.0f : 1.0f / d.z);
    // the quotient check covered |d| in [2^-39, 2^20]
    bool big = fabsf(d.x) > 0x1p20f || fabsf(d.y) > 0x1p20f || fabsf(d.z) > 0x1p20f;
    r.fast = bounds_tame && !(dx || dy || dz || big) && tame(o.x) && tame(o.y) && tame(o.z);
    return r;
}

// RN(a / b) from y = RN(1 / b): one product and two FMA corrections give
// the correctly rounded quotient (Markstein's theorem) whenever the
// remainder cannot underflow, i.e. for |a| >= 2^-96 (checked against IEEE
// division on 8e8 random pairs) or a == 0. When box bounds and origin
// components are all 0 or at least 2^-72, every nonzero lo - o is at least
// 2^-95, so the fast form needs no guard.
DEV float qdiv_fast(float a, float b, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-q, b, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, y, q);
}
DEV float qdiv(float a, float b, float y) {
    if (fabsf(a) < 0x1p-96f) return a / b;
    return qdiv_fast(a, b, y);
}

// The reference's box_hit (trace.wgsl:453-468) with the same quotients. The
// early exit after each axis is folded into one final test (the entry
// distance only grows and the exit distance only shrinks, so both forms
// agree). `nr` is the entry distance: a deferred child passes its pop-time
// test exactly when nr <= the current t.
DEV bool slab(float lx, float ly, float lz, float hx, float hy, float hz, const Ray& R, float limit, float& nr) {
    float near_ = 0.0f, far_ = limit;
    bool inside = true;
    if (fabsf(R.d.x) < 1e-12f) {
        inside = !(R.o.x < lx || R.o.x > hx);
    } else {
        float a = qdiv(lx - R.o.x, R.d.x, R.y.x), b = qdiv(hx - R.o.x, R.d.x, R.y.x);
        near_ = fmaxf(near_, fminf(a, b));
        far_ = fminf(far_, fmaxf(a, b));
    }
    if (fabsf(R.d.y) < 1e-12f) {
        inside = inside && !(R.o.y < ly || R.o.y > hy);
    } else {
        float a = qdiv(ly - R.o.y, R.d.y, R.y.y), b = qdiv(hy - R.o.y, R.d.y, R.y.y);
        near_ = fmaxf(near_, fminf(a, b));
        far_ = fminf(far_, fmaxf(a, b));
    }
    if (fabsf(R.d.z) < 1e-12f) {
        inside = inside && !(R.o.z < lz || R.o.z > hz);
    } else {
        float a = qdiv(lz - R.o.z, R.d.z, R.y.z), b = qdiv(hz - R.o.z, R.d.z, R.y.z);
        near_ = fmaxf(near_, fminf(a, b));
        far_ = fminf(far_, fmaxf(a, b));
    }
    nr = near_;
    return inside && !(near_ > far_);
}
// The same test for a fast ray: straight-line code, no branches.
DEV bool slab_fast(float lx, float ly, float lz, float hx, float hy, float hz, const Ray& R, float limit, float& nr) {
    float ax = qdiv_fast(lx - R.o.x, R.d.x, R.y.x), bx = qdiv_fast(hx - R.o.x, R.d.x, R.y.x);
    float ay = qdiv_fast(ly - R.o.y, R.d.y, R.y.y), by = qdiv_fast(hy - R.o.y, R.d.y, R.y.y);
    float az = qdiv_fast(lz - R.o.z, R.d.z, R.y.z), bz = qdiv_fast(hz - R.o.z, R.d.z, R.y.z);
    float near_ = fmaxf(fmaxf(fmaxf(0.0f, fminf(ax, bx)), fminf(ay, by)), fminf(az, bz));
    float far_ = fminf(fminf(fminf(limit, fmaxf(ax, bx)), fmaxf(ay, by)), fmaxf(az, bz));
    nr = near_;
    return !(near_ > far_);
}
// The conservative test of the default walk for a fast ray: quotients from
// the reciprocal, q' = RN(RN(lo - o) * RN(1 / d)), which lies within 3
// units of 2^-24 (relative) of the correctly rounded RN(RN(lo - o) / d) the
// exact test uses. Entry and exit distances are the same min/max of these,
// so each lies within that bound of its exact value too (min and max commute
// with the monotone bounds x -> x -+ eps|x|), and with eps = 2^-21 (8 units,
// enough to cover the rounding of the bounds themselves) the test accepts
// every box the exact test accepts: near' (1 - eps) <= far' + eps |far'|.
// `nr` is near' (1 - eps), a lower bound of the exact entry distance, so the
// pop-time test nr <= t defers no child the exact walk would take. A box
// accepted beyond the exact test is an inner node whose children the exact
// entry distance already culls, or a leaf the ray passes within rounding of
// its faces.
DEV bool slab_wide(float lx, float ly, float lz, float hx, float hy, float hz, const Ray& R, float limit, float& nr) {
    float ax = (lx - R.o.x) * R.y.x, bx = (hx - R.o.x) * R.y.x;
    float ay = (ly - R.o.y) * R.y.y, by = (hy - R.o.y) * R.y.y;
    float az = (lz - R.o.z) * R.y.z, bz = (hz - R.o.z) * R.y.z;
    float near_ = fmaxf(fmaxf(fmaxf(0.0f, fminf(ax, bx)), fminf(ay, by)), fminf(az, bz));
    float far_ = fminf(fminf(fminf(limit, fmaxf(ax, bx)), fmaxf(ay, by)), fmaxf(az, bz));
    constexpr float EPS = 0x1p-21f;
    nr = near_ * (1.0f - EPS);
    return !(nr > __builtin_fmaf(EPS, fabsf(far_), far_));
}

// trace.wgsl:449-452 to_object: rows . (p, w), summed x, y, z, w
DEV v3 to_object(const float* r, v3 p, float w) {
    return mk(r[0] * p.x + r[1] * p.y + r[2] * p.z + r[3] * w, r[4] * p.x + r[5] * p.y + r[6] * p.z + r[7] * w,
              r[8] * p.x + r[9] * p.y + r[10] * p.z + r[11] * w);
}

struct Stack {
    uint2* lds;    // this lane's column in the block's LDS stack
    uint2* spill;  // this lane's column in the global overflow area
    uint32_t spill_stride;
    int top;
    DEV void push(uint32_t ref, float nr) {
        uint2 e = make_uint2(ref, fbits(nr));
        if (top < STACK_SHORT)
            lds[top * TRACE_BLOCK] = e;
        else
            spill[(size_t)(top - STACK_SHORT) * spill_stride] = e;
        ++top;
    }
    DEV uint2 pop() {
        --top;
        return top < STACK_SHORT ? lds[top * TRACE_BLOCK] : spill[(size_t)(top - STACK_SHORT) * spill_stride];
    }
};

// A closest hit: distance, barycentrics, triangle slot and the top-level
// group it lies in (0 in a scene without instances). Groups are numbered in
// the depth-first order of the top-level tree.
struct Closest {
    float t, u, v;
    uint32_t slot, grp;
};

// Children whose boxes a visit of node `index` tests: the binary node
// visits of the reference's walk that the 4-wide visit stands for.
template <class G>
DEV uint32_t child_boxes(const G& S, uint32_t index) {
    uint4 rf = reinterpret_cast<const uint4*>(S.nodes + index)[6];
    return 1u + (rf.y != REF_NONE) + (rf.z != REF_NONE) + (rf.w != REF_NONE);
}

// Sorts two (key, ref) pairs by key.
DEV void order2(uint32_t& ka, uint32_t& ra, uint32_t& kb, uint32_t& rb) {
    bool swap = kb < ka;
    uint32_t k = swap ? kb : ka, r = swap ? rb : ra;
    kb = swap ? ka : kb;
    rb = swap ? ra : rb;
    ka = k;
    ra = r;
}

// Test the four children of a Node4. With NEAR (closest-hit walks by
// default) continue with the nearest hit child (smallest slab entry
// distance) and push the other hit children farthest first, so they pop
// nearest first; otherwise (any-hit walks, and closest-hit walks of a scene
// built with the left-first order) continue with the first hit child and
// push the others right to left, the reference's left-first order
// (trace.wgsl:1049-1054). Returns false when no child was hit.
// Left first, the walk tests the reference's triangles in the reference's
// order: a child box is tested at its parent's visit, and its deferred
// entry distance is compared with the best hit when it is popped, which is
// the reference's pop-time box test (the intermediate binary node's box
// contains its children's, so its own test never decides anything).
// Nearest first (with slab_wide), the walk still tests every triangle whose
// box is entered within the best hit so far, and ties at one distance go to
// the triangle the left-first walk meets first (Walk::closer); the result
// can differ from the reference's only where the reference's own answer
// depends on its order: a hit within an ulp of a box face, where a box's
// entry distance and the triangle's distance round apart (measured on rays
// aimed at the vertices and edges of a closed icosphere: 171 of 46,080
// Moller-Trumbore and 109 watertight rays keep another triangle, each
// within 4 ulps of the oracle's distance). Nearest first cuts the closest-hit walk on the 4K
// terrain from 25.0 to 20.3 node visits per ray; shadow rays start on a
// surface whose own boxes are nearest, and walk fewer nodes left first
// (22.6 against 23.2), which is also where any-hit may stop.
// WIDE: a fast ray tests boxes with slab_wide (the default walks), else with
// the exact quotients (left-first walks, the reference's arithmetic).
template <bool NEAR>
DEV bool choose4(bool h0, bool h1, bool h2, bool h3, float n0, float n1, float n2, float n3, uint4 rf, Stack& st,
                 uint32_t& cur);
template <bool NEAR, bool WIDE, class G>
DEV bool visit4(const G& S, uint32_t index, const Ray& R, float t, Stack& st, uint32_t& cur) {
    const float4* np = reinterpret_cast<const float4*>(S.nodes + index);
    float4 lx = np[0], ly = np[1], lz = np[2], hx = np[3], hy = np[4], hz = np[5];
    uint4 rf = reinterpret_cast<const uint4*>(np)[6];
    float n0, n1, n2, n3;
    bool h0, h1, h2, h3;
    if (WIDE && R.fast) {
        h0 = slab_wide(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, R, t, n0);
        h1 = slab_wide(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, R, t, n1) && rf.y != REF_NONE;
        h2 = slab_wide(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, R, t, n2) && rf.z != REF_NONE;
        h3 = slab_wide(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, R, t, n3) && rf.w != REF_NONE;
    } else if (R.fast) {
        h0 = slab_fast(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, R, t, n0);
        h1 = slab_fast(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, R, t, n1) && rf.y != REF_NONE;
        h2 = slab_fast(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, R, t, n2) && rf.z != REF_NONE;
        h3 = slab_fast(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, R, t, n3) && rf.w != REF_NONE;
    } else {
        h0 = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, R, t, n0);
        h1 = rf.y != REF_NONE && slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, R, t, n1);
        h2 = rf.z != REF_NONE && slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, R, t, n2);
        h3 = rf.w != REF_NONE && slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, R, t, n3);
    }
    return choose4<NEAR>(h0, h1, h2, h3, n0, n1, n2, n3, rf, st, cur);
}
// The children a visit continues with and defers, from their box tests (hit
// flags and entry distances) -- see visit4.
template <bool NEAR>
DEV bool choose4(bool h0, bool h1, bool h2, bool h3, float n0, float n1, float n2, float n3, uint4 rf, Stack& st,
                 uint32_t& cur) {
    if (!NEAR) {
        if (h3 && (h0 || h1 || h2)) st.push(rf.w, n3);
        if (h2 && (h0 || h1)) st.push(rf.z, n2);
        if (h1 && h0) st.push(rf.y, n1);
        cur = h0 ? rf.x : (h1 ? rf.y : (h2 ? rf.z : rf.w));
        return h0 || h1 || h2 || h3;
    }
    // entry distances are >= 0, so their bits order as integers; misses last
    const uint32_t MISSED = 0xffffffffu;
    uint32_t k0 = h0 ? fbits(n0) : MISSED, k1 = h1 ? fbits(n1) : MISSED;
    uint32_t k2 = h2 ? fbits(n2) : MISSED, k3 = h3 ? fbits(n3) : MISSED;
    uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
    order2(k0, r0, k1, r1);
    order2(k2, r2, k3, r3);
    order2(k0, r0, k2, r2);
    order2(k1, r1, k3, r3);
    order2(k1, r1, k2, r2);
    if (k3 != MISSED) st.push(r3, ubits(k3));
    if (k2 != MISSED) st.push(r2, ubits(k2));
    if (k1 != MISSED) st.push(r1, ubits(k1));
    cur = r0;
    return k0 != MISSED;
}

// A plane value of a QNode4: byte c of the plane's word, decoded.
DEV float dq(uint32_t word, uint32_t c, float scale, float origin) {
    return __builtin_fmaf((float)((word >> (8u * c)) & 0xffu), scale, origin);
}
// The visit of the default walks on a scene with quantized nodes: each
// child's box is decoded from its bytes just before its test (a box that
// contains the exact one: the test stays conservative, the walk keeps every
// triangle the exact walk would test) and tested as visit4 tests a fast
// ray's boxes, or with the exact quotients for another ray. The node takes
// four 16-byte loads and 16 registers, against seven and 28 for a Node4.
template <bool NEAR, class G>
DEV bool visit4q(const G& S, uint32_t index, const Ray& R, float t, Stack& st, uint32_t& cur) {
    const uint4* np = reinterpret_cast<const uint4*>(S.qnodes + index);
    const uint4 a = np[0], b = np[1], c = np[2], rf = np[3];
    const float ox = __uint_as_float(a.x), oy = __uint_as_float(a.y), oz = __uint_as_float(a.z);
    const float sx = __uint_as_float(a.w), sy = __uint_as_float(b.x), sz = __uint_as_float(b.y);
    // q_lo x, y, z = b.z, b.w, c.x; q_hi x, y, z = c.y, c.z, c.w
    float n[4];
    bool h[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const float lx = dq(b.z, k, sx, ox), ly = dq(b.w, k, sy, oy), lz = dq(c.x, k, sz, oz);
        const float hx = dq(c.y, k, sx, ox), hy = dq(c.z, k, sy, oy), hz = dq(c.w, k, sz, oz);
        h[k] = R.fast ? slab_wide(lx, ly, lz, hx, hy, hz, R, t, n[k]) : slab(lx, ly, lz, hx, hy, hz, R, t, n[k]);
    }
    return choose4<NEAR>(h[0], h[1] && rf.y != REF_NONE, h[2] && rf.z != REF_NONE, h[3] && rf.w != REF_NONE, n[0],
                         n[1], n[2], n[3], rf, st, cur);
}

// The traversal of one ray as a resumable state: closest hit
// (trace.wgsl:999-1075 `scene`, with tri_hit :780-804) or, with ANY, any hit
// (`occluded`, :1096-1168, tri_blocks :1076-1095). Each ray walks the tree
// depth first, left child first, with the same deferred pop-time distance
// test, so it tests the reference's leaves in the reference's order and
// equal-distance ties resolve to the same triangle. Keeping the walk in a
// state lets a wave retire a finished ray and start a new one in its lane
// while the other lanes keep walking.
// The watertight test's per-ray frame: kz is the axis of the largest
// direction component (the first of equals), kx and ky follow it cyclically
// and swap when that component is negative; the shear takes the direction
// to (0, 0, 1) in (kx, ky, kz).
struct Shear {
    uint32_t kx, ky, kz;
    float sx, sy, sz;
};
DEV float pick(v3 v, uint32_t k) { return k == 0u ? v.x : (k == 1u ? v.y : v.z); }
DEV Shear shear_of(v3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    Shear s;
    s.kz = (ax >= ay && ax >= az) ? 0u : (ay >= az ? 1u : 2u);
    s.kx = s.kz == 2u ? 0u : s.kz + 1u;
    s.ky = s.kx == 2u ? 0u : s.kx + 1u;
    float dz = pick(d, s.kz);
    if (dz < 0.0f) {
        uint32_t k = s.kx;
        s.kx = s.ky;
        s.ky = k;
    }
    s.sx = pick(d, s.kx) / dz;
    s.sy = pick(d, s.ky) / dz;
    s.sz = 1.0f / dz;
    return s;
}

template <bool ANY>
struct Walk {
    Ray world, R;
    Closest h;
    float limit;
    uint32_t cur, base, grp, word;
    float inv_det;
    int inst_top;

    // Begin a ray; false when it misses the scene's root box.
    template <class G>
    DEV bool start(const G& S, v3 o, v3 d, float lim, Stack& st) {
        limit = lim;
        h.t = lim;
        h.u = 0.0f;
        h.v = 0.0f;
        h.slot = 0xffffffffu;
        h.grp = 0u;
        world = make_ray(o, d, (S.geo_flags & GEO_TAME) != 0u);
        R = world;
        base = 0;
        grp = 0;
        word = 0;
        inv_det = 1.0f;
        inst_top = -1;
        st.top = 0;
        cur = S.root.ref;
        float nr;
        return S.root.ref != REF_NONE && slab(S.root.lo[0], S.root.lo[1], S.root.lo[2], S.root.hi[0],
                                              S.root.hi[1], S.root.hi[2], world, lim, nr);
    }
    DEV bool inner() const { return (cur & (REF_LEAF | REF_INST)) == 0u; }
    DEV bool leafy() const { return (cur & REF_LEAF) != 0u; }
    // Continue with the next deferred child whose entry distance is within
    // the current hit distance, leaving an instance when its subtree is done;
    // false when none is left.
    DEV bool next(Stack& st) {
        for (;;) {
            if (st.top == inst_top) {
                R = world;
                base = 0;
                grp = 0;
                word = 0;
                inv_det = 1.0f;
                inst_top = -1;
            }
            if (st.top == 0) return false;
            uint2 e = st.pop();
            if (ubits(e.y) <= h.t) {
                cur = e.x;
                return true;
            }
        }
    }
    // Whether a hit at distance t on triangle slot i of the current group
    // replaces the best one: nearer, or as near and met earlier by the
    // reference's left-first walk -- an earlier group of the top-level tree,
    // or in the same group a lower slot (slots follow the bottom-level
    // tree's depth-first leaf order). A tie with the distance limit itself
    // (no hit yet) does not count, as in the reference's strict test.
    DEV bool closer(float t, uint32_t i) const {
        return t < h.t || (t == h.t && h.slot != 0xffffffffu && (grp < h.grp || (grp == h.grp && i < h.slot)));
    }
    // The record of the best hit's group (0 in a scene without instances).
    template <class G>
    DEV uint32_t record(const G& S) const { return S.root.top ? S.groups[h.grp].record : 0u; }
    // Test the current leaf's triangles in order; with ANY, true on the
    // first blocking hit.
    template <bool COUNT, class G>
    DEV bool leaf(const G& S, uint32_t& prims_ctr) {
        bool skip = ANY ? (word & 2u) != 0u : (word & 1u) != 0u;  // casts-no-shadow / shadow-only
        if (skip) return false;
        uint32_t first = base + (cur & REF_FIRST_MASK);
        uint32_t count = ((cur >> 27) & 15u) + 1u;
        if (S.geo_flags & GEO_WATERTIGHT) return leaf_watertight<COUNT>(S, first, count, prims_ctr);
        if (ANY) {
            // two triangles' loads in flight at once (a shadow walk has the
            // registers to spare), tested in order
            for (uint32_t j = 0; j < count; j += 2) {
                uint32_t i = first + j;
                bool two = j + 1u < count;
                const float4* tp = reinterpret_cast<const float4*>(S.tris + i);
                const float4* tq = reinterpret_cast<const float4*>(S.tris + (two ? i + 1u : i));
                float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
                float4 s0 = tq[0], s1 = tq[1], s2 = tq[2];
                if (COUNT) prims_ctr++;
                if (tri_mt(t0, t1, t2, i)) return true;
                if (two) {
                    if (COUNT) prims_ctr++;
                    if (tri_mt(s0, s1, s2, i + 1u)) return true;
                }
            }
            return false;
        }
        for (uint32_t j = 0; j < count; ++j) {
            uint32_t i = first + j;
            if (COUNT) prims_ctr++;
            const float4* tp = reinterpret_cast<const float4*>(S.tris + i);
            tri_mt(tp[0], tp[1], tp[2], i);
        }
        return false;
    }
    // One triangle (edge form a, e1, e2) with the reference's test; with ANY
    // true when it blocks, else the best hit is updated when it is closer.
    DEV bool tri_mt(float4 t0, float4 t1, float4 t2, uint32_t i) {
        v3 A = mk(t0.x, t0.y, t0.z), e1 = mk(t0.w, t1.x, t1.y), e2 = mk(t1.z, t1.w, t2.x);
        v3 p = pfx::cross(R.d, e2);
        float det = dot(e1, p);
        float seen = det * inv_det;
        if (fabsf(seen) < 1e-9f) return false;
        if (!ANY && (word & ONE_SIDED) != 0u && seen < 0.0f) return false;
        float inv = 1.0f / det;
        v3 s = R.o - A;
        float u = dot(s, p) * inv;
        if (u < 0.0f || u > 1.0f) return false;
        v3 q = pfx::cross(s, e1);
        float v = dot(R.d, q) * inv;
        if (v < 0.0f || u + v > 1.0f) return false;
        float t = dot(e2, q) * inv;
        if (ANY) return t > 0.0001f && t < limit;
        if (t > 0.0001f && closer(t, i)) {
            h.t = t;
            h.u = u;
            h.v = v;
            h.slot = i;
            h.grp = grp;
        }
        return false;
    }
    // The same leaf with the watertight test (Woop, Benthin and Wald, "Watertight
    // Ray/Triangle Intersection", JCGT 2(1), 2013): vertices are moved to the
    // ray origin, sheared so the ray runs along +z of a permuted frame (kz the
    // largest direction component, kx/ky swapped when it is negative to keep
    // the winding), and the three 2D edge functions decide the hit. A vertex
    // or edge shared by two triangles gives bit-identical edge functions of
    // opposite sign in both, so a ray can not pass between them. FP32 only:
    // an edge function that rounds to 0 counts as inside (the paper's FP64
    // recheck of that case is not made; a ray can then hit both triangles at
    // the same t, and the first in order is kept, as for any tie).
    template <bool COUNT, class G>
    DEV bool leaf_watertight(const G& S, uint32_t first, uint32_t count, uint32_t& prims_ctr) {
        const Shear sh = shear_of(R.d);
        for (uint32_t j = 0; j < count; ++j) {
            uint32_t i = first + j;
            if (COUNT) prims_ctr++;
            const float4* tp = reinterpret_cast<const float4*>(S.tris + i);
            float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
            v3 A = mk(t0.x, t0.y, t0.z) - R.o, B = mk(t0.w, t1.x, t1.y) - R.o, C = mk(t1.z, t1.w, t2.x) - R.o;
            float az = pick(A, sh.kz), bz = pick(B, sh.kz), cz = pick(C, sh.kz);
            float ax = pick(A, sh.kx) - sh.sx * az, ay = pick(A, sh.ky) - sh.sy * az;
            float bx = pick(B, sh.kx) - sh.sx * bz, by = pick(B, sh.ky) - sh.sy * bz;
            float cx = pick(C, sh.kx) - sh.sx * cz, cy = pick(C, sh.ky) - sh.sy * cz;
            float U = cx * by - cy * bx;
            float V = ax * cy - ay * cx;
            float W = bx * ay - by * ax;
            if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) continue;
            float det = U + V + W;
            if (det == 0.0f) continue;
            if (!ANY && (word & ONE_SIDED) != 0u && det * inv_det < 0.0f) continue;
            float T = U * (sh.sz * az) + V * (sh.sz * bz) + W * (sh.sz * cz);
            float rcp = 1.0f / det;
            float t = T * rcp;
            if (ANY) {
                if (t <= 0.0001f || t >= limit) continue;
                return true;
            }
            if (t > 0.0001f && closer(t, i)) {
                h.t = t;
                h.u = V * rcp;
                h.v = W * rcp;
                h.slot = i;
                h.grp = grp;
            }
        }
        return false;
    }
    // Enter the current instance: move the ray into its object space and
    // continue at its bottom-level root; false when its box is missed.
    template <class G>
    DEV bool enter(const G& S, Stack& st) {
        grp = cur & REF_INDEX_MASK;
        const Group& g = S.groups[grp];
        base = g.base;
        word = g.word;
        inv_det = g.inv_det;
        if (g.has_xform)
            R = make_ray(to_object(g.inv, world.o, 1.0f), to_object(g.inv, world.d, 0.0f), (S.geo_flags & GEO_TAME) != 0u);
        inst_top = st.top;
        float nr;
        if (slab(g.lo[0], g.lo[1], g.lo[2], g.hi[0], g.hi[1], g.hi[2], R, h.t, nr)) {
            cur = g.root;
            return true;
        }
        return false;
    }
    // One step: an inner node, a leaf or an instance, then the next deferred
    // child if needed. Returns false when the walk is over; `blocked` is
    // set when an any-hit walk found an occluder.
    template <bool COUNT, class G>
    DEV bool step(const G& S, Stack& st, bool& blocked, uint32_t& nodes_ctr, uint32_t& prims_ctr) {
        if (inner()) {
            if (COUNT) nodes_ctr++;
            if ((S.geo_flags & GEO_LEFT_FIRST) != 0u) {
                if (visit4<false, false>(S, cur, R, h.t, st, cur)) return true;
            } else if (S.qnodes) {
                if (visit4q<!ANY>(S, cur, R, h.t, st, cur)) return true;
            } else if (ANY) {
                if (visit4<false, true>(S, cur, R, h.t, st, cur)) return true;
            } else if (visit4<true, true>(S, cur, R, h.t, st, cur)) {
                return true;
            }
        } else if (leafy()) {
            if (leaf<COUNT, G>(S, prims_ctr)) {
                blocked = true;
                return false;
            }
        } else if (enter(S, st)) {
            return true;
        }
        return next(st);
    }
};

// Whole walks for one ray (used where lanes do not refill).
template <class G>
DEV Closest trace_closest(const G& S, v3 o, v3 d, float tmax, Stack& st) {
    Walk<false> w;
    uint32_t nodes = 0, prims = 0;
    bool blocked = false;
    if (w.start(S, o, d, tmax, st))
        while (w.template step<false, G>(S, st, blocked, nodes, prims)) {
        }
    return w.h;
}

// What a walk reads of a scene. In a batch each lane walks its own still's
// scene, so a lane keeps this much of it in registers; a single trace walks
// with the uniform SceneDev itself.
struct Geo {
    const Node4* nodes;
    const QNode4* qnodes;
    const TriPos* tris;
    const Group* groups;
    RootEntry root;
    uint32_t geo_flags;
};
DEV Geo geo_of(const SceneDev& S) { return Geo{S.nodes, S.qnodes, S.tris, S.groups, S.root, S.geo_flags}; }
template <bool BATCH>
DEV decltype(auto) walk_scene(const SceneDev& S, const Geo& g) {
    if constexpr (BATCH)
        return (g);
    else
        return (S);
}

// Persistent waves refill finished lanes. The queue [0, n) is split into
// PARTS contiguous parts with a counter each (on its own slot); a wave
// starts on the part of its XCD (workgroups go to the XCDs round-robin) and
// moves to the next part once its part is used up. Once at least `refill`
// lanes of a wave are idle, the wave claims that many positions of its part
// with one atomic and hands them out in lane order.
struct Fetch {
    uint32_t part;  // wave-uniform
    uint32_t hops;  // parts found used up; PARTS means the queue is drained
};
DEV Fetch fetch_start() { return Fetch{blockIdx.x % PARTS, 0u}; }
DEV bool fetch_drained(const Fetch& f) { return f.hops >= PARTS; }

// Returns the queue position for an idle lane that got one, else
// 0xffffffff. Every lane of the wave must call it.
DEV uint32_t refill_lanes(bool idle, uint32_t refill, uint32_t n, uint32_t* counters, Fetch& f) {
    unsigned long long mask = __ballot(idle);
    uint32_t nidle = (uint32_t)__popcll(mask);
    if (fetch_drained(f) || nidle < refill || nidle == 0u) return 0xffffffffu;
    uint32_t lane = __lane_id();
    unsigned long long below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
    uint32_t rank = (uint32_t)__popcll(below);
    for (;;) {
        uint32_t lo = (uint32_t)((uint64_t)n * f.part / PARTS);
        uint32_t hi = (uint32_t)((uint64_t)n * (f.part + 1u) / PARTS);
        uint32_t first = 0;
        if (lane == 0) first = atomicAdd(counters + f.part * SLOT, nidle);
        first = (uint32_t)__shfl((int)first, 0);
        if (first < hi - lo) {
            uint32_t k = lo + first + rank;
            return (idle && k < hi) ? k : 0xffffffffu;
        }
        f.part = (f.part + 1u) % PARTS;
        if (++f.hops >= PARTS) return 0xffffffffu;
    }
}

// A wave runs one phase per round: inner-node steps when the lanes waiting on
// inner nodes are at least two thirds as many as those waiting on leaves or
// instances, else the leaf phase. Every lane calls it (wave-uniform result).
DEV bool inner_round(bool in, bool out) {
    uint32_t ni = (uint32_t)__popcll(__ballot(in)), nl = (uint32_t)__popcll(__ballot(out));
    return ni * 3u >= nl * 2u;
}

// ---- surface at a hit (trace.wgsl:754-779 tri_fill, :681-694 world_hit) -----
struct Surf {
    v3 n, geometric;
    uint32_t material;
};
DEV Surf surface_at(const SceneDev& S, uint32_t slot, uint32_t rec, float u, float v) {
    const float4* tp = reinterpret_cast<const float4*>(S.tris + slot);
    float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
    v3 e1 = mk(t0.w, t1.x, t1.y), e2 = mk(t1.z, t1.w, t2.x);
    if (S.geo_flags & GEO_WATERTIGHT) {  // vertex form: the host's edge subtractions, here
        v3 a = mk(t0.x, t0.y, t0.z);
        e1 = e1 - a;
        e2 = e2 - a;
    }
    const float4* ap = reinterpret_cast<const float4*>(S.attrs + slot);
    float4 a0 = ap[0], a1 = ap[1], a2 = ap[2];
    Surf s;
    v3 geometric = normalize(pfx::cross(e1, e2));
    float w = (1.0f - u) - v;
    v3 shaded = f3(a0) * w + f3(a1) * u + f3(a2) * v;
    v3 n;
    if (dot(shaded, shaded) > 1e-12f)
        n = normalize(shaded);
    else
        n = rec > 0u ? mk(0.0f, 0.0f, 0.0f) : geometric;
    const Surface& sf = S.surfaces[rec];
    s.material = sf.material == NO_MATERIAL ? fbits(a0.w) : sf.material;
    if (rec == 0u) {
        s.geometric = geometric;
        s.n = n;
        return s;
    }
    float mirror = sf.inv_det < 0.0f ? -1.0f : 1.0f;
    v3 i0 = mk(sf.inv[0], sf.inv[1], sf.inv[2]), i1 = mk(sf.inv[4], sf.inv[5], sf.inv[6]),
       i2 = mk(sf.inv[8], sf.inv[9], sf.inv[10]);
    s.geometric = normalize(i0 * geometric.x + i1 * geometric.y + i2 * geometric.z) * mirror;
    v3 sh = i0 * n.x + i1 * n.y + i2 * n.z;
    s.n = pfx::all_zero(n) ? s.geometric : normalize(sh);
    return s;
}

// ---- BRDF (crates/materials/src/wgsl/brdf.wgsl) -------------------------------
struct Mat {
    v3 base;
    float roughness, metalness, specular, clearcoat, clearcoat_roughness, sheen, fresnel_power;
    v3 emission;
};
DEV Mat load_mat(const SceneDev& S, uint32_t m) {
    const float4* p = reinterpret_cast<const float4*>(S.materials + m);
    float4 a = p[0], b = p[1], c = p[2], e = p[3];
    return Mat{mk(a.x, a.y, a.z), a.w, b.x, b.y, b.z, b.w, c.x, c.y, mk(e.x, e.y, e.z)};
}
DEV v3 mix3(v3 a, v3 b, float t) { return a * (1.0f - t) + b * t; }
DEV v3 schlick(v3 f0, float cos_theta, float power) {
    float k = fm_pow(1.0f - fm_clamp(cos_theta, 0.0f, 1.0f), power);
    return mix3(f0, mk(1, 1, 1), k);
}
DEV float ggx_d(float nh, float a2) {
    float n = fmaxf(nh, 0.0f);
    float k = n * n * (a2 - 1.0f) + 1.0f;
    return a2 / (PI * k * k);
}
DEV float smith_g1(float ndot, float a2) {
    float n = fmaxf(ndot, 0.0f);
    return 2.0f * n / (n + sqrtf(a2 + (1.0f - a2) * n * n));
}
DEV float alpha_rough(float r) {
    float rough = fm_clamp(r, 0.02f, 1.0f);
    return rough * rough;
}
DEV float fresnel_exponent(const Mat& s) { return s.fresnel_power <= 0.0f ? 5.0f : s.fresnel_power; }
DEV v3 reflection(const Mat& s, float ndh_cos) {
    float power = fresnel_exponent(s);
    v3 f0 = mix3(mk(s.specular, s.specular, s.specular), s.base, fm_clamp(s.metalness, 0.0f, 1.0f));
    if (f0.x <= 0.0f && f0.y <= 0.0f && f0.z <= 0.0f) return mk(0, 0, 0);
    return schlick(f0, ndh_cos, power);
}
DEV v3 ggx_spec(float ndv, float ndl, float ndh, float a2, v3 fresnel) {
    return fresnel * ggx_d(ndh, a2) * (smith_g1(ndv, a2) * smith_g1(ndl, a2)) / (4.0f * ndv * ndl);
}
DEV v3 eval(const Mat& s, v3 wo, v3 wi) {
    float ndv = wo.z;
    if (ndv <= 1e-5f) return mk(0, 0, 0);
    float ndl = wi.z;
    if (fabsf(ndl) <= 1e-8f) return mk(0, 0, 0);
    float metal = fm_clamp(s.metalness, 0.0f, 1.0f);
    float sheen = fm_clamp(s.sheen, 0.0f, 1.0f);
    float sheen_cover = sheen * fm_pow(1.0f - ndv, 4.0f);
    float view = fm_clamp(ndv, 0.0f, 1.0f);
    if (ndl < 0.0f) return mk(0, 0, 0);
    v3 h = normalize(wo + wi);
    float ndh = fmaxf(h.z, 0.0f);
    float vdh = fmaxf(dot(wo, h), 0.0f);
    v3 fresnel = reflection(s, vdh);
    float a = alpha_rough(s.roughness);
    v3 spec = ggx_spec(ndv, ndl, ndh, a * a, fresnel);
    v3 transmitted = mk(1, 1, 1) - fresnel;
    float body = (1.0f - metal) * 1.0f * 1.0f;
    float diffuse_w = body * (1.0f - sheen_cover);
    float sheen_w = body * sheen_cover * fm_pow(1.0f - ndh, 5.0f);
    v3 colour = spec + s.base * transmitted * (diffuse_w + sheen_w) / PI;
    float coat = fm_clamp(s.clearcoat, 0.0f, 1.0f);
    if (coat > 0.0f) {
        float coat_a = alpha_rough(s.clearcoat_roughness);
        v3 coat_f = schlick(mk(0.04f, 0.04f, 0.04f), vdh, fresnel_exponent(s));
        v3 coat_spec = ggx_spec(ndv, ndl, ndh, coat_a * coat_a, coat_f);
        float fc = schlick(mk(0.04f, 0.04f, 0.04f), view, fresnel_exponent(s)).x;
        colour = coat_spec * coat + colour * (1.0f - coat * fc);
    }
    return colour;
}
DEV float coat_weight(const Mat& s, v3 wo) {
    return fm_clamp(s.clearcoat * (0.12f + 0.88f * (0.04f + 0.96f * fm_pow(1.0f - fmaxf(wo.z, 0.0f), 5.0f))), 0.0f,
                    0.9f);
}
DEV float bsdf_pdf(const Mat& s, v3 wo, v3 wi) {  // trace.wgsl:1774-1788
    if (wi.z <= 0.0f) return 0.0f;
    float coat = coat_weight(s, wo);
    float spec = (1.0f - coat) * fm_clamp(s.metalness + s.specular, 0.0f, 0.9f);
    v3 h = normalize(wo + wi);
    float a = alpha_rough(s.roughness);
    float ca = alpha_rough(s.clearcoat_roughness);
    float denom = fmaxf(4.0f * fabsf(dot(wi, h)), 1e-8f);
    float hz = fmaxf(h.z, 0.0f);
    float base_pdf = ggx_d(hz, a * a) * hz / denom;
    float coat_pdf = ggx_d(hz, ca * ca) * hz / denom;
    return 1.0f * (coat * coat_pdf + spec * base_pdf + (1.0f - coat - spec) * wi.z / PI);
}
DEV v3 reflect(v3 e1, v3 e2) { return e1 - e2 * (2.0f * dot(e2, e1)); }

// ---- lights (crates/trace/src/lights.wgsl) --------------------------------------
// A path's random numbers (trace.wgsl:423-446, sobol.wgsl:44-55): with the
// PCG sampler (`table` null) every draw is the next PCG value, in the
// reference's order; with the Sobol sampler `at(d)` takes the point of
// dimension base + d for the path's pixel seed and sample index, and a
// draw is one of its components (rnd(state, u.x) and its kind).
struct Draw {
    uint32_t state;      // PCG state, or the Sobol pixel seed (rng.x)
    uint32_t index;      // Sobol sample index (rng.y)
    uint32_t base;       // Sobol dimension of the bounce (rng.z = 16 (bounce + 1))
    const float* table;  // the Sobol patterns, or null for PCG
    uint32_t length;
    pfx::Sobol4 u;
    DEV void at(uint32_t dim) {
        if (table) u = pfx::sobol_point(table, length, state, index, base + dim);
    }
    DEV float x() { return table ? u.x : pfx::rnd(&state); }
    DEV float y() { return table ? u.y : pfx::rnd(&state); }
    DEV float z() { return table ? u.z : pfx::rnd(&state); }
    DEV float w() { return table ? u.w : pfx::rnd(&state); }
};
// The Sobol dimension of NEE slot k (trace.wgsl:442-446)
DEV uint32_t light_dim(uint32_t k) { return k == 0u ? 1u : (k < 13u ? 3u + k : k << 16); }

struct Cand {
    v3 wi, scale;
    float pdf, roughness;
    v3 direction;
    float limit;
};
DEV Cand none() { return Cand{mk(0, 0, 1), mk(0, 0, 0), 0.0f, 0.0f, mk(0, 0, 1), -1.0f}; }
DEV float mis_power(float chosen, float other) {
    if (chosen <= 0.0f) return 0.0f;
    float ratio = other / chosen;
    return 1.0f / (1.0f + ratio * ratio);
}
DEV float4 row(const float4* rows, uint32_t r, int k) { return rows[r * 4u + (uint32_t)k]; }
DEV void shadow_to(v3 origin, v3 aim, bool shadow, v3& dir, float& limit) {  // lights.wgsl:95-102
    dir = mk(0, 0, 1);
    limit = 0.0f;
    if (!shadow) return;
    v3 ray = aim - origin;
    float distance = pfx::length(ray);
    float l = distance * 0.9999f - 0.0002f;
    if (l <= 0.0f) return;
    dir = ray / distance;
    limit = l;
}
DEV Cand emitter_candidate(const SceneDev& S, const m3& T, v3 p, v3 origin, Draw& dr) {  // :103-144
    uint32_t count = (uint32_t)S.emitters[0].x;
    float uz = dr.z();
    uint32_t low = 0, high = count;
    while (low < high) {
        uint32_t mid = (low + high) / 2u;
        if (S.cdf[mid] <= uz)
            low = mid + 1u;
        else
            high = mid;
    }
    uint32_t index = (low < count - 1u ? low : count - 1u) + 1u;
    float4 ea = row(S.emitters, index, 0), eb = row(S.emitters, index, 1), ec = row(S.emitters, index, 2),
           ee = row(S.emitters, index, 3);
    float r1 = sqrtf(dr.x());
    float r2 = dr.y();
    v3 a = f3(ea), b = f3(eb), c = f3(ec);
    v3 x = a * (1.0f - r1) + b * (r1 * (1.0f - r2)) + c * (r1 * r2);
    v3 to = x - p;
    float distance2 = dot(to, to);
    v3 direction = to / sqrtf(fmaxf(distance2, 1e-30f));
    v3 local = pfx::tmul(T, direction);
    if (local.z <= 0.0f) return none();
    v3 normal = normalize(pfx::cross(b - a, c - a));
    if (fabsf(dot(normal, p - a)) <= 1e-5f) return none();
    float cosine = fabsf(dot(normal, direction));
    if (cosine <= 1e-7f) return none();
    Cand r;
    shadow_to(origin, x, true, r.direction, r.limit);
    r.wi = local;
    r.scale = f3(ee);
    r.pdf = ee.w / ec.w * distance2 / cosine;
    r.roughness = 0.0f;
    return r;
}
// lights.wgsl:57-63: (unit normal, area) of a rect or disc light row
DEV float4 area_frame(float4 la, float4 lc, float4 ld) {
    if ((uint32_t)la.w == 3u) return f4(f3(lc), PI * ld.x * ld.x);
    v3 spanned = pfx::cross(f3(lc), f3(ld)) * 4.0f;
    float area = pfx::length(spanned);
    return f4(spanned / area, area);
}
// lights.wgsl:64-91: radiance of the shadowed rect and disc lights among the
// looped rows that the ray (o, d) meets before `limit`, MIS-weighted
// against light sampling when the ray was BSDF-sampled with previous_pdf.
DEV v3 light_hits(const SceneDev& S, v3 o, v3 d, float limit, v3 previous_point, float previous_pdf) {
    v3 total = mk(0, 0, 0);
    for (uint32_t i = 0; i < S.looped; ++i) {
        float4 la = row(S.lights, i, 0), lb = row(S.lights, i, 1), lc = row(S.lights, i, 2),
               ld = row(S.lights, i, 3);
        uint32_t kind = (uint32_t)la.w;
        if ((kind != 2u && kind != 3u) || ld.w < 0.5f) continue;
        float4 plane = area_frame(la, lc, ld);
        v3 pn = f3(plane);
        float facing = dot(d, pn);
        if (fabsf(facing) < 1e-9f || (lc.w < 0.5f && facing >= 0.0f)) continue;
        float t = dot(f3(la) - o, pn) / facing;
        if (t <= 0.0001f || t >= limit) continue;
        v3 x = o + d * t;
        v3 q = x - f3(la);
        if (kind == 3u) {
            if (dot(q, q) > ld.x * ld.x) continue;
        } else if (fabsf(dot(q, f3(lc))) > dot(f3(lc), f3(lc)) || fabsf(dot(q, f3(ld))) > dot(f3(ld), f3(ld))) {
            continue;
        }
        float weight = 1.0f;
        if (previous_pdf > 0.0f) {
            v3 to = x - previous_point;
            float distance2 = dot(to, to);
            float cosine = fabsf(dot(pn, to)) / sqrtf(fmaxf(distance2, 1e-30f));
            weight = mis_power(previous_pdf, distance2 / (plane.w * fmaxf(cosine, 1e-7f)));
        }
        total = total + f3(lb) * weight;
    }
    return total;
}
DEV Cand light_candidate(const SceneDev& S, uint32_t li, const m3& T, v3 p, v3 origin, v3 geometric,
                         Draw& dr) {  // :145-200
    float4 la = row(S.lights, li, 0), lb = row(S.lights, li, 1), lc = row(S.lights, li, 2), ld = row(S.lights, li, 3);
    bool shadow = ld.w > 0.5f;
    uint32_t kind = (uint32_t)la.w;
    if (kind == 2u || kind == 3u) {  // :149-170: a point on the rect or disc, its area pdf in solid angle
        float u1 = dr.x();
        float u2 = dr.y();
        float4 plane = area_frame(la, lc, ld);
        v3 x = f3(la) + f3(lc) * (2.0f * u1 - 1.0f) + f3(ld) * (2.0f * u2 - 1.0f);
        if (kind == 3u) {
            m3 disc = basis(f3(lc));
            float r = ld.x * sqrtf(u1);
            float phi = 2.0f * PI * u2;
            x = f3(la) + disc.c0 * (r * fm_cos(phi)) + disc.c1 * (r * fm_sin(phi));
        }
        v3 to = x - p;
        float distance2 = dot(to, to);
        v3 direction = to / sqrtf(fmaxf(distance2, 1e-30f));
        v3 local = pfx::tmul(T, direction);
        if (local.z <= 0.0f) return none();
        float facing = -dot(f3(plane), direction);
        float cosine = lc.w > 0.5f ? fabsf(facing) : facing;
        if (cosine <= 1e-7f) return none();
        float pdf = distance2 / (plane.w * cosine);
        Cand r;
        shadow_to(origin, x, shadow, r.direction, r.limit);
        r.wi = local;
        r.roughness = 0.0f;
        if (!shadow) {
            r.scale = f3(lb) / pdf;
            r.pdf = 0.0f;
        } else {
            r.scale = f3(lb);
            r.pdf = pdf;
        }
        return r;
    }
    v3 to = f3(la) - p;
    float distance = pfx::length(to);
    float range = lb.w;
    if (distance >= range || distance <= 1e-6f) return none();
    v3 l = to / distance;
    v3 local = pfx::tmul(T, l);
    if (local.z <= 0.0f || dot(geometric, l) <= 0.0f) return none();
    float cone = 1.0f;
    if ((uint32_t)la.w == 1u) {
        float along = fm_clamp(dot(-l, f3(lc)) * ld.x + ld.y, 0.0f, 1.0f);
        cone = along * along;
        if (cone <= 0.0f) return none();
    }
    float radius = lc.w;
    v3 aim = f3(la);
    if (shadow && radius > 0.0f) {
        m3 disc = basis(l);
        float r = radius * sqrtf(dr.x());
        float phi = 2.0f * PI * dr.y();
        aim = aim + (disc.c0 * (r * fm_cos(phi)) + disc.c1 * (r * fm_sin(phi)));
    }
    float ratio = distance / range;
    float window = fm_clamp(1.0f - ratio * ratio * ratio * ratio, 0.0f, 1.0f);
    float reach = fmaxf(distance, radius);
    Cand r;
    shadow_to(origin, aim, shadow, r.direction, r.limit);
    r.wi = local;
    r.scale = f3(lb) * (window * window * cone / (reach * reach));
    r.pdf = 0.0f;
    r.roughness = sqrtf(fm_clamp(radius / (2.0f * distance), 0.0f, 1.0f));
    return r;
}
DEV float sin_from_cos(float c) { return sqrtf(fmaxf(1.0f - c * c, 0.0f)); }
// `cone` and `angles` are the row's light_cones entries: cos θo, sin θo,
// cos(θo + θe), cos(θe − θo) and θo, θe, computed on the host with the same
// functions this code would call.
DEV float2 tree_importance(v3 n, v3 to, float cos_u, v3 axis, float4 cone, float4 angles, float far_, float near_,
                           float energy) {  // :204-237
    const float theta_o = angles.x, theta_e = angles.y;
    float sin_u = sin_from_cos(cos_u);
    float cos_i = dot(to, n);
    float sin_i = sin_from_cos(cos_i);
    float cos_min_i = cos_i >= cos_u ? 1.0f : cos_i * cos_u + sin_i * sin_u;
    if (cos_min_i < 0.0f) return make_float2(0.0f, 0.0f);
    float cos_max_i = fmaxf(cos_i * cos_u - sin_i * sin_u, 0.0f);
    float cos_t = 1.0f, sin_t = 0.0f;
    v3 nt = -to;
    if (axis.x != nt.x || axis.y != nt.y || axis.z != nt.z) {
        cos_t = dot(axis, nt);
        sin_t = sin_from_cos(cos_t);
    }
    float cos_t_minus_u = cos_t * cos_u + sin_t * sin_u;
    float cos_o = cone.x;
    float sin_o = cone.y;
    float cos_min_out;
    if (cos_t >= cos_u || cos_t_minus_u >= cos_o) {
        cos_min_out = 1.0f;
    } else if (theta_o + theta_e > PI || cos_t_minus_u > cone.z) {
        cos_min_out = cos_t_minus_u * cos_o + sin_from_cos(cos_t_minus_u) * sin_o;
    } else {
        return make_float2(0.0f, 0.0f);
    }
    float most = fabsf(cos_min_i * energy * cos_min_out / (near_ * near_));
    float cos_t_plus_u = cos_t * cos_u - sin_t * sin_u;
    if (theta_e - theta_o < 0.0f || cos_t < 0.0f || cos_u < 0.0f || cos_t_plus_u < cone.w)
        return make_float2(most, 0.0f);
    float cos_max_out = cos_t_plus_u * cos_o - sin_from_cos(cos_t_plus_u) * sin_o;
    return make_float2(most, fabsf(cos_max_i * energy * cos_max_out / (far_ * far_)));
}
DEV float2 node_importance(const SceneDev& S, uint32_t r, float4 a, float4 b, float4 c, v3 p, v3 n) {  // :238-250
    v3 lo = f3(a), hi = f3(b);
    v3 centroid = (lo + hi) * 0.5f;
    v3 extent = hi - centroid;
    if (dot(n, centroid - p) + dot(pfx::vabs(n), pfx::vabs(extent)) <= 0.0f) return make_float2(0.0f, 0.0f);
    if (pfx::length(p - pfx::vmin(pfx::vmax(p, lo), hi)) >= b.w) return make_float2(0.0f, 0.0f);
    v3 to = centroid - p;
    float reach = pfx::length(to);
    float radius2 = dot(extent, extent);
    float cos_u = -1.0f;
    if (reach * reach > radius2) cos_u = sqrtf(fmaxf(1.0f - radius2 / (reach * reach), 0.0f));
    float distance = fmaxf(fmaxf(0.5f * sqrtf(radius2), reach), 1e-6f);
    return tree_importance(n, to / fmaxf(reach, 1e-30f), cos_u, f3(c), S.light_cones[2u * r], S.light_cones[2u * r + 1u],
                           distance, distance, a.w);
}
DEV float2 emitter_importance(const SceneDev& S, uint32_t li, v3 p, v3 n) {  // :255-271
    float4 la = row(S.lights, li, 0), lb = row(S.lights, li, 1), lc = row(S.lights, li, 2), ld = row(S.lights, li, 3);
    v3 to = f3(la) - p;
    float near_ = pfx::length(to);
    if (near_ >= lb.w || near_ <= 1e-6f) return make_float2(0.0f, 0.0f);
    v3 toward = to / near_;
    float radius = lc.w;
    float far_ = sqrtf(radius * radius + near_ * near_);
    float cos_u = near_ / far_;
    float energy = (lb.x + lb.y + lb.z) / 3.0f;
    // θo = 0 and θe = acos(-ld.y / ld.x) for a spot, π/2 otherwise: in light_cones
    float4 cone = S.light_cones[2u * li], angles = S.light_cones[2u * li + 1u];
    if ((uint32_t)la.w == 1u) {
        // cos(max(0, acos(cone . -toward) - acos(cos_u))), lights.wgsl:262-264
        float cos_min_out = pfx::fm_cos_acos_gap(fm_clamp(dot(f3(lc), -toward), -1.0f, 1.0f), cos_u);
        float along = fm_clamp(cos_min_out * ld.x + ld.y, 0.0f, 1.0f);
        energy *= along * along;
        return tree_importance(n, toward, cos_u, f3(lc), cone, angles, far_, near_, energy);
    }
    return tree_importance(n, toward, cos_u, -toward, cone, angles, far_, near_, energy);
}
DEV float2 child_importance(const SceneDev& S, uint32_t r, v3 p, v3 n) {  // :272-275
    float4 a = row(S.lights, r, 0), b = row(S.lights, r, 1), c = row(S.lights, r, 2), d = row(S.lights, r, 3);
    if ((uint32_t)d.y == 1u) return emitter_importance(S, (uint32_t)d.z, p, n);
    return node_importance(S, r, a, b, c, p, n);
}
DEV float2 tree_leaf(const SceneDev& S, uint32_t first, uint32_t count, v3 p, v3 n, float start) {  // :276-349
    float rand = start;
    bool sample_max = rand > 0.5f;
    if (count > 1u) rand = rand * 2.0f - (sample_max ? 1.0f : 0.0f);
    int selected = -1;
    float chosen = 0.0f, chosen_other = 0.0f, total = 0.0f, total_other = 0.0f;
    uint32_t has = 0;
    for (uint32_t i = 0; i < count; ++i) {
        int index = (int)(first + i);
        float2 imp = emitter_importance(S, first + i, p, n);
        float weight = sample_max ? imp.x : imp.y;
        float other = sample_max ? imp.y : imp.x;
        if (weight > 0.0f) {
            total += weight;
            if (selected < 0) {
                selected = index;
                chosen = weight;
            } else {
                float thresh = weight / total;
                if (rand <= thresh) {
                    selected = index;
                    chosen = weight;
                    rand = rand / thresh;
                } else {
                    rand = (rand - thresh) / (1.0f - thresh);
                }
                rand = fm_clamp(rand, 0.0f, 1.0f);
            }
        }
        if (selected == index) chosen_other = other;
        total_other += other;
        has |= (imp.x > 0.0f ? 1u : 0u) << i;
    }
    if (has == 0u) return make_float2(0.0f, 0.0f);
    float chosen_max = sample_max ? chosen : chosen_other;
    float total_max = sample_max ? total : total_other;
    float chosen_min = sample_max ? chosen_other : chosen;
    float total_min = sample_max ? total_other : total;
    if (total_min == 0.0f) {
        if (sample_max) {
            chosen_min = 1.0f;
            total_min = (float)__popc(has);
        } else {
            selected = -1;
            float valid = 0.0f;
            uint32_t bits = has;
            for (uint32_t i = 0; i < count; ++i) {
                if ((bits & 1u) != 0u) {
                    valid += 1.0f;
                    if (selected < 0) {
                        selected = (int)(first + i);
                    } else {
                        float thresh = 1.0f / valid;
                        if (rand <= thresh) {
                            selected = (int)(first + i);
                            rand = rand / thresh;
                        } else {
                            rand = (rand - thresh) / (1.0f - thresh);
                        }
                        rand = fm_clamp(rand, 0.0f, 1.0f);
                    }
                }
                bits >>= 1u;
            }
            chosen_min = 1.0f;
            total_min = valid;
            chosen_max = emitter_importance(S, (uint32_t)selected, p, n).x;
        }
    }
    return make_float2((float)selected, 0.5f * (chosen_max / total_max + chosen_min / total_min));
}
DEV float2 tree_pick(const SceneDev& S, v3 p, v3 n, float start) {  // :350-386
    uint32_t at = (uint32_t)(S.emitters[0].y + S.emitters[0].w);
    float rand = start;
    float pdf = 1.0f;
    for (;;) {
        float4 d = row(S.lights, at, 3);
        uint32_t count = (uint32_t)d.y;
        if (count > 0u) {
            float2 leaf = tree_leaf(S, (uint32_t)d.z, count, p, n, rand);
            return make_float2(leaf.x, leaf.y * pdf);
        }
        uint32_t left = at + 1u;
        uint32_t right = (uint32_t)d.z;
        float2 l = child_importance(S, left, p, n);
        float2 r = child_importance(S, right, p, n);
        float total_max = l.x + r.x;
        if (total_max == 0.0f) return make_float2(0.0f, 0.0f);
        float total_min = l.y + r.y;
        float probability_min = 0.5f * ((l.x > 0.0f ? 1.0f : 0.0f) + (r.x == 0.0f ? 1.0f : 0.0f));
        if (total_min > 0.0f) probability_min = l.y / total_min;
        float left_probability = 0.5f * (l.x / total_max + probability_min);
        float right_weight = 1.0f - left_probability;
        at = left;
        if (right_weight > 0.0f) {
            float thresh = right_weight / (left_probability + right_weight);
            if (rand <= thresh) {
                at = right;
                rand = rand / thresh;
            } else {
                rand = (rand - thresh) / (1.0f - thresh);
            }
            rand = fm_clamp(rand, 0.0f, 1.0f);
        }
        pdf *= at == left ? left_probability : 1.0f - left_probability;
    }
}
DEV Cand direct_light(const SceneDev& S, uint32_t slot, const m3& T, v3 p, v3 origin, v3 geometric,
                      Draw& dr) {  // :387-409
    uint32_t looped = (uint32_t)S.emitters[0].y;
    if (slot > looped) {
        float2 picked = tree_pick(S, p, T.c2, dr.z());
        if (!(picked.y > 0.0f)) return none();
        Cand c = light_candidate(S, (uint32_t)picked.x, T, p, origin, geometric, dr);
        c.scale = c.scale / picked.y;
        return c;
    }
    if (slot > 0u) return light_candidate(S, slot - 1u, T, p, origin, geometric, dr);
    if (S.emitters[0].x < 0.5f) return none();
    return emitter_candidate(S, T, p, origin, dr);
}

// ---- sky (crates/trace/src/sky.rs SKY_SAMPLING_WGSL) ---------------------------
DEV v3 env_radiance(const SceneDev& S, v3 direction) {
    if (S.env_black) return mk(0, 0, 0);
    v3 d = normalize(direction);
    float u = pfx::fm_atan2(d.x, -d.z) / 6.28318531f + 0.5f;
    u = u - floorf(u);
    float v = fm_acos(fm_clamp(d.y, -1.0f, 1.0f)) / 3.14159265f;
    float x = u * (float)S.env_w - 0.5f, y = v * (float)S.env_h - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float tx = x - x0, ty = y - y0;
    int w = (int)S.env_w, h = (int)S.env_h;
    int c0 = (((int)x0) % w + w) % w, c1 = (((int)(x0 + 1.0f)) % w + w) % w;
    int r0 = (int)y0, r1 = (int)(y0 + 1.0f);
    r0 = r0 < 0 ? 0 : (r0 > h - 1 ? h - 1 : r0);
    r1 = r1 < 0 ? 0 : (r1 > h - 1 ? h - 1 : r1);
    const float4* t = reinterpret_cast<const float4*>(S.env);
    float4 A = t[r0 * w + c0], B = t[r0 * w + c1], C = t[r1 * w + c0], D = t[r1 * w + c1];
    v3 a = f3(A), b = f3(B), c = f3(C), dd = f3(D);
    return (a * (1.0f - tx) + b * tx) * (1.0f - ty) + (c * (1.0f - tx) + dd * tx) * ty;
}
DEV uint32_t cdf_find(const float* cdf, uint32_t n, float needle) {
    uint32_t low = 0, high = n;
    while (low < high) {
        uint32_t mid = (low + high) / 2u;
        if (cdf[mid] <= needle)
            low = mid + 1u;
        else
            high = mid;
    }
    return low < n - 1u ? low : n - 1u;
}
constexpr uint32_t SKY_W = 256, SKY_H = 128;
DEV void sky_sample(const SceneDev& S, float rx, float ry, v3& direction, float& pdf) {
    float u = fm_clamp(rx, 0.0f, 0.99999994f);
    float v = fm_clamp(ry, 0.0f, 0.99999994f);
    uint32_t rowi = cdf_find(S.sky_rows, SKY_H, v);
    float row_before = rowi > 0u ? S.sky_rows[rowi - 1u] : 0.0f;
    float fy = (v - row_before) / fmaxf(S.sky_rows[rowi] - row_before, 0.000001f);
    float needle = fmaxf(u - 0.0f, 0.0f) / fmaxf(1.0f - 0.0f, 0.000001f);
    const float* cols = S.sky_columns + rowi * SKY_W;
    uint32_t col = cdf_find(cols, SKY_W, needle);
    float col_before = col > 0u ? cols[col - 1u] : 0.0f;
    float fx = (needle - col_before) / fmaxf(cols[col] - col_before, 0.000001f);
    float theta = 3.14159265f * ((float)rowi + fy) / (float)SKY_H;
    float phi = 6.28318531f * ((float)col + fx) / (float)SKY_W;
    direction = mk(fm_sin(theta) * fm_cos(phi), fm_cos(theta), fm_sin(theta) * fm_sin(phi));
    float area = 6.28318531f / (float)SKY_W * 3.14159265f / (float)SKY_H * fmaxf(fm_sin(theta), 0.00001f);
    float mass = (S.sky_rows[rowi] - row_before) * (cols[col] - col_before);
    pdf = (1.0f - 0.0f) * mass / area + 0.0f;
}
DEV float sky_pdf(const SceneDev& S, v3 direction) {
    v3 d = normalize(direction);
    float theta = fm_acos(fm_clamp(d.y, -1.0f, 1.0f));
    float a = pfx::fm_atan2(d.z, d.x) + 6.28318531f;
    float phi = a - 6.28318531f * truncf(a / 6.28318531f);
    float fr = theta / 3.14159265f * (float)SKY_H, fc = phi / 6.28318531f * (float)SKY_W;
    uint32_t rowi = fr >= (float)SKY_H ? SKY_H - 1u : (uint32_t)fmaxf(fr, 0.0f);
    uint32_t col = fc >= (float)SKY_W ? SKY_W - 1u : (uint32_t)fmaxf(fc, 0.0f);
    if (rowi > SKY_H - 1u) rowi = SKY_H - 1u;
    if (col > SKY_W - 1u) col = SKY_W - 1u;
    float row_before = rowi > 0u ? S.sky_rows[rowi - 1u] : 0.0f;
    const float* cols = S.sky_columns + rowi * SKY_W;
    float col_before = col > 0u ? cols[col - 1u] : 0.0f;
    float mass = (S.sky_rows[rowi] - row_before) * (cols[col] - col_before);
    float area = 6.28318531f / (float)SKY_W * 3.14159265f / (float)SKY_H * fmaxf(fm_sin(theta), 0.00001f);
    return (1.0f - 0.0f) * mass / area + 0.0f;
}

// ---- kernels ------------------------------------------------------------------------
// Camera rays (trace.wgsl:2504-2561): one path per (sample, active pixel).
// Orthographic and equirect cameras (trace.wgsl:2546-2555, probe.wgsl:18-26):
// an orthographic ray starts on the view rectangle and runs along forward;
// an equirect ray takes its direction from longitude and latitude.
DEV void project(const SceneDev& S, float ndcx, float ndcy, v3 fwd, v3 right, v3 up, v3& origin, v3& direction) {
    if (S.projection == 1u) {
        direction = normalize(fwd);
        origin = origin + (normalize(right) * ((ndcx + S.lens[0]) * S.half_view[0]) +
                           normalize(up) * ((-ndcy + S.lens[1]) * S.half_view[1]));
    } else {
        float phi = ((ndcx * 0.5f) * 2.0f) * PI;
        float theta = (ndcy * 0.5f + 0.5f) * PI;
        float st = fm_sin(theta);
        v3 local = mk(st * fm_sin(phi), fm_cos(theta), -st * fm_cos(phi));
        direction = normalize(normalize(right) * local.x + normalize(up) * local.y - normalize(fwd) * local.z);
    }
}

// Start path i at pixel `pix` of the still with sample index `sample`.
DEV void camera_path(const SceneDev& S, const WaveDev& W, uint32_t i, uint32_t pix, uint32_t sample, uint32_t seed) {
    uint32_t y = pix / S.width, x = pix - y * S.width;
    uint32_t px = x + S.window[2], py = y + S.window[3];
    uint32_t stream = py * S.window[0] + px;
    Draw dr{pfx::pcg(stream * 1973u + pfx::pcg(seed * 9277u + sample * 7919u + 26699u)), 0u, 0u, S.sobol_table,
            S.sobol_len, {}};
    if (S.sobol_table) {  // trace.wgsl:2522
        dr.state = pfx::hash_iqnt2d(px, py) ^ seed;
        dr.index = sample;
    }
    dr.at(0u);
    float jx = dr.x();
    float jy = dr.y();
    float ndcx = ((float)px + jx) / (float)S.window[0] * 2.0f - 1.0f;
    float ndcy = ((float)py + jy) / (float)S.window[1] * 2.0f - 1.0f;
    v3 fwd = mk(S.forward[0], S.forward[1], S.forward[2]), right = mk(S.right[0], S.right[1], S.right[2]),
       up = mk(S.up[0], S.up[1], S.up[2]);
    v3 direction = normalize(fwd + right * (ndcx + S.lens[0]) + up * (-ndcy + S.lens[1]));
    v3 origin = mk(S.origin[0], S.origin[1], S.origin[2]);
    if (S.projection != 0u) {
        project(S, ndcx, ndcy, fwd, right, up, origin, direction);
    } else if (S.lens[2] > 0.0f) {
        v3 focus = origin + direction * (S.lens[3] / dot(direction, normalize(fwd)));
        dr.at(1u);
        float radius = sqrtf(dr.x()) * S.lens[2];
        float phi = 2.0f * PI * dr.y();
        origin = origin + (normalize(right) * (radius * fm_cos(phi)) + normalize(up) * (radius * fm_sin(phi)));
        direction = normalize(focus - origin);
    }
    W.o[i] = f4(origin, 0.0f);
    W.d[i] = f4(direction, ubits(0u));
    W.thr[i] = make_float4(1.0f, 1.0f, 1.0f, ubits(dr.state));
    W.res[i] = make_float4(0.0f, 0.0f, 0.0f, ubits(dr.index));
    W.alb[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    W.nrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(256) void k_camera(SceneDev S, WaveDev W, const uint32_t* active, uint32_t npix,
                                                uint32_t nsamples, uint32_t acc, uint32_t seed) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix * nsamples) return;
    uint32_t s = i / npix, p = i - s * npix;
    uint32_t pix = active ? active[p] : p;
    camera_path(S, W, i, pix, acc + s, seed);
}

// The unit of a batch wave that holds path (or pixel slot) k: units are in
// increasing order of both.
DEV uint32_t unit_of(const WavePlan& P, uint32_t k, bool by_pixel) {
    uint32_t lo = 0, hi = P.units;  // the last unit whose first is <= k
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        uint32_t first = by_pixel ? P.u[mid].pfirst : P.u[mid].first;
        if (first <= k)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Batch: every path of the current wave, each unit a run of samples of one
// still (sample-major over its pixels, as a single trace lays out a wave).
__global__ __launch_bounds__(256) void k_camera_batch(BatchDev B, WaveDev W) {
    const WavePlan& P = B.waves[*B.wave];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < P.paths; i += gridDim.x * 256u) {
        const Unit& u = P.u[unit_of(P, i, false)];
        uint32_t local = i - u.first, s = local / u.npix, pix = local - s * u.npix;
        W.sid[i] = u.sid;
        camera_path(B.scenes[u.sid], W, i, pix, u.acc + s, u.seed);
    }
}

DEV void store_closest(const WaveDev& W, uint32_t i, const Closest& h, uint32_t record) {
    W.hit[i] = make_float4(h.t, h.u, h.v, ubits(h.slot));
    W.hit_rec[i] = record;
}

// Closest hit of every queued path. Persistent waves refill finished lanes
// from a shared counter, and each round walks inner nodes until every live
// lane stands on a leaf or an instance, then handles those together, so
// lanes diverge as little as the rays allow.
// Six waves per SIMD: the walk is bound by dependent node fetches, and the
// extra resident waves hide more of that latency than the few registers
// spilled at this limit cost.
// In a batch (BATCH) each path belongs to a still of its own (W.sid) whose
// scene is scenes[sid], and the queue length is the one k_roll left in the
// counters.
// NEAR: nearest-first walk (the default); otherwise the reference's
// left-first order (a scene built with PFXH_WALK_LEFT_FIRST).
// What the instrumented closest-hit pass counts per wave round (wave-uniform
// sums, added by lane 0): inner rounds, leaf rounds, lanes stepping in
// each, lanes live over all rounds, and the leaf rounds' triangle-loop
// trips (the most triangles any stepping lane tests).
struct WaveCount {
    unsigned long long v[6] = {0, 0, 0, 0, 0, 0};
    DEV void round(bool inner, bool step, bool live) {
        v[inner ? 0 : 1] += 1;
        v[inner ? 2 : 3] += (unsigned long long)__popcll(__ballot(step));
        v[4] += (unsigned long long)__popcll(__ballot(live));
    }
    DEV void trips(uint32_t n) {
        for (int o = 32; o > 0; o >>= 1) n = max(n, (uint32_t)__shfl_xor((int)n, o));
        v[5] += n;
    }
    DEV void flush(uint32_t* counters) {
        if (__lane_id() != 0u) return;
        for (int k = 0; k < 6; ++k)
            atomicAdd(reinterpret_cast<unsigned long long*>(counters + C_WAVE) + k, v[k]);
    }
};

// Q: the scene's nodes are quantized (QNode4) and the walk is nearest first.
template <bool COUNT, bool BATCH, bool NEAR, bool Q = false>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_extend(SceneDev S, const SceneDev* scenes, WaveDev W, const uint32_t* queue, uint32_t n) {
    __shared__ uint2 lds[STACK_SHORT * TRACE_BLOCK];
    uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    Stack st{lds + threadIdx.x, W.spill + gtid, W.spill_threads, 0};
    uint32_t nodes = 0, prims = 0, boxes = 0;
    WaveCount wc;
    Walk<false> w;
    Geo g{};
    if (BATCH) n = W.counters[C_N];
    uint32_t ray = 0;
    bool live = false;
    Fetch f = fetch_start();
    for (;;) {
        uint32_t k = refill_lanes(!live, W.refill, n, W.counters + C_FETCH, f);
        if (k != 0xffffffffu) {
            ray = queue ? queue[k] : k;
            float4 o4 = W.o[ray], d4 = W.d[ray];
            if (BATCH) g = geo_of(scenes[W.sid[ray]]);
            live = w.start(walk_scene<BATCH>(S, g), f3(o4), f3(d4), 1e20f, st);
            if (!live) store_closest(W, ray, w.h, w.record(walk_scene<BATCH>(S, g)));
        }
        if (!__any(live)) {
            if (fetch_drained(f)) break;
            continue;
        }
        // one phase per round: inner nodes, or leaves and instances, whichever
        // has more lanes waiting (inner steps favoured 3:2)
        bool in = live && w.inner();
        bool out = live && !in;
        bool inner_phase = inner_round(in, out);
        if (COUNT) {
            wc.round(inner_phase, inner_phase ? in : out, live);
            if (!inner_phase) wc.trips(out && w.leafy() ? ((w.cur >> 27) & 15u) + 1u : 0u);
        }
        if (inner_phase) {
            if (in) {
                if (COUNT) {
                    nodes++;
                    boxes += child_boxes(walk_scene<BATCH>(S, g), w.cur);
                }
                bool more;
                if constexpr (Q)
                    more = visit4q<true>(walk_scene<BATCH>(S, g), w.cur, w.R, w.h.t, st, w.cur);
                else
                    more = visit4<NEAR, NEAR>(walk_scene<BATCH>(S, g), w.cur, w.R, w.h.t, st, w.cur);
                if (!more && !w.next(st)) {
                    store_closest(W, ray, w.h, w.record(walk_scene<BATCH>(S, g)));
                    live = false;
                }
            }
        } else if (out) {
            bool go = true;
            if (w.leafy())
                w.template leaf<COUNT>(walk_scene<BATCH>(S, g), prims);
            else
                go = !w.enter(walk_scene<BATCH>(S, g), st);
            if (go && !w.next(st)) {
                store_closest(W, ray, w.h, w.record(walk_scene<BATCH>(S, g)));
                live = false;
            }
        }
    }
    if (COUNT) {
        atomicAdd(reinterpret_cast<unsigned long long*>(W.counters + C_NODES), (unsigned long long)nodes);
        atomicAdd(reinterpret_cast<unsigned long long*>(W.counters + C_PRIMS), (unsigned long long)prims);
        atomicAdd(reinterpret_cast<unsigned long long*>(W.counters + C_BOXES), (unsigned long long)boxes);
        wc.flush(W.counters);
    }
}

// One path vertex (trace.wgsl:2169-2501 loop body).
// SOBOL: the scene uses the Sobol sampler (with PCG the sampler code compiles out)
template <bool SOBOL>
DEV void shade_path(const SceneDev& S, const WaveDev& W, uint32_t i, uint32_t bounce, bool& cont, uint32_t& entries,
                    uint32_t& tested) {
    float4 o4 = W.o[i], d4 = W.d[i], t4 = W.thr[i], r4 = W.res[i], hv = W.hit[i];
    v3 o = f3(o4), d = f3(d4), throughput = f3(t4), result = f3(r4);
    float previous_pdf = o4.w;
    uint32_t lobes = fbits(d4.w);
    Draw dr{fbits(t4.w), fbits(r4.w), 16u * (bounce + 1u), SOBOL ? S.sobol_table : nullptr, S.sobol_len, {}};
    float t = hv.x;
    bool done = false;
    if (S.area) result = result + throughput * light_hits(S, o, d, t, pfx::spawn_point(o, -d, o, 0.0f), previous_pdf);
    if (t >= 1e19f) {
        if (!S.env_black) {
            float pdf = sky_pdf(S, d);
            float weight = (bounce > 0u && previous_pdf > 0.0f) ? previous_pdf / (previous_pdf + pdf) : 1.0f;
            result = result + throughput * env_radiance(S, d) * weight;
        }
        done = true;
    }
    if (!done) {
        uint32_t slot = fbits(hv.w), rec = W.hit_rec[i];
        Surf sf = surface_at(S, slot, rec, hv.y, hv.z);
        v3 p = o + d * t;
        v3 origin_point = pfx::spawn_point(o, -d, o, 0.0f);
        v3 n = dot(sf.n, d) > 0.0f ? -sf.n : sf.n;
        Mat s = load_mat(S, sf.material);
        v3 geometric = dot(sf.geometric, d) > 0.0f ? -sf.geometric : sf.geometric;
        if (dot(n, geometric) < 0.05f) n = normalize(n + geometric * 0.1f);
        s.roughness = fmaxf(s.roughness, 0.05f);
        s.clearcoat_roughness = fmaxf(s.clearcoat_roughness, 0.05f);
        if (bounce == 0u) {
            W.alb[i] = f4(s.base, 1.0f);
            W.nrm[i] = f4(n, 1.0f);
        }
        // emitter_weight (lights.wgsl:38-44)
        if (S.lit) {
            float weight = 1.0f;
            bool emissive = !(s.emission.x <= 0.0f && s.emission.y <= 0.0f && s.emission.z <= 0.0f);
            if (previous_pdf > 0.0f && emissive) {
                uint32_t index = S.attrs[slot].emitter;
                if (index != 0u) {
                    float4 ee = row(S.emitters, index, 3), ec = row(S.emitters, index, 2);
                    v3 to = p - origin_point;
                    float distance2 = dot(to, to);
                    float cosine = fabsf(dot(sf.geometric, to)) / sqrtf(fmaxf(distance2, 1e-30f));
                    float epdf = ee.w / fmaxf(ec.w, 1e-30f) * distance2 / fmaxf(cosine, 1e-7f);
                    weight = mis_power(previous_pdf, epdf);
                }
            }
            s.emission = s.emission * weight;
        }
        result = result + throughput * s.emission;
        if (lobes & LAST_BIT) {
            done = true;
        } else {
            m3 T = basis(n);
            v3 wo = pfx::tmul(T, -d);
            v3 origin = pfx::spawn_point(p, n, o, t);
            uint32_t slots = 2u + S.light_slots;
            for (uint32_t kk = 0; kk < slots; ++kk) {
                v3 local = mk(0, 0, 1), direction = mk(0, 0, 1), strength = mk(0, 0, 0);
                float limit = 1e19f, pdf = 0.0f;
                Mat lit = s;
                dr.at(light_dim(kk));
                if (kk == 0u) {
                    if (S.sun_intensity <= 0.0f) continue;
                    direction = normalize(mk(S.sun_dir[0], S.sun_dir[1], S.sun_dir[2]));
                    if (S.sun_w < 1.0f) {
                        float height = 1.0f - dr.x() * (1.0f - S.sun_w);
                        float radius = sqrtf(fmaxf(0.0f, 1.0f - height * height));
                        float phi = 2.0f * PI * dr.y();
                        m3 orient = basis(direction);
                        direction =
                            normalize(pfx::mul(orient, mk(radius * fm_cos(phi), radius * fm_sin(phi), height)));
                    }
                    local = pfx::tmul(T, direction);
                    if (local.z <= 0.0f) continue;
                    strength = mk(S.sun_color[0], S.sun_color[1], S.sun_color[2]) * (S.sun_intensity * 1.0f);
                } else if (kk == 1u) {
                    float rx = dr.x();
                    float ry = dr.y();
                    if (S.env_black) continue;  // strength 0: the term adds nothing
                    float spdf;
                    sky_sample(S, rx, ry, direction, spdf);
                    local = pfx::tmul(T, direction);
                    if (local.z <= 0.0f) continue;
                    strength = env_radiance(S, direction) * 1.0f;
                    pdf = spdf;
                    if (!(pdf > 0.0f)) continue;
                } else {
                    Cand c = direct_light(S, kk - 2u, T, p, origin, geometric, dr);
                    if (c.limit < 0.0f) continue;
                    local = c.wi;
                    direction = c.direction;
                    strength = c.scale;
                    pdf = c.pdf;
                    limit = c.limit;
                    lit.roughness = fmaxf(s.roughness, c.roughness);
                }
                if (pdf > 0.0f) {
                    float bsdf = bsdf_pdf(s, wo, local);
                    if (kk == 1u) {
                        strength = strength / (pdf + bsdf);
                    } else {
                        float ratio = bsdf / pdf;
                        strength = strength / (pdf * (1.0f + ratio * ratio));
                    }
                }
                v3 add = throughput * eval(lit, wo, local) * 1.0f * local.z * strength;
                if (add.x == 0.0f && add.y == 0.0f && add.z == 0.0f) continue;
                if (entries < S.entries) {
                    size_t e = (size_t)entries * W.capacity + i;
                    W.e_o[e] = f4(origin, limit > 0.0f ? limit : 0.0f);
                    W.e_d[e] = f4(direction, 0.0f);
                    W.e_c[e] = f4(add, 0.0f);
                    if (limit > 0.0f) tested |= 1u << entries;
                    entries++;
                }
            }
            // sample_bsdf (trace.wgsl:1716-1773), diffuse + GGX + coat
            dr.at(3u);
            (void)dr.w();  // transmission pick: never below zero transmission
            float cw = coat_weight(s, wo);
            float sw = (1.0f - cw) * fm_clamp(s.metalness + s.specular, 0.0f, 0.9f);
            float pick = dr.z();
            v3 wi;
            uint32_t lobe;
            if (pick < cw + sw) {
                bool coat = pick < cw;
                float alpha = alpha_rough(coat ? s.clearcoat_roughness : s.roughness);
                float u1 = dr.x();
                float u2 = dr.y();
                float phi = 2.0f * PI * u1;
                float cos_theta = sqrtf((1.0f - u2) / (1.0f + (alpha * alpha - 1.0f) * u2));
                float stt = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
                v3 h = mk(fm_cos(phi) * stt, fm_sin(phi) * stt, cos_theta);
                wi = reflect(-wo, h);
                lobe = 1u;
            } else {
                float u1 = dr.x();
                float u2 = dr.y();
                float phi = 2.0f * PI * u1;
                float r = sqrtf(u2);
                wi = mk(fm_cos(phi) * r, fm_sin(phi) * r, sqrtf(1.0f - u2));
                lobe = 0u;
            }
            float spdf = bsdf_pdf(s, wo, wi);
            uint32_t dif = (lobes & 0xffu) + (lobe == 0u ? 1u : 0u);
            uint32_t glo = ((lobes >> 8) & 0xffu) + (lobe == 1u ? 1u : 0u);
            uint32_t tra = (lobes >> 16) & 0xffu;
            bool last = bounce >= S.total || dif > S.diffuse || glo > S.glossy || tra > S.transmission;
            lobes = (dif & 0xffu) | ((glo & 0xffu) << 8) | ((tra & 0xffu) << 16) | (last ? LAST_BIT : 0u);
            if (wi.z <= 0.0f || spdf <= 0.0f) {
                done = true;
            } else {
                throughput = throughput * (eval(s, wo, wi) * 1.0f * wi.z / spdf * 1.0f);
                d = normalize(pfx::mul(T, wi));
                previous_pdf = spdf;
                o = pfx::spawn_point(p, d, o, t);
                if (bounce >= 3u) {
                    float keep = fm_clamp(fmaxf(throughput.x, fmaxf(throughput.y, throughput.z)), 0.05f, 0.95f);
                    dr.at(0u);
                    if (dr.x() > keep)
                        done = true;
                    else
                        throughput = throughput / keep;
                }
            }
        }
    }
    cont = !done;
    W.o[i] = f4(o, previous_pdf);
    W.d[i] = f4(d, ubits(lobes));
    W.thr[i] = f4(throughput, ubits(dr.state));
    W.res[i] = f4(result, ubits(dr.index));
}

// Queue position k of the bounce (every thread of the block calls it).
template <bool BATCH, bool SOBOL>
DEV void shade_step(const SceneDev& S, const SceneDev* __restrict__ scenes, const WaveDev& W, const uint32_t* queue, uint32_t n,
                    uint32_t bounce, uint32_t k) {
    bool valid = k < n;
    bool cont = false;
    uint32_t entries = 0, tested = 0;
    uint32_t i = 0;
    if (BATCH) {
        // when every lane of the wave shades the same still (the usual case:
        // queues keep a wave's paths together) its scene table row is read
        // with wave-uniform loads; otherwise each lane reads its own
        uint32_t sid = 0;
        if (valid) {
            i = queue ? queue[k] : k;
            sid = W.sid[i];
        }
        const uint32_t s0 = __builtin_amdgcn_readfirstlane(sid);
        if (__all(!valid || sid == s0)) {
            if (valid) shade_path<SOBOL>(scenes[s0], W, i, bounce, cont, entries, tested);
        } else if (valid) {
            shade_path<SOBOL>(scenes[sid], W, i, bounce, cont, entries, tested);
        }
    } else if (valid) {
        i = queue ? queue[k] : k;
        shade_path<SOBOL>(S, W, i, bounce, cont, entries, tested);
    }
    bool sh = valid && entries > 0u;
    uint3 at = block_append3(W.counters, cont ? 1u : 0u, sh ? 1u : 0u, (uint32_t)__popc(tested));
    if (cont) W.q_next[at.x] = i;
    if (sh) W.q_sh[at.y] = i | (entries << 28);
    for (uint32_t bits = tested, r = at.z; bits; bits &= bits - 1u, ++r)
        W.q_sray[r] = ((uint32_t)(__ffs(bits) - 1) << 28) | i;
}

// In a batch (BATCH) the path's scene is scenes[W.sid[i]] and the queue
// length comes from the counters; blocks then stride over the queue.
// Four waves per SIMD (128 registers, a few spilled) hide more of the
// shading's memory latency than three with no spills (+4% Mrays/s on the
// 4K terrain).
// A block compacts the chunk [lo, hi) of a queue into `out` in two passes:
// it counts the entries to keep (remembering each thread's answers in a
// bit mask), takes their room with one atomic on `counter`, then writes
// them in queue order. One atomic per chunk, not per 256 entries:
// device-scope atomics on one counter serialize across the XCDs (one per
// 256 entries cost 0.3 ms per 8M-entry launch). Every thread of the block
// calls it with the same chunk.
template <class Keep>
DEV void compact_chunk(uint32_t lo, uint32_t hi, const uint32_t* queue, Keep keep, uint32_t* counter,
                       uint32_t* out) {
    __shared__ uint32_t tot[4];
    __shared__ uint32_t base;
    uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    unsigned long long bits = 0ull;
    uint32_t count = 0, it = 0;
    for (uint32_t k = lo + threadIdx.x; k < hi; k += 256u, ++it) {
        bool kp = keep(queue ? queue[k] : k);
        if (kp && it < 64u) bits |= 1ull << it;
        count += kp ? 1u : 0u;
    }
    uint32_t ws = wave_scan(count);
    if (lane == 63) tot[wave] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = tot[0] + tot[1] + tot[2] + tot[3];
        base = sum ? atomicAdd(counter, sum) : 0u;
    }
    __syncthreads();
    uint32_t pos = base;
    it = 0;
    for (uint32_t t0 = lo; t0 < hi; t0 += 256u, ++it) {
        uint32_t k = t0 + threadIdx.x, i = 0;
        bool kp = false;
        if (k < hi) {
            i = queue ? queue[k] : k;
            kp = it < 64u ? ((bits >> it) & 1ull) != 0ull : keep(i);
        }
        uint32_t sc = wave_scan(kp ? 1u : 0u);
        __syncthreads();  // the previous tile's totals have been read
        if (lane == 63) tot[wave] = sc;
        __syncthreads();
        uint32_t at = pos + sc - (kp ? 1u : 0u);
        for (uint32_t w = 0; w < wave; ++w) at += tot[w];
        if (kp) out[at] = i;
        pos += tot[0] + tot[1] + tot[2] + tot[3];
    }
}

// The paths of a bounce whose ray hit something, for shading. Under a black
// sky a missed path ends with no shading work at all (no draws, no state
// change), so only hits need shading lanes; the order of the hit queue does
// not matter, every path being shaded on its own.
__global__ __launch_bounds__(256) void k_hits(WaveDev W, const uint32_t* queue, uint32_t n, uint32_t* hitq) {
    uint32_t chunk = (n + gridDim.x - 1u) / gridDim.x;
    uint32_t lo = min(n, blockIdx.x * chunk), hi = min(n, lo + chunk);
    auto hit = [&](uint32_t i) { return !(W.hit[i].x >= 1e19f); };  // shade_path's miss test, negated
    compact_chunk(lo, hi, queue, hit, W.counters + C_HITS, hitq);
}

// The same for a batch wave: the queue length is on the device, and a
// missed path is dropped only when its own still's sky is black.
__global__ __launch_bounds__(256) void k_hits_batch(BatchDev B, WaveDev W, const uint32_t* queue, uint32_t* hitq) {
    uint32_t n = W.counters[C_N];
    uint32_t chunk = (n + gridDim.x - 1u) / gridDim.x;
    uint32_t lo = min(n, blockIdx.x * chunk), hi = min(n, lo + chunk);
    auto keep = [&](uint32_t i) {
        const SceneDev& S = B.scenes[W.sid[i]];
        return !(W.hit[i].x >= 1e19f) || !S.env_black || S.area;
    };
    compact_chunk(lo, hi, queue, keep, W.counters + C_HITS, hitq);
}

template <bool BATCH, bool SOBOL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(SceneDev S, const SceneDev* __restrict__ scenes, WaveDev W, const uint32_t* queue,
                                               uint32_t n, uint32_t bounce, const uint32_t* count) {
    if (!BATCH) {
        if (count) n = *count;
        if (blockIdx.x * 256u >= n) return;  // the whole block is past the queue
        shade_step<false, SOBOL>(S, scenes, W, queue, n, bounce, blockIdx.x * 256u + threadIdx.x);
        return;
    }
    n = count ? *count : W.counters[C_N];
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u)
        shade_step<true, SOBOL>(S, scenes, W, queue, n, bounce, base + threadIdx.x);
}

// Any-hit test of every queued shadow ray (one NEE entry each). An occluded
// entry is marked in its contribution's w; k_fold adds the others.
// WIDE: the default walk's conservative box test (a trace built with the
// left-first walk keeps the exact quotients for its shadow rays too).
// Seven waves per SIMD for a single trace (72 registers, none spilled, and an
// 11-entry LDS stack so 14 blocks fit a CU): 3% faster than six on the 4K
// terrain. A batch walk holds its still's scene in registers and keeps six.
// Q: the scene's nodes are quantized (QNode4), with the default walk.
template <bool COUNT, bool BATCH, bool WIDE, bool Q = false>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(BATCH ? 6 : 7))) void k_shadow(SceneDev S, const SceneDev* scenes, WaveDev W, uint32_t n) {
    __shared__ uint2 lds[STACK_SHORT * TRACE_BLOCK];
    uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    Stack st{lds + threadIdx.x, W.spill + gtid, W.spill_threads, 0};
    uint32_t nodes = 0, prims = 0;
    Walk<true> w;
    Geo g{};
    if (BATCH) n = W.counters[C_N + 1];
    size_t at = 0;
    bool live = false;
    Fetch f = fetch_start();
    for (;;) {
        uint32_t k = refill_lanes(!live, W.refill, n, W.counters + C_SFETCH, f);
        if (k != 0xffffffffu) {
            uint32_t q = W.q_sray[k];
            at = (size_t)(q >> 28) * W.capacity + (q & 0x0fffffffu);
            float4 eo = W.e_o[at], ed = W.e_d[at];
            if (BATCH) g = geo_of(scenes[W.sid[q & 0x0fffffffu]]);
            live = w.start(walk_scene<BATCH>(S, g), f3(eo), f3(ed), eo.w, st);
        }
        if (!__any(live)) {
            if (fetch_drained(f)) break;
            continue;
        }
        bool in = live && w.inner();
        bool out = live && !in;
        if (inner_round(in, out)) {
            if (in) {
                if (COUNT) nodes++;
                bool more;
                if constexpr (Q)
                    more = visit4q<false>(walk_scene<BATCH>(S, g), w.cur, w.R, w.h.t, st, w.cur);
                else
                    more = visit4<false, WIDE>(walk_scene<BATCH>(S, g), w.cur, w.R, w.h.t, st, w.cur);
                if (!more && !w.next(st)) live = false;
            }
        } else if (out) {
            bool go = true;
            if (w.leafy()) {
                if (w.template leaf<COUNT>(walk_scene<BATCH>(S, g), prims)) {
                    reinterpret_cast<float*>(W.e_c + at)[3] = 1.0f;
                    live = false;
                    go = false;
                }
            } else {
                go = !w.enter(walk_scene<BATCH>(S, g), st);
            }
            if (go && !w.next(st)) live = false;
        }
    }
    if (COUNT) {
        atomicAdd(reinterpret_cast<unsigned long long*>(W.counters + C_SNODES), (unsigned long long)nodes);
        atomicAdd(reinterpret_cast<unsigned long long*>(W.counters + C_SPRIMS), (unsigned long long)prims);
    }
}

// Add each queued path's unoccluded light contributions in slot order
// (trace.wgsl:2465-2470).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_fold(WaveDev W, uint32_t n) {
    if (BATCH) n = W.counters[C_N + 2];
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u) {
        uint32_t q = W.q_sh[k];
        uint32_t i = q & 0x0fffffffu, m = q >> 28;
        float4 r4 = W.res[i];
        v3 result = f3(r4);
        for (uint32_t e = 0; e < m; ++e) {
            float4 c = W.e_c[(size_t)e * W.capacity + i];
            if (c.w == 0.0f) result = result + f3(c);
        }
        W.res[i] = f4(result, r4.w);  // w: the Sobol sample index
    }
}

DEV void neumaier(float& s, float& c, float x) {
    float t = s + x;
    if (fabsf(s) >= fabsf(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// Fold a wave's samples into each pixel's sums, in sample order
// (trace.wgsl:2562-2571).
// Pixel `pix` of a still whose samples in the wave are paths
// first + s * npix + p, s < nsamples.
DEV float lum_sq(float4 r) {
    const float l = (r.x + r.y + r.z) / 3.0f;
    return l * l;
}
DEV void finish_pixel(const WaveDev& W, const AccumDev& A, uint32_t pix, uint32_t first, uint32_t npix, uint32_t p,
                      uint32_t nsamples, float count_after) {
    p += first;
    // the luminance moment, in sample order, as the oracle sums it
    float m2 = A.m2[pix];
    for (uint32_t s = 0; s < nsamples; ++s) m2 = m2 + lum_sq(W.res[s * npix + p]);
    A.m2[pix] = m2;
    if (A.mode == PFXH_ACCUM_PLAIN) {
        float4 c = A.sum[pix * 3u], a = A.sum[pix * 3u + 1u], nn = A.sum[pix * 3u + 2u];
        for (uint32_t s = 0; s < nsamples; ++s) {
            uint32_t i = s * npix + p;
            float4 r = W.res[i], al = W.alb[i], no = W.nrm[i];
            c.x = c.x + r.x;
            c.y = c.y + r.y;
            c.z = c.z + r.z;
            a.x = a.x + al.x;
            a.y = a.y + al.y;
            a.z = a.z + al.z;
            a.w = a.w + al.w;
            nn.x = nn.x + no.x;
            nn.y = nn.y + no.y;
            nn.z = nn.z + no.z;
            nn.w = nn.w + no.w;
        }
        c.w = count_after;
        A.sum[pix * 3u] = c;
        A.sum[pix * 3u + 1u] = a;
        A.sum[pix * 3u + 2u] = nn;
    } else if (A.mode == 2u) {
        double* D = A.sum64 + (size_t)pix * 12u;
        double acc[12];
        for (int k = 0; k < 12; ++k) acc[k] = D[k];
        for (uint32_t s = 0; s < nsamples; ++s) {
            uint32_t i = s * npix + p;
            float4 r = W.res[i], al = W.alb[i], no = W.nrm[i];
            float v[12] = {r.x, r.y, r.z, 0.0f, al.x, al.y, al.z, al.w, no.x, no.y, no.z, no.w};
            for (int k = 0; k < 12; ++k)
                if (k != 3) acc[k] = acc[k] + (double)v[k];
        }
        acc[3] = (double)count_after;
        for (int k = 0; k < 12; ++k) D[k] = acc[k];
        float4 c = A.sum[pix * 3u];
        c.w = count_after;
        A.sum[pix * 3u] = c;
    } else {
        float4 S0 = A.sum[pix * 3u], S1 = A.sum[pix * 3u + 1u], S2 = A.sum[pix * 3u + 2u];
        float4 C0 = A.comp[pix * 3u], C1 = A.comp[pix * 3u + 1u], C2 = A.comp[pix * 3u + 2u];
        for (uint32_t s = 0; s < nsamples; ++s) {
            uint32_t i = s * npix + p;
            float4 r = W.res[i], al = W.alb[i], no = W.nrm[i];
            neumaier(S0.x, C0.x, r.x);
            neumaier(S0.y, C0.y, r.y);
            neumaier(S0.z, C0.z, r.z);
            neumaier(S1.x, C1.x, al.x);
            neumaier(S1.y, C1.y, al.y);
            neumaier(S1.z, C1.z, al.z);
            neumaier(S1.w, C1.w, al.w);
            neumaier(S2.x, C2.x, no.x);
            neumaier(S2.y, C2.y, no.y);
            neumaier(S2.z, C2.z, no.z);
            neumaier(S2.w, C2.w, no.w);
        }
        S0.w = count_after;
        A.sum[pix * 3u] = S0;
        A.sum[pix * 3u + 1u] = S1;
        A.sum[pix * 3u + 2u] = S2;
        A.comp[pix * 3u] = C0;
        A.comp[pix * 3u + 1u] = C1;
        A.comp[pix * 3u + 2u] = C2;
    }
}

__global__ __launch_bounds__(256) void k_finish(WaveDev W, AccumDev A, const uint32_t* active, uint32_t npix,
                                                uint32_t nsamples, float count_after) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= npix) return;
    finish_pixel(W, A, active ? active[p] : p, 0u, npix, p, nsamples, count_after);
}

// Batch: every pixel of every unit of the current wave, into its still's sums.
__global__ __launch_bounds__(256) void k_finish_batch(BatchDev B, WaveDev W) {
    const WavePlan& P = B.waves[*B.wave];
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < P.pixels; k += gridDim.x * 256u) {
        const Unit& u = P.u[unit_of(P, k, true)];
        uint32_t p = k - u.pfirst;
        finish_pixel(W, B.accums[u.sid], p, u.first, u.npix, p, u.ns, (float)(u.acc + u.ns));
    }
}

// Queue lengths of a batch bounce, kept on the device so the bounce loop
// needs no host round trip (and can be captured in a graph):
//   phase 0, before k_extend: the queue length (all paths of the wave at
//            bounce 0, else what the last k_shade appended) to C_N, the
//            traversal work counters cleared, rays counted
//   phase 1, before k_shade: the append counters cleared
//   phase 2, before k_shadow: shadow rays and shadow paths to C_N + 1 and
//            C_N + 2, the shadow work counters cleared, rays counted
__global__ void k_roll(BatchDev B, uint32_t* counters, uint32_t phase, uint32_t bounce) {
    uint32_t t = threadIdx.x;
    if (phase == 0u) {
        uint32_t n = bounce == 0u ? B.waves[*B.wave].paths : counters[C_NEXT];
        __syncthreads();
        if (t < PARTS) counters[C_FETCH + t * SLOT] = 0u;
        if (t == 0) {
            counters[C_N] = n;
            *reinterpret_cast<unsigned long long*>(counters + C_STATS) += n;
        }
    } else if (phase == 1u) {
        if (t == 0) {
            counters[C_NEXT] = 0u;
            counters[C_SHADOW] = 0u;
            counters[C_SRAYS] = 0u;
            counters[C_HITS] = 0u;
        }
    } else {
        uint32_t rays = counters[C_SRAYS], paths = counters[C_SHADOW];
        __syncthreads();
        if (t < PARTS) counters[C_SFETCH + t * SLOT] = 0u;
        if (t == 0) {
            counters[C_N + 1] = rays;
            counters[C_N + 2] = paths;
            *reinterpret_cast<unsigned long long*>(counters + C_STATS + 2) += rays;
        }
    }
}

__global__ void k_next_wave(BatchDev B) {
    if (threadIdx.x == 0) *B.wave += 1u;
}

DEV float total_of(const AccumDev& A, uint32_t pix, int k) {
    if (A.mode == 2u) return (float)A.sum64[(size_t)pix * 12u + k];
    const float* s = reinterpret_cast<const float*>(A.sum + pix * 3u);
    if (A.mode == PFXH_ACCUM_PLAIN) return s[k];
    const float* c = reinterpret_cast<const float*>(A.comp + pix * 3u);
    return s[k] + c[k];
}

// The standard error of each pixel's mean luminance: the learned denoiser's
// noise channel (a firefly is a pixel whose mean a few samples dominate).
__global__ __launch_bounds__(256) void k_deviation(AccumDev A, uint32_t pixels, float* out) {
    uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= pixels) return;
    const float n = fabsf(A.sum[pix * 3u].w);
    float se = 0.0f;
    if (n > 0.0f) {
        const float l = (total_of(A, pix, 0) / n + total_of(A, pix, 1) / n + total_of(A, pix, 2) / n) / 3.0f;
        se = sqrtf(fmaxf(A.m2[pix] / n - l * l, 0.0f) / n);
    }
    out[pix] = se;
}

// The three RGBA32F images (trace.wgsl:2572-2578).
__global__ __launch_bounds__(256) void k_develop(AccumDev A, uint32_t pixels, float4* color, float4* albedo,
                                                 float4* normal) {
    uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= pixels) return;
    float w = A.sum[pix * 3u].w;
    float count = fmaxf(fabsf(w), 1.0f);
    if (A.mode == 2u) {
        const double* D = A.sum64 + (size_t)pix * 12u;
        double c = (double)count;
        if (color) color[pix] = make_float4((float)(D[0] / c), (float)(D[1] / c), (float)(D[2] / c), 1.0f);
        if (albedo) albedo[pix] = make_float4((float)(D[4] / c), (float)(D[5] / c), (float)(D[6] / c), (float)(D[7] / c));
        v3 n = mk((float)(D[8] / c), (float)(D[9] / c), (float)(D[10] / c));
        v3 nn = pfx::length(n) > 1e-6f ? normalize(n) : mk(0, 0, 0);
        if (normal) normal[pix] = f4(nn, (float)(D[11] / c));
        return;
    }
    if (color)
        color[pix] = make_float4(total_of(A, pix, 0) / count, total_of(A, pix, 1) / count, total_of(A, pix, 2) / count,
                                 1.0f);
    if (albedo)
        albedo[pix] = make_float4(total_of(A, pix, 4) / count, total_of(A, pix, 5) / count,
                                  total_of(A, pix, 6) / count, total_of(A, pix, 7) / count);
    v3 n = mk(total_of(A, pix, 8) / count, total_of(A, pix, 9) / count, total_of(A, pix, 10) / count);
    v3 nn = pfx::length(n) > 1e-6f ? normalize(n) : mk(0, 0, 0);
    if (normal) normal[pix] = f4(nn, total_of(A, pix, 11) / count);
}

__global__ __launch_bounds__(256) void k_clear(AccumDev A, uint32_t pixels) {
    uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= pixels) return;
    float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int k = 0; k < 3; ++k) {
        A.sum[pix * 3u + k] = z;
        if (A.comp) A.comp[pix * 3u + k] = z;
    }
    if (A.sum64)
        for (int k = 0; k < 12; ++k) A.sum64[(size_t)pix * 12u + k] = 0.0;
    A.m2[pix] = 0.0f;
}

// Pixels not frozen by adaptive sampling (trace.wgsl:2517).
// Each block compacts one contiguous chunk of pixels with one atomic (one
// atomic per wave on the one counter serialised across the XCDs: 963 us per
// call on the 4K frame against 84 this way); the order of the active list
// does not matter, every pixel's samples being seeded by the pixel and the
// sample index.
__global__ __launch_bounds__(256) void k_active(AccumDev A, uint32_t pixels, uint32_t* active, uint32_t* counter) {
    uint32_t chunk = (pixels + gridDim.x - 1u) / gridDim.x;
    uint32_t lo = min(pixels, blockIdx.x * chunk), hi = min(pixels, lo + chunk);
    auto live = [&](uint32_t pix) { return !(A.sum[pix * 3u].w < 0.0f); };
    compact_chunk(lo, hi, nullptr, live, counter, active);
}

// Pixel-centre primary hits (probe.wgsl:3-39) plus the caller's primitive.
__global__ __launch_bounds__(TRACE_BLOCK) void k_probe(SceneDev S, WaveDev W, pfxh_probe_hit* out) {
    __shared__ uint2 lds[STACK_SHORT * TRACE_BLOCK];
    uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    uint32_t stride = gridDim.x * TRACE_BLOCK;
    Stack st{lds + threadIdx.x, W.spill + gtid, W.spill_threads, 0};
    uint32_t pixels = S.width * S.height;
    v3 fwd = mk(S.forward[0], S.forward[1], S.forward[2]), right = mk(S.right[0], S.right[1], S.right[2]),
       up = mk(S.up[0], S.up[1], S.up[2]);
    for (uint64_t p64 = gtid; p64 < pixels; p64 += stride) {  // 64-bit: no wrap for pixels near 2^32
        uint32_t pix = (uint32_t)p64;
        uint32_t y = pix / S.width, x = pix - y * S.width;
        float ndcx = ((float)(x + S.window[2]) + 0.5f) / (float)S.window[0] * 2.0f - 1.0f;
        float ndcy = ((float)(y + S.window[3]) + 0.5f) / (float)S.window[1] * 2.0f - 1.0f;
        v3 d = normalize(fwd + right * (ndcx + S.lens[0]) + up * (-ndcy + S.lens[1]));
        v3 o = mk(S.origin[0], S.origin[1], S.origin[2]);
        if (S.projection != 0u) project(S, ndcx, ndcy, fwd, right, up, o, d);
        Closest h = trace_closest(S, o, d, 1e20f, st);
        pfxh_probe_hit r;
        r.depth = 0.0f;
        r.distance = 0.0f;
        r.material_id = 0u;
        r.primitive = 0xffffffffu;
        r.instance = 0u;
        r.normal[0] = r.normal[1] = r.normal[2] = 0.0f;
        if (h.t < 1e19f) {
            uint32_t rec = S.root.top ? S.groups[h.grp].record : 0u;
            Surf sf = surface_at(S, h.slot, rec, h.u, h.v);
            v3 n = dot(sf.n, d) > 0.0f ? -sf.n : sf.n;
            r.depth = S.projection == 2u ? h.t : h.t * dot(d, normalize(fwd));
            r.distance = h.t;
            r.material_id = sf.material + 1u;
            r.primitive = S.slot_prim[h.slot];
            r.instance = S.record_instance[rec];
            r.normal[0] = n.x;
            r.normal[1] = n.y;
            r.normal[2] = n.z;
        }
        out[pix] = r;
    }
}

// converge.wgsl:9-30
__global__ __launch_bounds__(256) void k_measure(AccumDev A, float4* snaps, uint32_t width, uint32_t height) {
    uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= width * height) return;
    float count = A.sum[pix * 3u].w;
    if (count <= 0.0f) {
        snaps[pix * 2u + 1u].y = 0.0f;
        return;
    }
    float now[4] = {total_of(A, pix, 0) / count, total_of(A, pix, 1) / count, total_of(A, pix, 2) / count,
                    total_of(A, pix, 7) / count};
    float4 mean = snaps[pix * 2u], info = snaps[pix * 2u + 1u];
    float prior = info.x;
    float m[4] = {mean.x, mean.y, mean.z, mean.w};
    float error = 1e9f;
    if (prior > 0.0f && count > prior) {
        float scale = sqrtf(1.0f / count) / sqrtf(1.0f / prior + 1.0f / (count - prior));
        float gap[4];
        for (int k = 0; k < 4; ++k) {
            float rest = (now[k] * count - m[k] * prior) / (count - prior);
            gap[k] = fabsf(m[k] - rest) * scale;
        }
        error = (gap[0] + gap[1] + gap[2]) / (1e-4f + sqrtf(fmaxf(now[0] + now[1] + now[2], 0.0f))) + gap[3];
    }
    snaps[pix * 2u] = make_float4(now[0], now[1], now[2], now[3]);
    snaps[pix * 2u + 1u] = make_float4(count, error, 0.0f, 0.0f);
}

// converge.wgsl:32-51
// Blocks walk the frame with a grid stride and add their running count with
// one atomic each (one per wave took 989 us per call on the 4K frame, 148
// this way).
__global__ __launch_bounds__(256) void k_mark(AccumDev A, const float4* snaps, uint32_t width, uint32_t height,
                                              uint32_t min_samples, float threshold, uint32_t* running) {
    __shared__ uint32_t tot[4];
    const uint32_t pixels = width * height;
    uint32_t runs = 0;
    for (uint32_t pix = blockIdx.x * 256u + threadIdx.x; pix < pixels; pix += gridDim.x * 256u) {
        float count = A.sum[pix * 3u].w;
        if (count > 0.0f) {
            int x = (int)(pix % width), y = (int)(pix / width);
            float worst = 0.0f;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = min(max(x + dx, 0), (int)width - 1), yy = min(max(y + dy, 0), (int)height - 1);
                    worst = fmaxf(worst, snaps[((uint32_t)yy * width + (uint32_t)xx) * 2u + 1u].y);
                }
            if (count >= (float)min_samples && worst < threshold)
                A.sum[pix * 3u].w = -count;
            else
                ++runs;
        }
    }
    uint32_t ws = wave_scan(runs);
    if (__lane_id() == 63u) tot[threadIdx.x >> 6] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = tot[0] + tot[1] + tot[2] + tot[3];
        if (sum) atomicAdd(running, sum);
    }
}

__global__ void k_zero(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- launchers ---------------------------------------------------------------------
// Streaming copy for the achievable-HBM probe: each thread moves float4s
// with a grid stride, so every wave reads and writes whole 1 KiB runs;
// nontemporal loads and stores keep the stream from churning the caches.
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) {
        float4 v;
        v.x = __builtin_nontemporal_load(&src[i].x);
        v.y = __builtin_nontemporal_load(&src[i].y);
        v.z = __builtin_nontemporal_load(&src[i].z);
        v.w = __builtin_nontemporal_load(&src[i].w);
        __builtin_nontemporal_store(v.x, &dst[i].x);
        __builtin_nontemporal_store(v.y, &dst[i].y);
        __builtin_nontemporal_store(v.z, &dst[i].z);
        __builtin_nontemporal_store(v.w, &dst[i].w);
    }
}

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

void launch_stream_copy(hipStream_t s, const float4* src, float4* dst, uint64_t n, uint32_t blocks) {
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s, src, dst, n);
}

void launch_camera(hipStream_t s, const SceneDev& S, const WaveDev& W, const uint32_t* active, uint32_t npix,
                   uint32_t nsamples, uint32_t acc, uint32_t seed) {
    uint64_t n = (uint64_t)npix * nsamples;
    if (!n) return;
    hipLaunchKernelGGL(k_camera, dim3(blocks_for(n, 256)), dim3(256), 0, s, S, W, active, npix, nsamples, acc, seed);
}
void launch_extend(hipStream_t s, const SceneDev& S, const WaveDev& W, const uint32_t* queue, uint32_t n, bool count,
                   uint32_t grid) {
    if (!n) return;
    uint32_t blocks = std::min(blocks_for(n, TRACE_BLOCK), grid);
    const SceneDev* none = nullptr;
    bool left = (S.geo_flags & GEO_LEFT_FIRST) != 0u, q = !left && S.qnodes;
    const dim3 gd(blocks), block(TRACE_BLOCK);
    if (count && left)
        hipLaunchKernelGGL((k_extend<true, false, false>), gd, block, 0, s, S, none, W, queue, n);
    else if (count && q)
        hipLaunchKernelGGL((k_extend<true, false, true, true>), gd, block, 0, s, S, none, W, queue, n);
    else if (count)
        hipLaunchKernelGGL((k_extend<true, false, true>), gd, block, 0, s, S, none, W, queue, n);
    else if (left)
        hipLaunchKernelGGL((k_extend<false, false, false>), gd, block, 0, s, S, none, W, queue, n);
    else if (q)
        hipLaunchKernelGGL((k_extend<false, false, true, true>), gd, block, 0, s, S, none, W, queue, n);
    else
        hipLaunchKernelGGL((k_extend<false, false, true>), gd, block, 0, s, S, none, W, queue, n);
}
void launch_hits(hipStream_t s, const WaveDev& W, const uint32_t* queue, uint32_t n, uint32_t* hitq) {
    if (!n) return;
    // chunks of at most 64 tiles of 256 (the bit mask), at least 2048 blocks
    uint32_t blocks = std::max(2048u, blocks_for(n, 64u * 256u));
    hipLaunchKernelGGL(k_hits, dim3(blocks), dim3(256), 0, s, W, queue, n, hitq);
}
void launch_shade(hipStream_t s, const SceneDev& S, const WaveDev& W, const uint32_t* queue, uint32_t n,
                  uint32_t bounce, const uint32_t* count) {
    if (!n) return;
    const SceneDev* none = nullptr;
    if (S.sobol_table)
        hipLaunchKernelGGL((k_shade<false, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s, S, none, W, queue, n,
                           bounce, count);
    else
        hipLaunchKernelGGL((k_shade<false, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s, S, none, W, queue, n,
                           bounce, count);
}
void launch_shadow(hipStream_t s, const SceneDev& S, const WaveDev& W, uint32_t n, bool count, uint32_t grid) {
    if (!n) return;
    uint32_t blocks = std::min(blocks_for(n, TRACE_BLOCK), grid);
    const SceneDev* none = nullptr;
    bool left = (S.geo_flags & GEO_LEFT_FIRST) != 0u, q = !left && S.qnodes;
    const dim3 gd(blocks), block(TRACE_BLOCK);
    if (count && left)
        hipLaunchKernelGGL((k_shadow<true, false, false>), gd, block, 0, s, S, none, W, n);
    else if (count && q)
        hipLaunchKernelGGL((k_shadow<true, false, true, true>), gd, block, 0, s, S, none, W, n);
    else if (count)
        hipLaunchKernelGGL((k_shadow<true, false, true>), gd, block, 0, s, S, none, W, n);
    else if (left)
        hipLaunchKernelGGL((k_shadow<false, false, false>), gd, block, 0, s, S, none, W, n);
    else if (q)
        hipLaunchKernelGGL((k_shadow<false, false, true, true>), gd, block, 0, s, S, none, W, n);
    else
        hipLaunchKernelGGL((k_shadow<false, false, true>), gd, block, 0, s, S, none, W, n);
}
void launch_fold(hipStream_t s, const WaveDev& W, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_fold<false>, dim3(blocks_for(n, 256)), dim3(256), 0, s, W, n);
}
void launch_finish(hipStream_t s, const WaveDev& W, const AccumDev& A, const uint32_t* active, uint32_t npix,
                   uint32_t nsamples, float count_after) {
    if (!npix) return;
    hipLaunchKernelGGL(k_finish, dim3(blocks_for(npix, 256)), dim3(256), 0, s, W, A, active, npix, nsamples,
                       count_after);
}
void launch_develop(hipStream_t s, const AccumDev& A, uint32_t pixels, float4* color, float4* albedo, float4* normal) {
    if (!pixels) return;
    hipLaunchKernelGGL(k_develop, dim3(blocks_for(pixels, 256)), dim3(256), 0, s, A, pixels, color, albedo, normal);
}
void launch_clear(hipStream_t s, const AccumDev& A, uint32_t pixels) {
    if (!pixels) return;
    hipLaunchKernelGGL(k_clear, dim3(blocks_for(pixels, 256)), dim3(256), 0, s, A, pixels);
}
void launch_deviation(hipStream_t s, const AccumDev& A, uint32_t pixels, float* out) {
    if (!pixels) return;
    hipLaunchKernelGGL(k_deviation, dim3(blocks_for(pixels, 256)), dim3(256), 0, s, A, pixels, out);
}
void launch_active(hipStream_t s, const AccumDev& A, uint32_t pixels, uint32_t* active, uint32_t* counter) {
    if (!pixels) return;
    // chunks of at most 64 tiles of 256 (the bit mask), at least 2048 blocks
    uint32_t blocks = std::max(2048u, blocks_for(pixels, 64u * 256u));
    hipLaunchKernelGGL(k_active, dim3(blocks), dim3(256), 0, s, A, pixels, active, counter);
}
// Closest hits of a caller's rays (origins and directions as float4, w
// unused): the caller's triangle id per ray, UINT32_MAX on a miss, and the
// distance (0 on a miss).
__global__ __launch_bounds__(TRACE_BLOCK) void k_cast(SceneDev S, WaveDev W, const float4* o, const float4* d,
                                                      uint32_t n, float limit, uint32_t* ids, float* t) {
    __shared__ uint2 lds[STACK_SHORT * TRACE_BLOCK];
    uint32_t gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * TRACE_BLOCK;
    Stack st{lds + threadIdx.x, W.spill + gtid, W.spill_threads, 0};
    for (uint64_t k = gtid; k < n; k += stride) {  // 64-bit: no wrap for n near 2^32
        Closest h = trace_closest(S, f3(o[k]), f3(d[k]), limit, st);
        bool hit = h.t < limit;
        ids[k] = hit ? S.slot_prim[h.slot] : 0xffffffffu;
        t[k] = hit ? h.t : 0.0f;
    }
}
void launch_cast(hipStream_t s, const SceneDev& S, const WaveDev& W, const float4* o, const float4* d, uint32_t n,
                 float limit, uint32_t* ids, float* t) {
    if (!n) return;
    uint32_t blocks = std::min(blocks_for(n, TRACE_BLOCK), W.spill_threads / TRACE_BLOCK);
    hipLaunchKernelGGL(k_cast, dim3(blocks), dim3(TRACE_BLOCK), 0, s, S, W, o, d, n, limit, ids, t);
}
void launch_probe(hipStream_t s, const SceneDev& S, const WaveDev& W, void* out) {
    uint32_t pixels = S.width * S.height;
    if (!pixels) return;
    uint32_t blocks = std::min(blocks_for(pixels, TRACE_BLOCK), W.spill_threads / TRACE_BLOCK);
    hipLaunchKernelGGL(k_probe, dim3(blocks), dim3(TRACE_BLOCK), 0, s, S, W, (pfxh_probe_hit*)out);
}
void launch_measure(hipStream_t s, const AccumDev& A, float4* snaps, uint32_t width, uint32_t height) {
    uint32_t pixels = width * height;
    if (!pixels) return;
    hipLaunchKernelGGL(k_measure, dim3(blocks_for(pixels, 256)), dim3(256), 0, s, A, snaps, width, height);
}
void launch_mark(hipStream_t s, const AccumDev& A, const float4* snaps, uint32_t width, uint32_t height,
                 uint32_t min_samples, float threshold, uint32_t* running) {
    uint32_t pixels = width * height;
    if (!pixels) return;
    hipLaunchKernelGGL(k_mark, dim3(std::min(2048u, blocks_for(pixels, 256))), dim3(256), 0, s, A, snaps, width,
                       height, min_samples, threshold, running);
}
void launch_zero(hipStream_t s, uint32_t* p, uint32_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
}

// ---- batch launchers: fixed grids, lengths from the device ------------------------
void launch_camera_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t grid) {
    hipLaunchKernelGGL(k_camera_batch, dim3(grid), dim3(256), 0, s, B, W);
}
void launch_roll(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t phase, uint32_t bounce) {
    hipLaunchKernelGGL(k_roll, dim3(1), dim3(64), 0, s, B, W.counters, phase, bounce);
}
void launch_extend_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, const uint32_t* queue, uint32_t grid,
                         bool left_first, bool quantized) {
    SceneDev none{};
    if (left_first)
        hipLaunchKernelGGL((k_extend<false, true, false>), dim3(grid), dim3(TRACE_BLOCK), 0, s, none, B.scenes, W, queue, 0u);
    else if (quantized)
        hipLaunchKernelGGL((k_extend<false, true, true, true>), dim3(grid), dim3(TRACE_BLOCK), 0, s, none, B.scenes, W,
                           queue, 0u);
    else
        hipLaunchKernelGGL((k_extend<false, true, true>), dim3(grid), dim3(TRACE_BLOCK), 0, s, none, B.scenes, W, queue, 0u);
}
void launch_hits_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, const uint32_t* queue, uint32_t* hitq,
                       uint32_t grid) {
    hipLaunchKernelGGL(k_hits_batch, dim3(grid), dim3(256), 0, s, B, W, queue, hitq);
}
void launch_shade_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, const uint32_t* queue, uint32_t bounce,
                        uint32_t grid, const uint32_t* count, bool sobol) {
    SceneDev none{};
    if (sobol)
        hipLaunchKernelGGL((k_shade<true, true>), dim3(grid), dim3(256), 0, s, none, B.scenes, W, queue, 0u, bounce,
                           count);
    else
        hipLaunchKernelGGL((k_shade<true, false>), dim3(grid), dim3(256), 0, s, none, B.scenes, W, queue, 0u, bounce,
                           count);
}
void launch_shadow_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t grid, bool left_first,
                         bool quantized) {
    SceneDev none{};
    if (left_first)
        hipLaunchKernelGGL((k_shadow<false, true, false>), dim3(grid), dim3(TRACE_BLOCK), 0, s, none, B.scenes, W, 0u);
    else if (quantized)
        hipLaunchKernelGGL((k_shadow<false, true, true, true>), dim3(grid), dim3(TRACE_BLOCK), 0, s, none, B.scenes, W,
                           0u);
    else
        hipLaunchKernelGGL((k_shadow<false, true, true>), dim3(grid), dim3(TRACE_BLOCK), 0, s, none, B.scenes, W, 0u);
}
void launch_fold_batch(hipStream_t s, const WaveDev& W, uint32_t grid) {
    hipLaunchKernelGGL(k_fold<true>, dim3(grid), dim3(256), 0, s, W, 0u);
}
void launch_finish_batch(hipStream_t s, const BatchDev& B, const WaveDev& W, uint32_t grid) {
    hipLaunchKernelGGL(k_finish_batch, dim3(grid), dim3(256), 0, s, B, W);
}
void launch_next_wave(hipStream_t s, const BatchDev& B) {
    hipLaunchKernelGGL(k_next_wave, dim3(1), dim3(64), 0, s, B);
}

}  // namespace pfxh
