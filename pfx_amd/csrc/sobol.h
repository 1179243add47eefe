// The Sobol sampler of the reference (`Sampler::Sobol`, crates/trace/src/
// sobol.rs and sobol.wgsl): a table of 256 scrambled 4-D patterns of
// `length` points each (a power of two, 256..8192), built on the host, and
// a per-dimension lookup on the device that shuffles the pattern, Owen-
// scrambles the sample index and Cranley-Patterson-rotates the point.
//
// The methods are published ones: the patterns are Cycles' progressive
// multi-jittered (0,2) construction of Christensen, Kensler and Kilpatrick
// ("Progressive Multi-Jittered Sample Sequences", EGSR 2018) driven by
// Cycles' hash_hp integer hash; the index scramble is Burley's hash-based
// Owen scramble ("Practical Hash-based Owen Scrambling", JCGT 9(4), 2020);
// the pattern shuffle is Kensler's hashed permutation ("Correlated
// Multi-Jittered Sampling", Pixar 2013); the rotations use Wang's integer
// hash. The constants and the XOR table are the reference's: the sample
// values must be its values, bit for bit.
#pragma once
#include <stdint.h>

#include "fmath.h"

namespace pfx {

constexpr uint32_t SOBOL_PATTERNS = 256u;  // sobol.rs:1
constexpr uint32_t SOBOL_MIN = 256u, SOBOL_MAX = 8192u;

// sequence_size (sobol.rs:33-39): the table length for a sample count
PFX_HD uint32_t sobol_length(uint32_t samples) {
    uint32_t n = samples < 1u ? 1u : samples;
    uint64_t p = 1;
    while (p < n) p <<= 1;
    if (p > SOBOL_MAX) p = SOBOL_MAX;
    return p < SOBOL_MIN ? SOBOL_MIN : (uint32_t)p;
}

PFX_HD uint32_t rev32(uint32_t v) {
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
    v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
    return (v >> 16) | (v << 16);
}
PFX_HD uint32_t clz32(uint32_t v) { return v == 0u ? 32u : (uint32_t)__builtin_clz(v); }
// sobol.wgsl:2-9
PFX_HD uint32_t reversed_bit_owen(uint32_t n, uint32_t seed) {
    n ^= n * 0x3d20adeau;
    n += seed;
    n *= (seed >> 16) | 1u;
    n ^= n * 0x05526c56u;
    n ^= n * 0x53a22864u;
    return n;
}
PFX_HD uint32_t nested_uniform_scramble(uint32_t i, uint32_t seed) {  // :10-12
    return rev32(reversed_bit_owen(rev32(i), seed));
}
PFX_HD uint32_t hash_wang_seeded(uint32_t i, uint32_t seed) {  // :13-19
    i = (i ^ 61u) ^ seed;
    i += i << 3;
    i ^= i >> 4;
    i *= 0x27d4eb2du;
    return i;
}
PFX_HD uint32_t hash_shuffle(uint32_t i, uint32_t length, uint32_t seed) {  // :20-43
    i = i % length;
    uint32_t mask = (1u << (32u - clz32(length - 1u))) - 1u;  // length >= 2
    for (;;) {
        i ^= seed;
        i *= 0xe170893du;
        i ^= seed >> 16;
        i ^= (i & mask) >> 4;
        i ^= seed >> 8;
        i *= 0x0929eb3fu;
        i ^= seed >> 23;
        i ^= (i & mask) >> 1;
        i *= 1u | (seed >> 27);
        i *= 0x6935fa69u;
        i ^= (i & mask) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & mask) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & mask) >> 2;
        i *= 0xc860a3dfu;
        i &= mask;
        i ^= i >> 5;
        if (i < length) break;
    }
    return i;
}
// trace.wgsl:428-432: the per-pixel seed of the Sobol patterns
PFX_HD uint32_t hash_iqnt2d(uint32_t x, uint32_t y) {
    uint32_t qx = 1103515245u * ((x >> 1) ^ y);
    uint32_t qy = 1103515245u * ((y >> 1) ^ x);
    return 1103515245u * (qx ^ (qy >> 3));
}
PFX_HD float fract(float x) { return x - floorf(x); }

PFX_HD uint32_t rotl32(uint32_t v, uint32_t k) { return (v << k) | (v >> (32u - k)); }
// hash_uint2 (sobol.rs:84-103): Bob Jenkins' lookup3 hash of two words
// (hashword with length 2 and initial value 13, as Cycles' hash_uint2):
// a = b = c = 0xdeadbeef + (2 << 2) + 13, then b += ky, a += kx, then
// lookup3's `final` mix.
PFX_HD uint32_t hash_uint2(uint32_t kx, uint32_t ky) {
    uint32_t a = 0xdeadbeefu + (2u << 2) + 13u;
    uint32_t b = a + ky, c = a;
    a += kx;
    c ^= b; c -= rotl32(b, 14);
    a ^= c; a -= rotl32(c, 11);
    b ^= a; b -= rotl32(a, 25);
    c ^= b; c -= rotl32(b, 16);
    a ^= c; a -= rotl32(c, 4);
    b ^= a; b -= rotl32(a, 14);
    c ^= b; c -= rotl32(b, 24);
    return c;
}
// animated_seed (sobol.rs:105-107): the trace seed of frame `frame` of an
// animation rendered with `seed` (still::Seed::Animated, still/src/trace.rs:77-82)
PFX_HD uint32_t animated_seed(uint32_t seed, uint32_t frame) { return hash_uint2(frame, seed); }

// sobol (sobol.wgsl:44-55): the point of dimension `dim` (rng.z + the
// dimension) for pixel seed `seed` (rng.x) and sample `index` (rng.y)
struct Sobol4 {
    float x, y, z, w;
};
PFX_HD Sobol4 sobol_point(const float* table, uint32_t length, uint32_t seed, uint32_t index, uint32_t dim) {
    uint32_t pattern = hash_shuffle(dim, SOBOL_PATTERNS, seed);
    uint32_t mask = length - 1u;
    uint32_t shuffled = nested_uniform_scramble(index, hash_wang_seeded(dim, seed));
    uint32_t sample = (index & ~mask) | (shuffled & mask);
    uint32_t at = (pattern * length + sample) % (length * SOBOL_PATTERNS);  // u32 arithmetic, as WGSL's
    const float* t = table + 4u * (size_t)at;
    const float scale = 2.3283064e-10f;
    float s0 = (float)hash_wang_seeded(dim, seed) * scale;
    float s1 = (float)hash_wang_seeded(dim, seed ^ 0xca0e1151u) * scale;
    float s2 = (float)hash_wang_seeded(dim, seed ^ 0xbf604c5au) * scale;
    float s3 = (float)hash_wang_seeded(dim, seed ^ 0x99634d1du) * scale;
    const float top = 0.99999994f;
    Sobol4 r;
    r.x = fminf(fract(t[0] + s0), top);
    r.y = fminf(fract(t[1] + s1), top);
    r.z = fminf(fract(t[2] + s2), top);
    r.w = fminf(fract(t[3] + s3), top);
    return r;
}

// The pattern table (sobol.rs:106-150): `length` points of each of the 256
// patterns, 4 floats per point, pattern-major. Host only.
void sobol_table(uint32_t length, float* out);
// generate_4d (sobol.rs:106-135) for one pattern of `count` points.
void sobol_pattern(float* points, uint32_t count, uint32_t seed);

}  // namespace pfx
