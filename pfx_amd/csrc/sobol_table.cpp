// The Sobol sampler's pattern table, built on the host (sobol.rs:40-150):
// progressive multi-jittered (0,2) points whose strata follow the XOR
// table of each axis, jittered by Cycles' hash_hp integer hash.
#include <algorithm>
#include <thread>
#include <vector>

#include "sobol.h"

namespace pfx {
namespace {

// sobol.rs:41-70: per axis, the partner index of each new point's stratum
const uint32_t XORS[4][32] = {
    {0},
    {0x00000000, 0x00000001, 0x00000001, 0x00000007, 0x00000001, 0x00000013, 0x00000015, 0x0000007f,
     0x00000001, 0x00000103, 0x00000105, 0x0000070f, 0x00000111, 0x00001333, 0x00001555, 0x00007fff,
     0x00000001, 0x00010003, 0x00010005, 0x0007000f, 0x00010011, 0x00130033, 0x00150055, 0x007f00ff,
     0x00010101, 0x01030303, 0x01050505, 0x070f0f0f, 0x01111111, 0x13333333, 0x15555555, 0x7fffffff},
    {0x00000000, 0x00000001, 0x00000003, 0x00000001, 0x00000005, 0x0000001f, 0x0000002b, 0x0000003d,
     0x00000011, 0x00000133, 0x00000377, 0x00000199, 0x00000445, 0x00001ccf, 0x00002ddb, 0x0000366d,
     0x00000101, 0x00010303, 0x00030707, 0x00010909, 0x00051515, 0x001f3f3f, 0x002b6b6b, 0x003dbdbd,
     0x00101011, 0x01303033, 0x03707077, 0x01909099, 0x04515145, 0x1cf3f3cf, 0x2db6b6db, 0x36dbdb6d},
    {0x00000000, 0x00000001, 0x00000000, 0x00000003, 0x0000000d, 0x0000000c, 0x00000005, 0x0000004f,
     0x00000014, 0x000000e7, 0x00000329, 0x0000039c, 0x00000011, 0x00001033, 0x00000044, 0x000030bb,
     0x0000d1cd, 0x0000c2ec, 0x00005415, 0x0004fc3f, 0x00015054, 0x000e5c97, 0x0032e5b9, 0x0039725c,
     0x00000101, 0x01000303, 0x00000404, 0x03000b0b, 0x0d001d1d, 0x0c002c2c, 0x05004545, 0x4f00cfcf},
};

uint32_t hash_hp(uint32_t i) {  // sobol.rs:72-79
    i ^= i >> 16;
    i *= 0x21f0aaadu;
    i ^= i >> 15;
    i *= 0xd35a2d97u;
    i ^= i >> 15;
    return i ^ 0xe6fe3bebu;
}
uint32_t hash_hp_seeded(uint32_t i, uint32_t seed) {  // :81-84
    seed ^= seed << 19;
    return hash_hp(i ^ seed);
}
float hash_hp_float(uint32_t i) { return (float)hash_hp(i) * (1.0f / 4294967808.0f); }  // :86-88

}  // namespace

void sobol_pattern(float* points, uint32_t size, uint32_t seed) {
    if (size == 0) return;
    uint32_t rng = hash_hp_seeded(seed, 0x44605a73u);
    auto next = [&] {
        float v = hash_hp_float(rng);
        rng += 1u;
        return v;
    };
    for (int a = 0; a < 4; ++a) points[a] = next();
    uint32_t log_n = 0, n = 1;
    while (n < size) {
        float strata = (float)(n * 2u);
        for (uint32_t i = 0; i < n && n + i < size; ++i) {
            for (int axis = 0; axis < 4; ++axis) {
                uint32_t partner = i ^ XORS[axis][log_n];
                uint32_t occupied = (uint32_t)(points[4 * partner + axis] * strata);
                points[4 * (n + i) + axis] = ((float)(occupied ^ 1u) + next()) / strata;
            }
        }
        log_n += 1;
        n *= 2u;
    }
}

void sobol_table(uint32_t length, float* out) {
    unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([=] {
            for (uint32_t p = t; p < SOBOL_PATTERNS; p += threads) sobol_pattern(out + 4 * (size_t)p * length, length, p);
        });
    for (auto& th : pool) th.join();
}

}  // namespace pfx
