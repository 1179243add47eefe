// PFXCKPT1 checkpoint bytes (crates/trace/src/checkpoint.rs:8-158): the
// same little-endian layout, so checkpoints move between this tracer and
// the reference's still path.
//
//   "PFXCKPT1" | u64 key length | key | u32 width | u32 height | u32 samples
//   | accumulator (48 B/px) | color, albedo, normal (16 B/px each)
//   | u32 round count or 0xffffffff | per round u32 samples, u64 active
//   | snapshots (32 B/px, only with rounds)
//
// Host code only; no device is touched here.
#include <stdint.h>
#include <string.h>

#include <string>

#include "../../include/pfxhip.h"

namespace {

constexpr char MAGIC[8] = {'P', 'F', 'X', 'C', 'K', 'P', 'T', '1'};
constexpr uint64_t ACCUM = 48, TEXEL = 16, SNAP = 32;

void put_err(char* err, size_t err_len, const std::string& msg) {
    if (!err || err_len == 0) return;
    size_t n = msg.size() < err_len - 1 ? msg.size() : err_len - 1;
    memcpy(err, msg.data(), n);
    err[n] = 0;
}

struct Writer {
    uint8_t* out;
    size_t at = 0;
    void bytes(const void* p, size_t n) {
        if (out && n) memcpy(out + at, p, n);
        at += n;
    }
    void u32(uint32_t v) {
        uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        bytes(b, 4);
    }
    void u64(uint64_t v) {
        u32((uint32_t)v);
        u32((uint32_t)(v >> 32));
    }
};

struct Reader {
    const uint8_t* p;
    size_t len, at = 0;
    bool take(size_t n, const uint8_t** where) {
        if (n > len - at) return false;
        *where = p + at;
        at += n;
        return true;
    }
    bool u32(uint32_t* v) {
        const uint8_t* b;
        if (!take(4, &b)) return false;
        *v = (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24;
        return true;
    }
    bool u64(uint64_t* v) {
        uint32_t lo, hi;
        if (!u32(&lo) || !u32(&hi)) return false;
        *v = (uint64_t)lo | (uint64_t)hi << 32;
        return true;
    }
};

}  // namespace

extern "C" {

pfxh_status pfxh_checkpoint_write(const pfxh_checkpoint* c, const char* key, uint8_t* out, size_t cap, size_t* len,
                                  char* err, size_t err_len) {
    if (!c || !key || !len) {
        put_err(err, err_len, "null argument");
        return PFXH_INVALID;
    }
    uint64_t pixels = (uint64_t)c->width * c->height;
    bool rounds = c->round_count != PFXH_NO_ROUNDS;
    if (!c->accum || !c->color || !c->albedo || !c->normal || (rounds && (!c->snaps || (c->round_count &&
                                                                          (!c->round_samples || !c->round_active))))) {
        put_err(err, err_len, "the checkpoint's buffers do not match its size");
        return PFXH_INVALID;
    }
    size_t keylen = strlen(key);
    Writer sizing{nullptr};
    sizing.at = 8 + 8 + keylen + 12 + pixels * (ACCUM + 3 * TEXEL) + 4 +
                (rounds ? (uint64_t)c->round_count * 12 + pixels * SNAP : 0);
    *len = sizing.at;
    if (!out) return PFXH_OK;
    if (cap < sizing.at) {
        put_err(err, err_len, "the output buffer is smaller than the checkpoint");
        return PFXH_INVALID;
    }
    Writer w{out};
    w.bytes(MAGIC, 8);
    w.u64(keylen);
    w.bytes(key, keylen);
    w.u32(c->width);
    w.u32(c->height);
    w.u32(c->samples);
    w.bytes(c->accum, pixels * ACCUM);
    w.bytes(c->color, pixels * TEXEL);
    w.bytes(c->albedo, pixels * TEXEL);
    w.bytes(c->normal, pixels * TEXEL);
    if (!rounds) {
        w.u32(PFXH_NO_ROUNDS);
    } else {
        w.u32(c->round_count);
        for (uint32_t i = 0; i < c->round_count; ++i) {
            w.u32(c->round_samples[i]);
            w.u64(c->round_active[i]);
        }
        w.bytes(c->snaps, pixels * SNAP);
    }
    return PFXH_OK;
}

pfxh_status pfxh_checkpoint_read(const uint8_t* bytes, size_t len, const char* key, pfxh_checkpoint* out,
                                 uint32_t* round_samples, uint64_t* round_active, uint32_t round_cap, char* err,
                                 size_t err_len) {
    if (!bytes || !key || !out) {
        put_err(err, err_len, "null argument");
        return PFXH_INVALID;
    }
    memset(out, 0, sizeof *out);
    Reader r{bytes, len};
    const uint8_t* magic;
    const char* cut = "the checkpoint is cut short";
    if (!r.take(8, &magic)) {
        put_err(err, err_len, cut);
        return PFXH_INVALID;
    }
    if (memcmp(magic, MAGIC, 8) != 0) {
        put_err(err, err_len, "not a pfx checkpoint");
        return PFXH_INVALID;
    }
    uint64_t keylen;
    const uint8_t* stored;
    if (!r.u64(&keylen)) {
        put_err(err, err_len, cut);
        return PFXH_INVALID;
    }
    if (keylen != strlen(key)) {
        put_err(err, err_len, "the checkpoint was made by another render");
        return PFXH_INVALID;
    }
    if (!r.take(keylen, &stored)) {
        put_err(err, err_len, cut);
        return PFXH_INVALID;
    }
    if (memcmp(stored, key, keylen) != 0) {
        put_err(err, err_len, "the checkpoint was made by another render");
        return PFXH_INVALID;
    }
    if (!r.u32(&out->width) || !r.u32(&out->height) || !r.u32(&out->samples)) {
        put_err(err, err_len, cut);
        return PFXH_INVALID;
    }
    uint64_t pixels = (uint64_t)out->width * out->height;
    if (pixels > (uint64_t)len) {  // every pixel holds at least 96 bytes
        put_err(err, err_len, cut);
        return PFXH_INVALID;
    }
    if (!r.take(pixels * ACCUM, &out->accum) || !r.take(pixels * TEXEL, &out->color) ||
        !r.take(pixels * TEXEL, &out->albedo) || !r.take(pixels * TEXEL, &out->normal) ||
        !r.u32(&out->round_count)) {
        put_err(err, err_len, cut);
        return PFXH_INVALID;
    }
    if (out->round_count != PFXH_NO_ROUNDS) {
        for (uint32_t i = 0; i < out->round_count; ++i) {
            uint32_t s;
            uint64_t a;
            if (!r.u32(&s) || !r.u64(&a)) {
                put_err(err, err_len, cut);
                return PFXH_INVALID;
            }
            if (i < round_cap && round_samples && round_active) {
                round_samples[i] = s;
                round_active[i] = a;
            }
        }
        if (!r.take(pixels * SNAP, &out->snaps)) {
            put_err(err, err_len, cut);
            return PFXH_INVALID;
        }
        out->round_samples = round_samples;
        out->round_active = round_active;
    }
    if (r.at != len) {
        put_err(err, err_len, "the checkpoint runs on past its end");
        return PFXH_INVALID;
    }
    return PFXH_OK;
}

}  // extern "C"
