// CPU oracle — shared FP32 vector helpers.
//
// TEST INFRASTRUCTURE ONLY. This directory restates the reference's host
// (CPU) arithmetic so the HIP product can be checked against it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
// product library never links, calls or falls back to anything here.
//
// Rounding: every operation is IEEE FP32 round-to-nearest with no contraction
// (built with -ffp-contract=off), which is what the reference's host branches
// of vec3 compute (CudaPhotonMapping/vec3.cuh, the non-USE_INTRINSICS paths).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct V3 {
    float x, y, z;
};

static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 mul(V3 a, float f) { return V3{a.x * f, a.y * f, a.z * f}; }
static inline V3 mulv(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// vec3::dot host branch (vec3.cuh:258-274): x*x' + y*y' + z*z', left to right.
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vec3::cross host branch (vec3.cuh:276-300).
static inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x};
}
// vec3::length host branch (vec3.cuh:171-184).
static inline float length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline bool is_null(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
// static vec3::normalize (vec3.cuh:226-240): component-wise division by length.
static inline V3 normalize_div(V3 a) {
    if (is_null(a)) return a;
    float l = length(a);
    return V3{a.x / l, a.y / l, a.z / l};
}
// member vec3::normalize (vec3.cuh:211-224 -> operator/= :156-169): multiply
// by the reciprocal of the length.
static inline V3 normalize_rcp(V3 a) {
    if (is_null(a)) return a;
    float u = 1.0f / length(a);
    return V3{a.x * u, a.y * u, a.z * u};
}
static inline float maxc(V3 a) { return std::fmax(std::fmax(a.x, a.y), a.z); }

static inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11). Multipliers and
// Weyl key increments are the published constants of that paper.
struct Philox {
    static inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
        uint64_t p = (uint64_t)a * (uint64_t)b;
        hi = (uint32_t)(p >> 32);
        lo = (uint32_t)p;
    }
    static inline void block(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
        uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
        for (int r = 0; r < 10; ++r) {
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, c0, hi0, lo0);
            mulhilo(0xCD9E8D57u, c2, hi1, lo1);
            uint32_t n0 = hi1 ^ c1 ^ k0;
            uint32_t n1 = lo1;
            uint32_t n2 = hi0 ^ c3 ^ k1;
            uint32_t n3 = lo0;
            c0 = n0; c1 = n1; c2 = n2; c3 = n3;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
    }
};

// Uniform float in (0, 1]: 24 random bits, +1, times 2^-24. Exactly
// representable, so host and device produce identical values.
static inline float u01(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

// One random stream per (seed, stream id, path id): successive calls consume
// successive Philox blocks, four uniforms per block.
struct Stream {
    uint32_t k0, k1, path, sid, blk, pos;
    uint32_t buf[4];
    void init(uint64_t seed, uint32_t stream_id, uint32_t path_id) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        path = path_id;
        sid = stream_id;
        blk = 0;
        pos = 4;
    }
    float next() {
        if (pos == 4) {
            uint32_t ctr[4] = {path, blk, sid, 0x7068u};
            Philox::block(ctr, k0, k1, buf);
            ++blk;
            pos = 0;
        }
        return u01(buf[pos++]);
    }
};

}  // namespace orc
