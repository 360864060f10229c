// Wave64 selection of the k smallest (d2, index) keys, shared by the grid
// gather (knn.hip) and the kd-tree gather (kdknn.hip): an LDS key buffer
// filled with ballot + mbcnt, narrowed by float-bit histograms when it fills,
// and a register bitonic sort across the wave.
#pragma once
#include <cmath>

#include "pm_internal.hpp"

namespace pms {

// Largest float T with sqrtf(T) <= r: `d2 <= T` is then exactly the
// reference's `length(p - q) <= r` test (PhotonGrid.cu:246-249) on the same
// FP32 squared distance, with no square root per candidate.
inline float radius_threshold(float r) {
    float t = r * r;
    while (std::sqrt(t) > r) t = std::nextafter(t, 0.f);
    for (;;) {
        float u = std::nextafter(t, INFINITY);
        if (std::sqrt(u) <= r)
            t = u;
        else
            break;
    }
    return t;
}

template <int CAP>
struct WaveLds {
    uint32_t hist[64];
    uint64_t buf[CAP];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive wave64 prefix sum in registers (DPP, no LDS): shifts right by
// 1, 2, 4, 8 lanes inside each row of 16 (a lane with no source adds 0), then
// row 0's last lane into row 1 and row 2's into row 3 (row_bcast:15), then
// lane 31 into rows 2 and 3 (row_bcast:31).
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_add(uint32_t v) {
    return v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, kRowMask, 0xf, false);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t) {
    v = dpp_add<0x111, 0xf>(v);  // row_shr:1
    v = dpp_add<0x112, 0xf>(v);  // row_shr:2
    v = dpp_add<0x114, 0xf>(v);  // row_shr:4
    v = dpp_add<0x118, 0xf>(v);  // row_shr:8
    v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Wave64 reductions of floats by the same DPP steps, read back from lane 63
// (wave-uniform): the sum is associated as the scan adds it; the maximum is
// of nonnegative values (a lane with no source contributes 0).
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_src(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask, 0xf, false));
}
__device__ __forceinline__ float wave_sum_f32(float v) {
    v += dpp_src<0x111, 0xf>(v);
    v += dpp_src<0x112, 0xf>(v);
    v += dpp_src<0x114, 0xf>(v);
    v += dpp_src<0x118, 0xf>(v);
    v += dpp_src<0x142, 0xa>(v);
    v += dpp_src<0x143, 0xc>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_max_nonneg_f32(float v) {
    v = fmaxf(v, dpp_src<0x111, 0xf>(v));
    v = fmaxf(v, dpp_src<0x112, 0xf>(v));
    v = fmaxf(v, dpp_src<0x114, 0xf>(v));
    v = fmaxf(v, dpp_src<0x118, 0xf>(v));
    v = fmaxf(v, dpp_src<0x142, 0xa>(v));
    v = fmaxf(v, dpp_src<0x143, 0xc>(v));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t lo = __shfl_xor((uint32_t)v, m, 64);
    uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Lane mask of a predicate (the ballot builtin on the i1 directly, without
// the int round trip of __ballot).
__device__ __forceinline__ uint64_t wballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }

__device__ __forceinline__ uint32_t mbcnt(uint64_t bal) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
}

// Squared distance in the reference's host order, never contracted to FMA.
__device__ __forceinline__ float dist2(float px, float py, float pz, float qx, float qy, float qz) {
    float dx = px - qx, dy = py - qy, dz = pz - qz;
    float xx = __fmul_rn(dx, dx), yy = __fmul_rn(dy, dy), zz = __fmul_rn(dz, dz);
    return __fadd_rn(__fadd_rn(xx, yy), zz);
}

__device__ __forceinline__ uint64_t make_key(float d2, uint32_t slot) {
    return ((uint64_t)__float_as_uint(d2) << 32) | slot;
}

// Narrow the bucket of keys kept from LDS buffer entries [0, cnt) — all of
// which are < khi — to the smallest bin edge that keeps >= k keys and at most
// `target` when the keys allow, then compact the buffer to those keys.
// Round 0 bins by the top float bits of d2 (two bins per octave below the
// radius threshold); later rounds split the crossing bin 64 ways.
template <int CAP>
__device__ uint32_t buf_select(WaveLds<CAP>& L, uint32_t lane, uint32_t cnt, uint32_t k, uint32_t target, uint32_t bT,
                               uint64_t& khi_io) {
    if (cnt <= target || cnt <= k) return cnt;
    L.hist[lane] = 0;
    wave_sync();
    for (uint32_t e = lane; e < cnt; e += 64) {
        uint32_t h = bT - (uint32_t)(L.buf[e] >> 54);
        atomicAdd(&L.hist[h < 63u ? h : 63u], 1u);
    }
    wave_sync();
    uint32_t cum = wave_incl_scan(L.hist[63 - lane], lane);
    int ls = __builtin_ctzll(wballot(cum >= k));
    uint32_t hstar = 63u - (uint32_t)ls;
    // (ls is wave-uniform: read the lanes directly)
    uint32_t A = __builtin_amdgcn_readlane(cum, ls);
    uint32_t prev = __builtin_amdgcn_readlane(cum, ls > 0 ? ls - 1 : 0);
    uint64_t khi = (uint64_t)(bT - hstar + 1u) << 54, klo = 0;
    uint32_t below = 0;
    if (hstar < 63u) {
        klo = (uint64_t)(bT - hstar) << 54;
        below = ls > 0 ? prev : 0u;
    }
    if (khi > khi_io) khi = khi_io;
    while (A > target) {
        uint64_t width = khi - klo;
        int s = width <= 64 ? 0 : 64 - __builtin_clzll(width - 1) - 6;
        L.hist[lane] = 0;
        wave_sync();
        for (uint32_t e = lane; e < cnt; e += 64) {
            uint64_t key = L.buf[e];
            if (key >= klo && key < khi) atomicAdd(&L.hist[(uint32_t)((key - klo) >> s)], 1u);
        }
        wave_sync();
        uint32_t cu = wave_incl_scan(L.hist[lane], lane);
        uint32_t need = k - below;
        int js = __builtin_ctzll(wballot(cu >= need));
        uint32_t pv = __builtin_amdgcn_readlane(cu, js > 0 ? js - 1 : 0);
        uint32_t at = __builtin_amdgcn_readlane(cu, js);
        uint64_t nlo = klo + ((uint64_t)js << s);
        uint64_t nhi = klo + ((uint64_t)(js + 1) << s);
        if (nhi > khi) nhi = khi;
        A = below + at;
        below += js > 0 ? pv : 0u;
        klo = nlo;
        khi = nhi;
    }
    // in-place compaction, one row of 64 at a time (writes never pass reads)
    uint32_t out = 0;
    for (uint32_t base = 0; base < cnt; base += 64) {
        uint32_t e = base + lane;
        uint64_t key = e < cnt ? L.buf[e] : ~0ull;
        wave_sync();
        bool keep = e < cnt && key < khi;
        uint64_t bal = wballot(keep);
        if (keep) L.buf[out + mbcnt(bal)] = key;
        out += (uint32_t)__popcll(bal);
        wave_sync();
    }
    khi_io = khi;
    return out;
}

// buf_select to exactly the k smallest keys, for at most 64 R buffered keys
// read into registers once (v[j] = entry 64 j + lane, ~0 past cnt): rounds
// of bins like buf_select's (finer in the first), counted from the
// registers instead of re-reading the buffer; the kept set is the same (the
// k smallest keys, which are unique). The kept
// keys are then written back to the buffer's front in buffer order (one
// pass, every read already done). Returns the count kept (min(cnt, k)).
template <int R, int CAP>
__device__ uint32_t reg_select(WaveLds<CAP>& L, uint32_t lane, const uint64_t (&v)[R], uint32_t cnt, uint32_t k,
                               uint32_t bT, uint64_t& khi_io) {
    if (cnt <= k) return cnt;
    // the first round bins four per octave (the top two mantissa bits of
    // d2; 2 bT + 1 bounds bits(T) >> 21), so the crossing bin holds fewer
    // keys for the rounds that split it
    const uint32_t bT2 = (bT << 1) | 1u;
    L.hist[lane] = 0;
    wave_sync();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        if ((uint32_t)j * 64 + lane < cnt) {
            const uint32_t h = bT2 - (uint32_t)(v[j] >> 53);
            atomicAdd(&L.hist[h < 63u ? h : 63u], 1u);
        }
    }
    wave_sync();
    const uint32_t cum = wave_incl_scan(L.hist[63 - lane], lane);
    const int ls = __builtin_ctzll(wballot(cum >= k));
    const uint32_t hstar = 63u - (uint32_t)ls;
    uint32_t A = __builtin_amdgcn_readlane(cum, ls);
    const uint32_t prev = __builtin_amdgcn_readlane(cum, ls > 0 ? ls - 1 : 0);
    uint64_t khi = (uint64_t)(bT2 - hstar + 1u) << 53, klo = 0;
    uint32_t below = 0;
    if (hstar < 63u) {
        klo = (uint64_t)(bT2 - hstar) << 53;
        below = ls > 0 ? prev : 0u;
    }
    if (khi > khi_io) khi = khi_io;
    while (A > k) {
        const uint64_t width = khi - klo;
        const int s = width <= 64 ? 0 : 64 - __builtin_clzll(width - 1) - 6;
        L.hist[lane] = 0;
        wave_sync();
#pragma unroll
        for (int j = 0; j < R; ++j)
            if (v[j] >= klo && v[j] < khi) atomicAdd(&L.hist[(uint32_t)((v[j] - klo) >> s)], 1u);
        wave_sync();
        const uint32_t cu = wave_incl_scan(L.hist[lane], lane);
        const uint32_t need = k - below;
        const int js = __builtin_ctzll(wballot(cu >= need));
        const uint32_t pv = __builtin_amdgcn_readlane(cu, js > 0 ? js - 1 : 0);
        const uint32_t at = __builtin_amdgcn_readlane(cu, js);
        const uint64_t nlo = klo + ((uint64_t)js << s);
        uint64_t nhi = klo + ((uint64_t)(js + 1) << s);
        if (nhi > khi) nhi = khi;
        A = below + at;
        below += js > 0 ? pv : 0u;
        klo = nlo;
        khi = nhi;
    }
    // the kept keys (invalid entries are ~0, never below the bound), in order
    uint32_t out = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const bool keep = v[j] < khi;
        const uint64_t bal = wballot(keep);
        if (keep) L.buf[out + mbcnt(bal)] = v[j];
        out += (uint32_t)__popcll(bal);
    }
    khi_io = khi;
    return out;
}

// Ascending bitonic sort of buf[0, n) in LDS by one wave; n is a power of two.
__device__ __forceinline__ void lds_bitonic_sort(uint64_t* buf, uint32_t n, uint32_t lane) {
    for (uint32_t size = 2; size <= n; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = lane; i < n / 2; i += 64) {
                const uint32_t a = 2 * stride * (i / stride) + (i % stride), b = a + stride;
                const bool asc = (a & size) == 0;
                const uint64_t x = buf[a], y = buf[b];
                if (asc ? (x > y) : (x < y)) {
                    buf[a] = y;
                    buf[b] = x;
                }
            }
            wave_sync();
        }
    }
}

template <int R>
__device__ __forceinline__ void bitonic_sort(uint64_t (&v)[R], uint32_t lane) {
    constexpr int N = 64 * R;
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= 64) {
                const int rs = stride / 64;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    if ((j & rs) == 0) {
                        const int jp = j | rs;
                        const bool asc = (((j * 64) & size) == 0);
                        uint64_t a = v[j], b = v[jp];
                        bool sw = asc ? (a > b) : (a < b);
                        v[j] = sw ? b : a;
                        v[jp] = sw ? a : b;
                    }
                }
            } else {
                const bool lower = (lane & (uint32_t)stride) == 0;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const uint32_t e = (uint32_t)j * 64 + lane;
                    const bool asc = (e & (uint32_t)size) == 0;
                    uint64_t p = shfl_xor_u64(v[j], stride);
                    bool take_min = (asc == lower);
                    uint64_t mn = v[j] < p ? v[j] : p;
                    uint64_t mx = v[j] < p ? p : v[j];
                    v[j] = take_min ? mn : mx;
                }
            }
        }
    }
}

// Sort the cnt (<= 64 R) buffered keys and write the first k as sorted
// (index, distance) rows; m = number of valid results (<= k).
template <int R, int CAP>
__device__ __forceinline__ void sort_keys(WaveLds<CAP>& L, uint32_t lane, uint32_t cnt, uint64_t (&v)[R]) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
        uint32_t e = (uint32_t)j * 64 + lane;
        v[j] = e < cnt ? L.buf[e] : ~0ull;
    }
    bitonic_sort<R>(v, lane);
}

template <int R>
__device__ __forceinline__ void write_rows(const uint64_t (&v)[R], uint32_t lane, uint32_t m, int k,
                                           uint32_t* __restrict__ idx_row, float* __restrict__ dist_row) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
        uint32_t e = (uint32_t)j * 64 + lane;
        if (e < (uint32_t)k) {
            bool keep = e < m;
            idx_row[e] = keep ? (uint32_t)v[j] : 0xFFFFFFFFu;
            dist_row[e] = keep ? sqrtf(__uint_as_float((uint32_t)(v[j] >> 32))) : __builtin_inff();
        }
    }
    // slots beyond the sorted registers (k > 64*R >= cnt >= m) are empty
    for (uint32_t e = 64u * R + lane; e < (uint32_t)k; e += 64) {
        idx_row[e] = 0xFFFFFFFFu;
        dist_row[e] = __builtin_inff();
    }
}

}  // namespace pms
