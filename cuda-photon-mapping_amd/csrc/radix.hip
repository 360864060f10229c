// Stable LSD radix sort of (u32 key, u32 value) pairs, 8-bit digits.
//
// Each pass is count -> scan -> stable scatter: a per-tile digit histogram,
// the wave64 look-back scan over the digit-major histogram table, and a
// scatter that ranks keys inside the tile with wave ballots, stages the tile
// in LDS in digit order and writes each digit's run contiguously. Stability
// makes the grid build's in-cell order (and the kd-tree's tie order) follow
// the input index, which the reference's atomicAdd slots do not guarantee.
//
// Several equal-length arrays laid end to end (segments) sort in the same
// launches: workgroup b takes tile b % tiles of segment b / tiles, and the
// histogram table is segment-major, so the scan's offsets place every
// segment's keys in its own range of the output.
#include "pm_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;                 // keys per lane
constexpr int kWaveKeys = 64 * kItems;     // 1024 keys per wave
constexpr int kTile = kThreads * kItems;   // 4096 keys per tile
constexpr int kBins = 256;

// n: the segment length; n_dev (nullable, one segment only): the element
// count on the device, at most n (the launch covers n; tiles past the device
// count contribute nothing).
__global__ __launch_bounds__(kThreads) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                     uint32_t mask, uint32_t* __restrict__ hist, uint32_t tiles,
                                                     const uint32_t* __restrict__ n_dev) {
    __shared__ uint32_t s_hist[kBins];
    const uint32_t t = threadIdx.x, seg = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const uint32_t* sk = keys + (size_t)seg * n;
    if (n_dev) n = min(n, *n_dev);
    s_hist[t] = 0;
    __syncthreads();
    const size_t base = (size_t)tile * kTile;
#pragma unroll 4
    for (int i = 0; i < kItems; ++i) {
        size_t idx = base + (size_t)i * kThreads + t;
        if (idx < n) atomicAdd(&s_hist[(sk[idx] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[((size_t)seg * kBins + t) * tiles + tile] = s_hist[t];
}

__global__ __launch_bounds__(kThreads) void k_rs_scatter(const uint32_t* __restrict__ keys_in,
                                                        const uint32_t* __restrict__ vals_in,
                                                        uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                        uint32_t n, int shift, uint32_t mask,
                                                        const uint32_t* __restrict__ offsets, uint32_t tiles,
                                                        const uint32_t* __restrict__ n_dev) {
    const uint32_t seg = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const size_t sbase = (size_t)seg * n;  // the segment's first element
    if (n_dev) n = min(n, *n_dev);
    if ((size_t)tile * kTile >= n) return;  // (block-uniform, before any barrier)
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t s_whist[kWaves][kBins];
    __shared__ uint32_t s_dstart[kBins];
    __shared__ uint32_t s_goff[kBins];
    __shared__ uint32_t s_scan[kWaves];

    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int w = 0; w < kWaves; ++w) s_whist[w][t] = 0;
    s_goff[t] = offsets[((size_t)seg * kBins + t) * tiles + tile];
    __syncthreads();

    const size_t wbase = (size_t)tile * kTile + (size_t)wave * kWaveKeys;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        size_t idx = wbase + (size_t)i * 64 + lane;
        bool valid = idx < n;
        key[i] = valid ? keys_in[sbase + idx] : 0u;
        val[i] = valid ? (vals_in ? vals_in[sbase + idx] : (uint32_t)idx) : 0u;
        uint32_t d = (key[i] >> shift) & mask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t r = (uint32_t)__popcll(peers & lt_mask);
        bool leader = valid && r == 0;
        if (leader) s_whist[wave][d] += cnt;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        uint32_t after = valid ? s_whist[wave][d] : 0u;
        rank[i] = valid ? (after - cnt + r) : 0xFFFFFFFFu;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Per digit (thread t owns digit t): counts across waves, then the
    // tile-local start of each digit (exclusive scan over the 256 digits).
    uint32_t c = 0;
    uint32_t wex[kWaves];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        wex[w] = c;
        c += s_whist[w][t];
    }
    uint32_t inc = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += y;
    }
    if (lane == 63) s_scan[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += s_scan[w];
    uint32_t dstart = before + inc - c;
    s_dstart[t] = dstart;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_whist[w][t] = dstart + wex[w];
    __syncthreads();

#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        if (rank[i] != 0xFFFFFFFFu) {
            uint32_t d = (key[i] >> shift) & mask;
            uint32_t p = s_whist[wave][d] + rank[i];
            s_keys[p] = key[i];
            s_vals[p] = val[i];
        }
    }
    __syncthreads();

    const size_t tbase = (size_t)tile * kTile;
    const uint32_t tcount = (uint32_t)min((size_t)kTile, (size_t)n - tbase);
    for (uint32_t p = t; p < tcount; p += kThreads) {
        uint32_t k = s_keys[p];
        uint32_t d = (k >> shift) & mask;
        uint32_t dst = s_goff[d] + (p - s_dstart[d]);
        keys_out[dst] = k;
        vals_out[dst] = s_vals[p];
    }
}

}  // namespace

namespace pm {

size_t radix_sort_scratch_bytes(uint32_t n, uint32_t segs) {
    const uint32_t tiles = div_up(n ? n : 1, kTile);
    const size_t all = (size_t)n * segs;
    return carve_size({all * 4, all * 4, (size_t)segs * kBins * tiles * 4, (size_t)segs * kBins * tiles * 4});
}

pm_status radix_sort_pairs(pm_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                           uint32_t* vals_out, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int key_bits,
                           uint32_t* hist, uint32_t* offs, const uint32_t* n_dev, uint32_t segs) {
    if (n == 0 || segs == 0) return PM_OK;
    if (segs > 1 && n_dev) return fail(ctx, PM_ERR_INVALID, "radix sort: a device count sorts one segment");
    if (key_bits < 1) key_bits = 1;
    if (key_bits > 32) key_bits = 32;
    const uint32_t tiles = div_up(n, kTile);
    const size_t table = (size_t)segs * kBins * tiles;
    if (table > 0xFFFFFFFFull || (uint64_t)n * segs > 0xFFFFFFFFull)
        return fail(ctx, PM_ERR_UNSUPPORTED, "radix sort: %u x %u keys", segs, n);
    const int passes = (key_bits + 7) / 8;
    const uint32_t* src_k = keys_in;
    const uint32_t* src_v = vals_in;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const int bits = std::min(8, key_bits - shift);
        const uint32_t mask = (1u << bits) - 1u;
        // ping-pong so that the last pass lands in *_out
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* dk = to_out ? keys_out : keys_tmp;
        uint32_t* dv = to_out ? vals_out : vals_tmp;
        hipLaunchKernelGGL(k_rs_hist, dim3(segs * tiles), dim3(kThreads), 0, ctx->stream, src_k, n, shift, mask, hist,
                           tiles, n_dev);
        PM_CHECK_LAUNCH(ctx, "k_rs_hist");
        pm_status s = scan_u32(ctx, hist, offs, (uint32_t)table, false, nullptr);
        if (s != PM_OK) return s;
        hipLaunchKernelGGL(k_rs_scatter, dim3(segs * tiles), dim3(kThreads), 0, ctx->stream, src_k, src_v, dk, dv, n,
                           shift, mask, offs, tiles, n_dev);
        PM_CHECK_LAUNCH(ctx, "k_rs_scatter");
        src_k = dk;
        src_v = dv;
    }
    return PM_OK;
}

// Sort using a caller-provided work block of radix_sort_scratch_bytes(n, segs).
pm_status radix_sort_pairs_auto(pm_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                                uint32_t* vals_out, uint32_t n, int key_bits, void* work, const uint32_t* n_dev,
                                uint32_t segs) {
    const uint32_t tiles = div_up(n ? n : 1, kTile);
    const size_t all = (size_t)n * segs;
    Carve cv{static_cast<char*>(work)};
    uint32_t* kt = cv.take<uint32_t>(all);
    uint32_t* vt = cv.take<uint32_t>(all);
    uint32_t* hist = cv.take<uint32_t>((size_t)segs * kBins * tiles);
    uint32_t* offs = cv.take<uint32_t>((size_t)segs * kBins * tiles);
    return radix_sort_pairs(ctx, keys_in, vals_in, keys_out, vals_out, kt, vt, n, key_bits, hist, offs, n_dev, segs);
}

}  // namespace pm
