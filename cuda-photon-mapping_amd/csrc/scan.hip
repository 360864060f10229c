// Single-pass u32 prefix scan: wave64 shuffle scan inside a 4096-element
// tile, decoupled look-back across tiles (Merrill & Garland 2016). Tile
// words carry an epoch held in device memory (read by every tile, advanced
// by the last), so consecutive launches need no clearing pass and a launch
// replayed from a captured graph sees a fresh epoch. Replaces
// the reference's 44-block Blelloch scan with its software grid barrier
// (PrefixScanSum.cuh:12-279, CudaGridSynchronizer.cuh:14-27), which needs all
// blocks co-resident; here a tile only waits on tiles that took a lower
// ordered id, so any number of tiles may be in flight. Inputs of 4M
// elements and more take a two-level scan instead (tile sums, a scan of the
// sums, tiles rescanned from their offsets), whose workgroups never wait on
// each other.
#include "pm_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // 4096

constexpr uint64_t kStatusAgg = 1, kStatusIncl = 2;
// inputs of at least this many elements take the two-level scan
constexpr uint32_t kTwoLevelMin = 1u << 22;

__device__ __forceinline__ uint64_t pack_state(uint32_t epoch, uint64_t status, uint32_t value) {
    return ((uint64_t)epoch << 34) | (status << 32) | value;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

template <bool kInclusive>
__global__ __launch_bounds__(kThreads) void k_scan_u32(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        uint32_t n, uint64_t* __restrict__ state,
                                                        uint32_t* __restrict__ counter, uint32_t n_tiles,
                                                        uint64_t* __restrict__ total) {
    __shared__ uint32_t s_tile, s_epoch;
    __shared__ uint32_t s_wave[kThreads / 64];
    __shared__ uint32_t s_prefix;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) {
        s_epoch = pm::lookback_epoch(counter);
        uint32_t id = atomicAdd(counter, 1u);
        // Every tile has taken its id once the last id is handed out: reset
        // the counter for the next launch on this context.
        if (id == n_tiles - 1) atomicExch(counter, 0u);
        s_tile = id;
    }
    __syncthreads();
    const uint32_t tile = s_tile, epoch = s_epoch;
    const size_t base = (size_t)tile * kTile + (size_t)t * kItems;

    uint32_t v[kItems];
    const bool full = base + kItems <= n;
    if (full && ((reinterpret_cast<uintptr_t>(in) & 15) == 0)) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int j = 0; j < kItems / 4; ++j) {
            uint4 q = p[j];
            v[4 * j] = q.x;
            v[4 * j + 1] = q.y;
            v[4 * j + 2] = q.z;
            v[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j) v[j] = (base + j < n) ? in[base + j] : 0u;
    }
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) tsum += v[j];

    uint32_t incl = wave_incl_scan(tsum, lane);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_excl = 0, agg = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        uint32_t s = s_wave[w];
        if ((uint32_t)w < wave) wave_excl += s;
        agg += s;
    }

    if (wave == 0) {
        uint32_t prefix = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&state[0], pack_state(epoch, kStatusIncl, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&state[tile], pack_state(epoch, kStatusAgg, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            int64_t j = (int64_t)tile - 1;
            for (;;) {
                int64_t idx = j - (int64_t)lane;
                uint64_t w;
                uint32_t status;
                uint64_t incl_mask, inval_mask, upto;
                for (;;) {
                    if (idx >= 0) {
                        w = __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        status = ((uint32_t)(w >> 34) == epoch) ? (uint32_t)((w >> 32) & 3) : 0u;
                    } else {
                        w = 0;
                        status = (uint32_t)kStatusIncl;
                    }
                    incl_mask = __ballot(status == kStatusIncl);
                    inval_mask = __ballot(status == 0);
                    // lanes 0..first-inclusive must all be published
                    int first = incl_mask ? __builtin_ctzll(incl_mask) : 63;
                    upto = (first == 63) ? ~0ull : ((2ull << first) - 1);
                    if ((inval_mask & upto) == 0) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                uint32_t contrib = ((upto >> lane) & 1ull) ? (uint32_t)w : 0u;
                // wave reduction
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) contrib += __shfl_xor(contrib, d, 64);
                prefix += contrib;
                if (incl_mask & upto) break;
                j -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&state[tile], pack_state(epoch, kStatusIncl, prefix + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_prefix = prefix;
            if (tile == n_tiles - 1) {
                if (total) *total = (uint64_t)(uint32_t)(prefix + agg);
                // every tile has published (the look-back above reached an
                // inclusive tile through all of them), so every tile has
                // read this launch's epoch: advance it for the next launch
                pm::lookback_advance(counter, epoch);
            }
        }
    }
    __syncthreads();

    uint32_t run = s_prefix + wave_excl + (incl - tsum);
    uint32_t o[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (kInclusive) {
            run += v[j];
            o[j] = run;
        } else {
            o[j] = run;
            run += v[j];
        }
    }
    if (full && ((reinterpret_cast<uintptr_t>(out) & 15) == 0)) {
        uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int j = 0; j < kItems / 4; ++j) p[j] = make_uint4(o[4 * j], o[4 * j + 1], o[4 * j + 2], o[4 * j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (base + j < n) out[base + j] = o[j];
    }
}

__global__ void k_zero_u64(uint64_t* p) { *p = 0; }

// ---- two-level scan for large inputs: tile sums -> scan of the sums ->
// each tile rescanned from its offset. Three passes and no waiting between
// workgroups: past a few million elements the single pass is held up by
// its ordered tile counter, one device-scope atomic per tile.
__global__ __launch_bounds__(kThreads) void k_tile_sums(const uint32_t* __restrict__ in, uint32_t n,
                                                       uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_wave[kThreads / 64];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const size_t base = (size_t)blockIdx.x * kTile + (size_t)t * kItems;
    uint32_t sum = 0;
    if (base + kItems <= n && ((reinterpret_cast<uintptr_t>(in) & 15) == 0)) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int j = 0; j < kItems / 4; ++j) {
            const uint4 q = p[j];
            sum += q.x + q.y + q.z + q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j) sum += (base + j < n) ? in[base + j] : 0u;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if (lane == 0) s_wave[wave] = sum;
    __syncthreads();
    if (t == 0) {
        uint32_t a = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) a += s_wave[w];
        sums[blockIdx.x] = a;
    }
}

template <bool kInclusive>
__global__ __launch_bounds__(kThreads) void k_tile_rescan(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                         uint32_t n, const uint32_t* __restrict__ offs) {
    __shared__ uint32_t s_wave[kThreads / 64];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const size_t base = (size_t)blockIdx.x * kTile + (size_t)t * kItems;
    const bool full = base + kItems <= n;
    uint32_t v[kItems];
    if (full && ((reinterpret_cast<uintptr_t>(in) & 15) == 0)) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int j = 0; j < kItems / 4; ++j) {
            const uint4 q = p[j];
            v[4 * j] = q.x;
            v[4 * j + 1] = q.y;
            v[4 * j + 2] = q.z;
            v[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j) v[j] = (base + j < n) ? in[base + j] : 0u;
    }
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) tsum += v[j];
    const uint32_t incl = wave_incl_scan(tsum, lane);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t run = offs[blockIdx.x] + (incl - tsum);
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w)
        if ((uint32_t)w < wave) run += s_wave[w];
    uint32_t o[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (kInclusive) {
            run += v[j];
            o[j] = run;
        } else {
            o[j] = run;
            run += v[j];
        }
    }
    if (full && ((reinterpret_cast<uintptr_t>(out) & 15) == 0)) {
        uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int j = 0; j < kItems / 4; ++j) p[j] = make_uint4(o[4 * j], o[4 * j + 1], o[4 * j + 2], o[4 * j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (base + j < n) out[base + j] = o[j];
    }
}

}  // namespace

namespace pm {

pm_status lookback_begin(pm_ctx* ctx, uint32_t tiles, uint64_t** state, uint32_t** counter) {
    if (ctx->fixed_state) {
        if (tiles > ctx->fixed_state_tiles)
            return fail(ctx, PM_ERR_UNSUPPORTED, "look-back scan of %u tiles exceeds the fixed state of %zu tiles",
                        tiles, ctx->fixed_state_tiles);
        *state = ctx->fixed_state;
        *counter = ctx->fixed_counter;
        return PM_OK;
    }
    if (tiles > ctx->scan_state_tiles) {
        // (growth is the only host-side step: warm a context up before
        // capturing its calls into a graph)
        PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->scan_state) PM_HIP(ctx, hipFree(ctx->scan_state));
        ctx->scan_state = nullptr;
        size_t want = (size_t)tiles + tiles / 2 + 64;
        PM_HIP(ctx, hipMalloc(&ctx->scan_state, want * sizeof(uint64_t)));
        ctx->scan_state_tiles = want;
        // zeroed on the context's stream, ahead of the scan that uses it
        // (a hipMemset runs on the null stream, which a non-blocking stream
        // does not wait for: a tile word cleared after its tile published
        // would leave the next tile waiting for good)
        pm_status st = fill_u32(ctx, reinterpret_cast<uint32_t*>(ctx->scan_state), 0u, 2 * want);
        if (st != PM_OK) return st;
    }
    *state = ctx->scan_state;
    *counter = ctx->scan_counter;
    return PM_OK;
}

pm_status scan_u32(pm_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, uint64_t* total) {
    if (n == 0) {
        if (total) {
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, ctx->stream, total);
            PM_CHECK_LAUNCH(ctx, "k_zero_u64");
        }
        return PM_OK;
    }
    uint32_t tiles = div_up(n, kTile);
    if (n >= kTwoLevelMin && !ctx->fixed_state) {
        // tile sums and their offsets in the context's own block (callers'
        // scratch carves stay untouched)
        if (2 * (size_t)tiles > ctx->tile_sums_cap) {
            PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if (ctx->tile_sums) PM_HIP(ctx, hipFree(ctx->tile_sums));
            ctx->tile_sums = nullptr;
            const size_t want = 2 * (size_t)tiles + tiles;
            PM_HIP(ctx, hipMalloc(&ctx->tile_sums, want * 4));
            ctx->tile_sums_cap = want;
        }
        uint32_t* sums = ctx->tile_sums;
        uint32_t* offs = ctx->tile_sums + tiles;
        hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kThreads), 0, ctx->stream, in, n, sums);
        PM_CHECK_LAUNCH(ctx, "k_tile_sums");
        pm_status st = scan_u32(ctx, sums, offs, tiles, false, total);
        if (st != PM_OK) return st;
        if (inclusive)
            hipLaunchKernelGGL(k_tile_rescan<true>, dim3(tiles), dim3(kThreads), 0, ctx->stream, in, out, n, offs);
        else
            hipLaunchKernelGGL(k_tile_rescan<false>, dim3(tiles), dim3(kThreads), 0, ctx->stream, in, out, n, offs);
        PM_CHECK_LAUNCH(ctx, "k_tile_rescan");
        return PM_OK;
    }
    uint64_t* state;
    uint32_t* counter;
    pm_status st = lookback_begin(ctx, tiles, &state, &counter);
    if (st != PM_OK) return st;
    if (inclusive)
        hipLaunchKernelGGL(k_scan_u32<true>, dim3(tiles), dim3(kThreads), 0, ctx->stream, in, out, n, state, counter,
                           tiles, total);
    else
        hipLaunchKernelGGL(k_scan_u32<false>, dim3(tiles), dim3(kThreads), 0, ctx->stream, in, out, n, state, counter,
                           tiles, total);
    PM_CHECK_LAUNCH(ctx, "k_scan_u32");
    return PM_OK;
}

}  // namespace pm

extern "C" pm_status pm_scan_u32(pm_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint32_t n, int32_t inclusive,
                                 uint64_t* d_total) {
    if (!ctx) return PM_ERR_INVALID;
    if ((!d_in || !d_out) && n > 0) return pm::fail(ctx, PM_ERR_INVALID, "pm_scan_u32: null buffer");
    return pm::scan_u32(ctx, d_in, d_out, n, inclusive != 0, d_total);
}
