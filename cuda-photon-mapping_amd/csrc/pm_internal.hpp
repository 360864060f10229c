// Internal declarations shared by the pm355 HIP sources (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pm355.h"

#define PM_WAVE 64

namespace pm {
// One remembered gather tuning: k, photon density class, choice.
struct TuneEntry {
    int k;
    int cls;
    pm_gather_tuning t;
};
}  // namespace pm

struct pm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // Growable device scratch arena (carved per call; never shared between
    // two calls in flight because one context drives one stream).
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // Decoupled look-back state for pm scans: tile flags carry an epoch so
    // no clearing pass is needed between launches.
    uint64_t* scan_state = nullptr;
    size_t scan_state_tiles = 0;
    // [0]: ordered tile ids, reset by each launch's last tile; [32]: the
    // epoch of the next look-back launch (device-side, starts at 1)
    uint32_t* scan_counter = nullptr;
    // tile sums and offsets of the two-level scan (large inputs)
    uint32_t* tile_sums = nullptr;
    size_t tile_sums_cap = 0;
    int cu_count = 256;
    // Fixed device memory that replaces the scratch arena and the look-back
    // state for the duration of one call (the allocation-free entry point):
    // while `fixed_arena` is set, scratch_reserve and lookback_begin hand
    // these out and fail instead of allocating when they are too small.
    void* fixed_arena = nullptr;
    size_t fixed_arena_bytes = 0;
    uint64_t* fixed_state = nullptr;
    size_t fixed_state_tiles = 0;
    uint32_t* fixed_counter = nullptr;
    // Frame state of pm_render_frame (hit slots and query block survive the
    // gathers, which use `scratch`).
    pm_frame* frame_state = nullptr;
    // Device blocks of destroyed grids, reused by later builds on this
    // context (in stream order) instead of hipMalloc / hipFree per grid.
    std::vector<std::pair<void*, size_t>> block_cache;
    size_t block_cache_bytes = 0;
    size_t block_cache_cap = 0;  // a quarter of the device's memory
    // Gather tunings measured by pm_tune_gather on this context, and the
    // last tuning's table of measurements (cell sizes x variants, ms).
    std::vector<pm::TuneEntry> tunes;
    std::vector<float> tune_table;
};

struct pm_grid {
    pm_grid_desc desc{};
    void* mem = nullptr;  // one allocation backing starts/sizes/photons/perm
    size_t mem_bytes = 0;
    uint32_t n_cap = 0;   // photons the allocation holds (pm_grid_rebuild)
    pm_ctx* owner = nullptr;  // the context whose block cache takes `mem` back
};

struct pm_kdtree {
    uint32_t n = 0, cap = 0, levels = 0;
    void* mem = nullptr;
    size_t mem_bytes = 0;
    pm_ctx* owner = nullptr;  // the context whose block cache `mem` returns to
    uint32_t* node_idx = nullptr;  // [cap]
    int8_t* node_plane = nullptr;  // [cap]
    float* photons = nullptr;      // photons9 in node order [9][cap] (empty slots hold +inf positions)
    // one 16-byte record per node for the kNN descent: position and
    // (photon index | split plane << 30); 0xFFFFFFFF in w marks an empty slot
    float4* nodes = nullptr;       // [cap]
};

namespace pm {

pm_status fail(pm_ctx* ctx, pm_status code, const char* fmt, ...);
pm_status hip_fail(pm_ctx* ctx, hipError_t e, const char* what);

#define PM_HIP(ctx, call)                                      \
    do {                                                       \
        hipError_t _e = (call);                                \
        if (_e != hipSuccess) return pm::hip_fail(ctx, _e, #call); \
    } while (0)

#define PM_CHECK_LAUNCH(ctx, what)                               \
    do {                                                         \
        hipError_t _e = hipGetLastError();                       \
        if (_e != hipSuccess) return pm::hip_fail(ctx, _e, what); \
    } while (0)

// Scratch arena: returns at least `bytes` of device memory valid until the
// next scratch request on this context. Grows (after a stream sync) as needed.
pm_status scratch_reserve(pm_ctx* ctx, size_t bytes, void** out);

// Bump allocator over one reserved scratch block.
struct Carve {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
};
static inline size_t carve_size(std::initializer_list<size_t> bytes) {
    size_t t = 0;
    for (size_t b : bytes) t = ((t + 255) & ~size_t(255)) + b;
    return t + 256;
}

// Decoupled look-back state for one single-pass launch of `tiles` tiles (the
// context's epoch-tagged 64-bit tile words {epoch:30, status:2, value:32}
// and its counter words: the ordered tile counter, which the last tile
// resets, and the launch epoch). One launch at a time per context: launches
// on the context's stream never overlap. Each launch reads the epoch
// (lookback_epoch, before taking a tile id) and advances it exactly once,
// when no tile of the launch can still be about to read it.
pm_status lookback_begin(pm_ctx* ctx, uint32_t tiles, uint64_t** state, uint32_t** counter);
// The epoch word sits 128 bytes past the tile counter, on a cache line of
// its own: every tile reads it with a plain load (an L2 hit after the first;
// it was written by an earlier launch), while the counter line takes one
// device-scope atomic per tile, and those serialise.
constexpr int kEpochWord = 32;
__device__ __forceinline__ uint32_t lookback_epoch(const uint32_t* counter) {
    return *(const volatile uint32_t*)(counter + kEpochWord);
}
// (after 2^30 - 1 launches the epoch wraps to 1)
__device__ __forceinline__ void lookback_advance(uint32_t* counter, uint32_t epoch) {
    __hip_atomic_store(counter + kEpochWord, epoch + 1 < (1u << 30) ? epoch + 1 : 1u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// PM_KNN_TUNED: the flags with the tuning bits of the context's remembered
// choice for (k, the grid's photon density at this radius); without the
// flag, or without a remembered choice, the flags as given (less the flag).
uint32_t tuned_flags(const pm_ctx* ctx, const pm_grid* grid, float radius, int32_t k, uint32_t flags);
// The remembered cell size for (k, density), else `fallback`.
float tuned_cell(const pm_ctx* ctx, uint32_t n, const float aabb[6], float radius, int32_t k, float fallback);

// Scan over u32 (exclusive or inclusive), optional u64 device total.
pm_status scan_u32(pm_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, uint64_t* total);

// Stable LSD radix sort of (key, value) pairs over key bits [0, key_bits).
// Sorted output ends in keys_out / vals_out. vals_in may be null (identity:
// the index within the segment). segs arrays of n pairs laid end to end are
// sorted each on its own, in the same launches. n_dev (nullable, one
// segment): the element count on the device, at most n; only the first
// *n_dev elements are sorted (and written).
pm_status radix_sort_pairs(pm_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                           uint32_t* vals_out, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n, int key_bits,
                           uint32_t* hist, uint32_t* offs, const uint32_t* n_dev = nullptr, uint32_t segs = 1);
size_t radix_sort_scratch_bytes(uint32_t n, uint32_t segs = 1);
pm_status radix_sort_pairs_auto(pm_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                                uint32_t* vals_out, uint32_t n, int key_bits, void* work,
                                const uint32_t* n_dev = nullptr, uint32_t segs = 1);

pm_status fill_u32(pm_ctx* ctx, uint32_t* p, uint32_t value, size_t count);

// Device blocks for grids and frames: a cached block of this context of at least
// `bytes` (at most twice as large), else a fresh hipMalloc; *got = its size.
// A returned block may still be read by work enqueued earlier on this
// context's stream; work enqueued later is ordered after it.
hipError_t block_take(pm_ctx* ctx, size_t bytes, void** p, size_t* got);
// Give a block back to its context's cache (false when the context is gone
// or the cache is full: the caller frees it).
bool block_give(pm_ctx* owner, void* p, size_t bytes);
// Free every cached block of the context (after its stream drains).
hipError_t cache_release(pm_ctx* ctx);

// pm_grid_build / pm_grid_rebuild over the first n photons of a photons9
// block whose planes are `stride` floats apart (a trace's output map holds
// its count of photons at the plane stride of its capacity).
pm_status grid_build_strided(pm_ctx* ctx, const float* d_photons9, uint32_t n, size_t stride, const float aabb[6],
                             float cell_size, int32_t desired, int32_t layout, pm_grid** out);
pm_status grid_rebuild_strided(pm_ctx* ctx, pm_grid* grid, const float* d_photons9, uint32_t n, size_t stride);

// Planned photon traces (trace.hip): a fixed schedule of `paths` paths in
// batches, enqueued without host synchronisation; the device fill counts
// (u64 global, caustic) tell afterwards whether the maps filled.
struct TracePlan;
pm_status trace_plan_create(pm_ctx* ctx, const pm_scene* scene, uint32_t n_global, uint32_t n_caustic,
                            int32_t max_bounces, uint64_t paths, TracePlan** out);
pm_status trace_plan_enqueue(pm_ctx* ctx, TracePlan* T, uint64_t seed, float* d_global9, float* d_caustic9);
const uint64_t* trace_plan_counts(const TracePlan* T);
void trace_plan_destroy(TracePlan* T);

// Planned frames (render.hip): pm_render_frame's phases with the hit count
// kept on the device (query block sized for `hits`, the planning frame's).
struct FramePlan;
pm_status frame_plan_create(pm_ctx* ctx, const pm_scene* scene, const pm_camera* camera,
                            const pm_render_params* params, uint64_t hits, FramePlan** out);
pm_status frame_plan_enqueue(pm_ctx* ctx, FramePlan* F, const pm_grid* global_grid, const pm_grid* caustic_grid,
                             float* d_rgb3, uint8_t* d_rgb8);
const uint64_t* frame_plan_hits(const FramePlan* F);
uint32_t frame_plan_capacity(const FramePlan* F);
void frame_plan_destroy(FramePlan* F);

// Fused gather + radiance over a query block of capacity `cap` (the planes'
// stride) whose count lives on the device (*d_nq <= cap): nothing is read
// back to the host, so a caller can enqueue it before the count is known.
pm_status gather_radiance_counted(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, const float* d_normals3,
                                  uint32_t cap, const uint32_t* d_nq, float radius, int32_t k, uint32_t flags,
                                  float* d_rgb3, uint32_t* d_count);

// Device address of the allocation-free gather's fixed scratch on the
// context's device (taken once per device; called by pm_ctx_create).
pm_status noalloc_setup(pm_ctx* ctx);

// Per-model AABB over triangle ranges [tb[m], tb[m+1]) (device arrays).
pm_status model_aabb_launch(pm_ctx* ctx, const float* tris9, const uint32_t* tb, int n_models, float* out6);

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace pm

// ---------------------------------------------------------------------------
// Device helpers shared by the kernels.
// ---------------------------------------------------------------------------
namespace pmd {

struct GridGeom {
    float mn[3];
    float cell;
    int g[3];
    int layout;
};

// separated_hash_position (PhotonGrid.cu:10-23): i = (int)((p - min)/cell)
// clamped to [0, g-1]. The float is range-checked before conversion so the
// result is defined for every input (and equals the reference inside the
// AABB). Division is IEEE correctly rounded (hipcc's default for FP32).
__device__ __forceinline__ int cell_coord(float p, float mn, float cell, int g) {
    float f = (p - mn) / cell;
    if (!(f < (float)g)) return g - 1;
    if (f < 1.0f) return 0;
    int i = (int)f;
    return i < g - 1 ? i : g - 1;
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3FF;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

__device__ __forceinline__ uint32_t cell_index(const GridGeom& G, int x, int y, int z) {
    if (G.layout == PM_LAYOUT_MORTON) return spread3((uint32_t)x) | (spread3((uint32_t)y) << 1) | (spread3((uint32_t)z) << 2);
    return (uint32_t)x + (uint32_t)y * (uint32_t)G.g[0] + (uint32_t)z * (uint32_t)G.g[0] * (uint32_t)G.g[1];
}

// Map hardware workgroup id to a logical tile so that each of the 8 XCDs
// (which receive workgroups round-robin) works on one contiguous range of
// tiles and keeps its neighbouring cells in its own L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nb) {
    const uint32_t X = 8;
    uint32_t per = nb / X, rem = nb % X;
    uint32_t x = bid % X, i = bid / X;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

}  // namespace pmd
