// Allocation-free fused gather: pm_grid_gather_radiance for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops
// that must not call hipMalloc, hosts that forbid it). The call's scratch —
// query cell keys, the processing order, the radix-sort passes, the large-k
// overflow list and the look-back scan state — lives in one file-scope
// device array instead of the context's growable arena. Each GPU holds its
// own copy of the array until the process ends; its device address is taken
// once per device when the first context on that device is created.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents: calls that overlap on one GPU (two contexts on two streams, or
// two host threads) race. The kernels are those of pm_grid_gather_radiance.
#include <mutex>

#include "pm_internal.hpp"

namespace {

constexpr uint32_t kMaxQueries = 1u << 20;  // 1,048,576 queries (a 1280x720 frame's hits fit)
constexpr size_t kCounterBytes = 256;       // look-back tile counter and epoch word
constexpr size_t kStateTiles = 256;         // look-back tile words (the sort's scans need 32 at the limit)
constexpr size_t kArenaBytes = 26u << 20;   // keys, order, radix passes, overflow list at the limit
constexpr size_t kScratchBytes = kCounterBytes + kStateTiles * sizeof(uint64_t) + kArenaBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_base[kMaxDevices];  // device address of `scratch` on each device, null until set up

// Look-back words for a first launch: tile counter 0, epoch 1; tile words of
// epoch 0 never match a launch.
__global__ void k_noalloc_init(uint32_t* __restrict__ counter, uint64_t* __restrict__ state, uint32_t tiles) {
    for (uint32_t i = threadIdx.x; i < kCounterBytes / 4; i += blockDim.x)
        counter[i] = i == (uint32_t)pm::kEpochWord ? 1u : 0u;
    for (uint32_t i = threadIdx.x; i < tiles; i += blockDim.x) state[i] = 0;
}

}  // namespace

namespace pm {

// Called by pm_ctx_create with `device` current: the array's address on this
// device, and its look-back words initialised, once per device.
pm_status noalloc_setup(pm_ctx* ctx) {
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices) return fail(ctx, PM_ERR_UNSUPPORTED, "device %d >= %d", dev, kMaxDevices);
    // the largest call's scratch (sorted queries, then the large-k overflow
    // list behind the order, as the gather carves it) must fit the arena
    const size_t a = (size_t)kMaxQueries * 4, rs = radix_sort_scratch_bytes(kMaxQueries);
    if (carve_size({a, a, a, rs, a, 64}) > kArenaBytes)
        return fail(ctx, PM_ERR_UNSUPPORTED, "fixed scratch arena too small for %u queries", kMaxQueries);
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_base[dev]) return PM_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    unsigned char* base = static_cast<unsigned char*>(p);
    hipLaunchKernelGGL(k_noalloc_init, dim3(1), dim3(256), 0, nullptr, reinterpret_cast<uint32_t*>(base),
                       reinterpret_cast<uint64_t*>(base + kCounterBytes), (uint32_t)kStateTiles);
    PM_CHECK_LAUNCH(ctx, "k_noalloc_init");
    PM_HIP(ctx, hipStreamSynchronize(nullptr));
    g_base[dev] = base;
    return PM_OK;
}

}  // namespace pm

extern "C" pm_status pm_grid_gather_radiance_noalloc(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3,
                                                     const float* d_normals3, uint32_t nq, float radius, int32_t k,
                                                     uint32_t flags, float* d_rgb3, uint32_t* d_count) {
    if (!ctx) return PM_ERR_INVALID;
    if (nq > kMaxQueries)
        return pm::fail(ctx, PM_ERR_UNSUPPORTED,
                        "pm_grid_gather_radiance_noalloc: %u queries exceed the fixed scratch limit of %u", nq,
                        kMaxQueries);
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_base[ctx->device] : nullptr;
    }
    if (!base) return pm::fail(ctx, PM_ERR_HIP, "pm_grid_gather_radiance_noalloc: fixed scratch not set up");
    ctx->fixed_counter = reinterpret_cast<uint32_t*>(base);
    ctx->fixed_state = reinterpret_cast<uint64_t*>(base + kCounterBytes);
    ctx->fixed_state_tiles = kStateTiles;
    ctx->fixed_arena = base + kCounterBytes + kStateTiles * sizeof(uint64_t);
    ctx->fixed_arena_bytes = kArenaBytes;
    const pm_status st =
        pm_grid_gather_radiance(ctx, grid, d_queries3, d_normals3, nq, radius, k, flags, d_rgb3, d_count);
    ctx->fixed_counter = nullptr;
    ctx->fixed_state = nullptr;
    ctx->fixed_state_tiles = 0;
    ctx->fixed_arena = nullptr;
    ctx->fixed_arena_bytes = 0;
    return st;
}
