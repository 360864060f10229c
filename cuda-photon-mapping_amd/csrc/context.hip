// Context, error reporting, scratch arena and small utility kernels.
#include <cstdarg>
#include <cstring>

#include <mutex>
#include <unordered_set>

#include "pm_internal.hpp"

namespace pm {

pm_status fail(pm_ctx* ctx, pm_status code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

pm_status hip_fail(pm_ctx* ctx, hipError_t e, const char* what) {
    pm_status code = (e == hipErrorOutOfMemory) ? PM_ERR_NOMEM : PM_ERR_HIP;
    return fail(ctx, code, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

pm_status scratch_reserve(pm_ctx* ctx, size_t bytes, void** out) {
    if (ctx->fixed_arena) {
        if (bytes > ctx->fixed_arena_bytes)
            return fail(ctx, PM_ERR_UNSUPPORTED, "scratch of %zu bytes exceeds the fixed block of %zu bytes", bytes,
                        ctx->fixed_arena_bytes);
        *out = ctx->fixed_arena;
        return PM_OK;
    }
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) {
            PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            PM_HIP(ctx, hipFree(ctx->scratch));
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        size_t want = bytes + bytes / 4 + (1u << 20);
        hipError_t e = hipMalloc(&ctx->scratch, want);
        if (e != hipSuccess) {
            ctx->scratch = nullptr;
            return hip_fail(ctx, e, "hipMalloc(scratch)");
        }
        ctx->scratch_bytes = want;
    }
    *out = ctx->scratch;
    return PM_OK;
}


__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

namespace {
// live contexts (a grid may outlive the context it was built on) and their
// block caches, under one lock
std::mutex g_ctx_mu;
std::unordered_set<pm_ctx*> g_ctx_live;
}  // namespace

hipError_t block_take(pm_ctx* ctx, size_t bytes, void** p, size_t* got) {
    {
        std::lock_guard<std::mutex> lock(g_ctx_mu);
        auto& c = ctx->block_cache;
        size_t best = c.size();
        for (size_t i = 0; i < c.size(); ++i)
            if (c[i].second >= bytes && c[i].second <= 2 * bytes && (best == c.size() || c[i].second < c[best].second))
                best = i;
        if (best != c.size()) {
            *p = c[best].first;
            *got = c[best].second;
            ctx->block_cache_bytes -= c[best].second;
            c.erase(c.begin() + (std::ptrdiff_t)best);
            return hipSuccess;
        }
    }
    *got = bytes;
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory && ctx->block_cache_bytes > 0) {
        // the cache holds what the device lacks: free it and try again
        (void)hipGetLastError();
        if (cache_release(ctx) == hipSuccess) e = hipMalloc(p, bytes);
    }
    return e;
}

hipError_t cache_release(pm_ctx* ctx) {
    std::vector<std::pair<void*, size_t>> drop;
    {
        std::lock_guard<std::mutex> lock(g_ctx_mu);
        drop.swap(ctx->block_cache);
        ctx->block_cache_bytes = 0;
    }
    if (drop.empty()) return hipSuccess;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    for (auto& b : drop) (void)hipFree(b.first);
    return e;
}

bool block_give(pm_ctx* owner, void* p, size_t bytes) {
    if (!owner || !p) return false;
    std::lock_guard<std::mutex> lock(g_ctx_mu);
    if (!g_ctx_live.count(owner) || owner->block_cache_bytes + bytes > owner->block_cache_cap) return false;
    owner->block_cache.emplace_back(p, bytes);
    owner->block_cache_bytes += bytes;
    return true;
}

pm_status fill_u32(pm_ctx* ctx, uint32_t* p, uint32_t value, size_t count) {
    if (count == 0) return PM_OK;
    uint32_t blocks = (uint32_t)std::min<size_t>((count + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, ctx->stream, p, value, count);
    PM_CHECK_LAUNCH(ctx, "k_fill_u32");
    return PM_OK;
}

}  // namespace pm

extern "C" {

const char* pm_version(void) { return "pm355 0.1 (gfx950)"; }

pm_status pm_ctx_create(int32_t device, void* hip_stream, pm_ctx** out) {
    if (!out) return PM_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return PM_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return PM_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PM_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PM_ERR_NO_DEVICE;
    pm_ctx* c = new pm_ctx();
    c->device = device;
    c->stream = (hipStream_t)hip_stream;
    c->cu_count = prop.multiProcessorCount;
    c->block_cache_cap = prop.totalGlobalMem / 4;
    // Tile counter 0, first look-back epoch 1, written by fill kernels on the
    // context's own stream: the null stream (hipMemset / hipMemcpy) is not
    // ordered with a non-blocking stream, so a first scan could otherwise
    // start before the words land.
    if (hipMalloc(&c->scan_counter, 256) != hipSuccess) {
        c->scan_counter = nullptr;
        delete c;
        return PM_ERR_NOMEM;
    }
    if (pm::fill_u32(c, c->scan_counter, 0u, 64) != PM_OK ||
        pm::fill_u32(c, c->scan_counter + pm::kEpochWord, 1u, 1) != PM_OK) {
        (void)hipFree(c->scan_counter);
        delete c;
        return PM_ERR_HIP;
    }
    // (a failure leaves only pm_grid_gather_radiance_noalloc unavailable,
    // which reports it)
    (void)pm::noalloc_setup(c);
    c->err.clear();
    {
        std::lock_guard<std::mutex> lock(pm::g_ctx_mu);
        pm::g_ctx_live.insert(c);
    }
    *out = c;
    return PM_OK;
}

pm_status pm_ctx_destroy(pm_ctx* ctx) {
    if (!ctx) return PM_OK;
    (void)hipStreamSynchronize(ctx->stream);
    {
        std::lock_guard<std::mutex> lock(pm::g_ctx_mu);
        pm::g_ctx_live.erase(ctx);
        for (auto& b : ctx->block_cache) (void)hipFree(b.first);
        ctx->block_cache.clear();
        ctx->block_cache_bytes = 0;
    }
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    pm_frame_destroy(ctx->frame_state);
    if (ctx->scan_state) (void)hipFree(ctx->scan_state);
    if (ctx->tile_sums) (void)hipFree(ctx->tile_sums);
    if (ctx->scan_counter) (void)hipFree(ctx->scan_counter);
    delete ctx;
    return PM_OK;
}

pm_status pm_ctx_set_stream(pm_ctx* ctx, void* hip_stream) {
    if (!ctx) return PM_ERR_INVALID;
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream = (hipStream_t)hip_stream;
    return PM_OK;
}

pm_status pm_ctx_trim(pm_ctx* ctx) {
    if (!ctx) return PM_ERR_INVALID;
    PM_HIP(ctx, pm::cache_release(ctx));
    return PM_OK;
}

pm_status pm_ctx_sync(pm_ctx* ctx) {
    if (!ctx) return PM_ERR_INVALID;
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PM_OK;
}

const char* pm_last_error(const pm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

}  // extern "C"
