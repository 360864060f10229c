// Photon-map files: the binary format read by load() (Initializer.cu:102-141)
// — a uint64 photon count, then per photon 9 FP32 values (pos, power,
// inc_dir), 36 bytes — between files and device photon blocks (9 planes).
// The reference only reads this format; the writer persists resident maps.
//
// Files stream through a pinned host buffer in chunks of kChunk records;
// each chunk crosses PCIe as records and is transposed on the device, so the
// host does no per-photon work.
#include <cstdio>
#include <memory>

#include "pm_internal.hpp"

namespace {

constexpr uint32_t kChunk = 1u << 20;  // records per chunk (36 MiB)

// records [count][9] -> planes [9][n] at photon offset base
__global__ __launch_bounds__(256) void k_records_to_planes(const float* __restrict__ rec, uint32_t count,
                                                          uint32_t base, uint32_t n, float* __restrict__ planes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
#pragma unroll
    for (int p = 0; p < 9; ++p) planes[(size_t)p * n + base + i] = rec[(size_t)i * 9 + p];
}

// planes [9][n] at photon offset base -> records [count][9]
__global__ __launch_bounds__(256) void k_planes_to_records(const float* __restrict__ planes, uint32_t count,
                                                          uint32_t base, uint32_t n, float* __restrict__ rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
#pragma unroll
    for (int p = 0; p < 9; ++p) rec[(size_t)i * 9 + p] = planes[(size_t)p * n + base + i];
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

struct Pinned {
    void* p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace

extern "C" pm_status pm_photon_file_count(pm_ctx* ctx, const char* path, uint64_t* n) {
    if (!ctx || !path || !n) return PM_ERR_INVALID;
    File F;
    F.f = std::fopen(path, "rb");
    if (!F.f) return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_count: cannot open %s", path);
    uint64_t c = 0;
    if (std::fread(&c, 8, 1, F.f) != 1) return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_count: no count in %s", path);
    *n = c;
    return PM_OK;
}

extern "C" pm_status pm_photon_file_read(pm_ctx* ctx, const char* path, float* d_photons9, uint32_t capacity,
                                         uint32_t* n_read) {
    if (!ctx || !path || !n_read) return PM_ERR_INVALID;
    *n_read = 0;
    File F;
    F.f = std::fopen(path, "rb");
    if (!F.f) return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_read: cannot open %s", path);
    uint64_t count = 0;
    if (std::fread(&count, 8, 1, F.f) != 1)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_read: failed to read the photon count of %s", path);
    if (count > capacity)
        return pm::fail(ctx, PM_ERR_CAPACITY, "pm_photon_file_read: %llu photons in %s, capacity %u",
                        (unsigned long long)count, path, capacity);
    const uint32_t n = (uint32_t)count;
    if (n == 0) return PM_OK;
    if (!d_photons9) return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_read: null photon block");
    const uint32_t chunk = std::min(n, kChunk);
    Pinned host;
    PM_HIP(ctx, hipHostMalloc(&host.p, (size_t)chunk * 36, hipHostMallocDefault));
    void* work = nullptr;
    pm_status st;
    if ((st = pm::scratch_reserve(ctx, (size_t)chunk * 36 + 256, &work)) != PM_OK) return st;
    float* dev_rec = static_cast<float*>(work);
    for (uint32_t base = 0; base < n; base += chunk) {
        const uint32_t c = std::min(chunk, n - base);
        // the previous chunk's copy must be done before the buffer is refilled
        PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (std::fread(host.p, 36, c, F.f) != c)
            return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_read: failed to read photon data of %s", path);
        PM_HIP(ctx, hipMemcpyAsync(dev_rec, host.p, (size_t)c * 36, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_records_to_planes, dim3(pm::div_up(c, 256)), dim3(256), 0, ctx->stream, dev_rec, c, base, n,
                           d_photons9);
        PM_CHECK_LAUNCH(ctx, "k_records_to_planes");
    }
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *n_read = n;
    return PM_OK;
}

extern "C" pm_status pm_photon_file_write(pm_ctx* ctx, const char* path, const float* d_photons9, uint32_t n) {
    if (!ctx || !path || (n > 0 && !d_photons9)) return PM_ERR_INVALID;
    File F;
    F.f = std::fopen(path, "wb");
    if (!F.f) return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_write: cannot open %s", path);
    const uint64_t count = n;
    if (std::fwrite(&count, 8, 1, F.f) != 1) return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_write: write failed");
    if (n > 0) {
        const uint32_t chunk = std::min(n, kChunk);
        Pinned host;
        PM_HIP(ctx, hipHostMalloc(&host.p, (size_t)chunk * 36, hipHostMallocDefault));
        void* work = nullptr;
        pm_status st;
        if ((st = pm::scratch_reserve(ctx, (size_t)chunk * 36 + 256, &work)) != PM_OK) return st;
        float* dev_rec = static_cast<float*>(work);
        for (uint32_t base = 0; base < n; base += chunk) {
            const uint32_t c = std::min(chunk, n - base);
            hipLaunchKernelGGL(k_planes_to_records, dim3(pm::div_up(c, 256)), dim3(256), 0, ctx->stream, d_photons9, c,
                               base, n, dev_rec);
            PM_CHECK_LAUNCH(ctx, "k_planes_to_records");
            PM_HIP(ctx, hipMemcpyAsync(host.p, dev_rec, (size_t)c * 36, hipMemcpyDeviceToHost, ctx->stream));
            PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if (std::fwrite(host.p, 36, c, F.f) != c)
                return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_write: write failed");
        }
    }
    if (std::fflush(F.f) != 0) return pm::fail(ctx, PM_ERR_INVALID, "pm_photon_file_write: write failed");
    return PM_OK;
}
