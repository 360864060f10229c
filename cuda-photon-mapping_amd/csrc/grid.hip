// Photon hash grid build: cell keys -> stable radix sort of (key, index) ->
// cell sizes from the sorted runs' ends -> exclusive scan into starts -> one
// gather of each photon's record into the cell-ordered SoA planes.
// Replaces PhotonGrid::build (PhotonGrid.cuh:51-66; PhotonGrid.cu:41-218).
#include <cmath>

#include "pm_internal.hpp"

namespace {

using pmd::GridGeom;

// Cell index of every photon (the hash of PhotonGrid.cu:152-174 without the
// counting atomics: counts come from the sorted keys). The same pass copies
// each photon into a 48-byte record (three float4), so the scatter below
// gathers one record per photon instead of nine scattered 4-byte values.
// in9: nine planes of `stride` floats (stride >= n: a photon map that holds
// fewer photons than its capacity)
__global__ __launch_bounds__(256) void k_grid_keys(const float* __restrict__ in9, uint32_t n, size_t stride,
                                                  GridGeom G, uint32_t* __restrict__ keys, float4* __restrict__ rec) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) v[p] = in9[(size_t)p * stride + i];
    int x = pmd::cell_coord(v[0], G.mn[0], G.cell, G.g[0]);
    int y = pmd::cell_coord(v[1], G.mn[1], G.cell, G.g[1]);
    int z = pmd::cell_coord(v[2], G.mn[2], G.cell, G.g[2]);
    keys[i] = pmd::cell_index(G, x, y, z);
    rec[3 * (size_t)i + 0] = make_float4(v[0], v[1], v[2], v[3]);
    rec[3 * (size_t)i + 1] = make_float4(v[4], v[5], v[6], v[7]);
    rec[3 * (size_t)i + 2] = make_float4(v[8], 0.f, 0.f, 0.f);
}

// Cell sizes from the sorted keys (the counts of PhotonGrid.cu:152-174): a
// run of equal keys [s, e] adds e + 1 at its last position and subtracts s
// at its first (mod 2^32), so each occupied cell ends at e + 1 - s and every
// other cell keeps the 0 it was filled with. The exclusive scan of the
// sizes gives the starts (PhotonGrid.cu:195-196).
__global__ __launch_bounds__(256) void k_cell_sizes(const uint32_t* __restrict__ skeys, uint32_t n,
                                                   uint32_t* __restrict__ sizes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = skeys[i];
    const bool first = i == 0 || skeys[i - 1] != key;
    const bool last = i + 1 == n || skeys[i + 1] != key;
    if (first && last)
        sizes[key] = 1u;
    else if (first)
        atomicSub(&sizes[key], i);
    else if (last)
        atomicAdd(&sizes[key], i + 1);
}

// analyze_density_kernel's max (PhotonGrid.cu:70-90), as a reduction.
__global__ __launch_bounds__(256) void k_max_u32(const uint32_t* __restrict__ a, uint32_t n, uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, a[i]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// assign_photons_to_cells (PhotonGrid.cu:152-174) with the slot taken from
// the stable sort instead of an atomic: in-cell order = input order. One
// 48-byte record gathered per photon, nine coalesced plane writes.
__global__ __launch_bounds__(256) void k_grid_scatter(const float4* __restrict__ rec, uint32_t n,
                                                     const uint32_t* __restrict__ order, float* __restrict__ out9,
                                                     uint32_t* __restrict__ perm) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t src = order[j];
    perm[j] = src;
    const float4 a = rec[3 * (size_t)src + 0], b = rec[3 * (size_t)src + 1], c = rec[3 * (size_t)src + 2];
    const float v[9] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x};
#pragma unroll
    for (int p = 0; p < 9; ++p) out9[(size_t)p * n + j] = v[p];
}

struct HostDims {
    float cell;
    int g[3];
    uint32_t n_cells;
};

// compute_grid_and_cell_size (PhotonGrid.cu:41-56), host FP32.
float reference_cell(uint32_t n, const float aabb[6], int desired) {
    float sx = std::fabs(aabb[3] - aabb[0]);
    float sy = std::fabs(aabb[4] - aabb[1]);
    float sz = std::fabs(aabb[5] - aabb[2]);
    int N = ((int)n + desired - 1) / desired;
    float volume = sx * sy * sz;
    return std::cbrt(volume / (float)N);
}

bool make_dims(const float aabb[6], float cell, int layout, HostDims& d) {
    d.cell = cell;
    uint64_t total = 1;
    int mx = 1;
    for (int a = 0; a < 3; ++a) {
        float size = std::fabs(aabb[a + 3] - aabb[a]);
        float gf = std::ceil(size / cell);
        if (!(gf < 1048576.f)) return false;
        int g = (int)gf;
        d.g[a] = g < 1 ? 1 : g;
        total *= (uint64_t)d.g[a];
        mx = std::max(mx, d.g[a]);
    }
    if (layout == PM_LAYOUT_MORTON) {
        if (mx > 1024) return false;
        uint64_t side = 1;
        while ((int)side < mx) side <<= 1;
        total = side * side * side;
    }
    if (total > (1ull << 30)) return false;
    d.n_cells = (uint32_t)total;
    return true;
}

int bits_for(uint32_t n_cells) {
    int b = 1;
    while (b < 32 && (1ull << b) < (uint64_t)n_cells) ++b;
    return b;
}

GridGeom geom_of(const float aabb[6], const HostDims& d, int layout) {
    GridGeom G;
    for (int a = 0; a < 3; ++a) {
        G.mn[a] = aabb[a];
        G.g[a] = d.g[a];
    }
    G.cell = d.cell;
    G.layout = layout;
    return G;
}

// A grid over no photons (a trace that stored nothing): every cell is
// empty, so every gather finds nothing and every estimate is zero. The
// reference's sizing rule divides by ceil(0 / desired) = 0 there
// (PhotonGrid.cu:41-56); without a given cell size the grid is one cell.
pm_status build_empty(pm_ctx* ctx, const float aabb[6], float cell_size, int layout, pm_grid** out) {
    float cell = cell_size;
    if (!(cell > 0.f)) {
        cell = 0.f;
        for (int a = 0; a < 3; ++a) cell = std::fmax(cell, std::fabs(aabb[a + 3] - aabb[a]));
        if (!(cell > 0.f) || !std::isfinite(cell)) cell = 1.f;
    }
    HostDims d;
    if (!make_dims(aabb, cell, layout, d))
        return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_grid_build: grid too large for cell %g", (double)cell);
    pm_grid* g = new pm_grid();
    // the photon planes and the permutation are empty; they point at a
    // zeroed 256-byte block so no pointer is null
    hipError_t e = pm::block_take(ctx, pm::carve_size({(size_t)d.n_cells * 4, (size_t)d.n_cells * 4, 256}), &g->mem,
                                  &g->mem_bytes);
    if (e != hipSuccess) {
        delete g;
        return pm::hip_fail(ctx, e, "hipMalloc(grid)");
    }
    g->owner = ctx;
    pm::Carve gc{static_cast<char*>(g->mem)};
    uint32_t* starts = gc.take<uint32_t>(d.n_cells);
    uint32_t* sizes = gc.take<uint32_t>(d.n_cells);
    uint32_t* nothing = gc.take<uint32_t>(64);
    pm_status st;
    if ((st = pm::fill_u32(ctx, starts, 0u, d.n_cells)) != PM_OK ||
        (st = pm::fill_u32(ctx, sizes, 0u, d.n_cells)) != PM_OK || (st = pm::fill_u32(ctx, nothing, 0u, 64)) != PM_OK) {
        pm_grid_destroy(g);
        return st;
    }
    pm_grid_desc& D = g->desc;
    for (int a = 0; a < 6; ++a) D.aabb[a] = aabb[a];
    D.cell_size = d.cell;
    for (int a = 0; a < 3; ++a) D.dims[a] = d.g[a];
    D.layout = layout;
    D.n_cells = d.n_cells;
    D.n_photons = 0;
    D.d_cell_starts = starts;
    D.d_cell_sizes = sizes;
    D.d_photons = reinterpret_cast<const float*>(nothing);
    D.d_perm = nothing;
    *out = g;
    return PM_OK;
}

size_t grid_bytes(uint32_t n_cells, uint32_t n_cap) {
    return pm::carve_size({(size_t)n_cells * 4, (size_t)n_cells * 4, (size_t)n_cap * 36, (size_t)n_cap * 4});
}

// scratch of one build: the density counter, keys, sorted keys, sorted
// order, radix work, records
size_t build_scratch_bytes(uint32_t n) {
    return pm::carve_size({256, (size_t)n * 4, (size_t)n * 4, (size_t)n * 4, pm::radix_sort_scratch_bytes(n),
                           (size_t)n * 48});
}

// Enqueue the build of n photons at dims dd into g's memory (capacity
// g->n_cap >= n): keys + records, stable key sort, cell sizes from the sorted
// runs, starts by exclusive scan, record gather. No host synchronisation.
pm_status enqueue_build(pm_ctx* ctx, const float* d_photons9, uint32_t n, size_t stride, const float aabb[6],
                        const HostDims& dd,
                        int layout, pm_grid* g) {
    pm::Carve gc{static_cast<char*>(g->mem)};
    uint32_t* starts = gc.take<uint32_t>(dd.n_cells);
    uint32_t* sizes = gc.take<uint32_t>(dd.n_cells);
    float* photons = gc.take<float>((size_t)g->n_cap * 9);
    uint32_t* perm = gc.take<uint32_t>(g->n_cap);
    pm_grid_desc& D = g->desc;
    for (int a = 0; a < 6; ++a) D.aabb[a] = aabb[a];
    D.cell_size = dd.cell;
    for (int a = 0; a < 3; ++a) D.dims[a] = dd.g[a];
    D.layout = layout;
    D.n_cells = dd.n_cells;
    D.n_photons = n;
    D.d_cell_starts = starts;
    D.d_cell_sizes = sizes;
    D.d_photons = photons;
    D.d_perm = perm;
    pm_status st;
    if (n == 0) {
        if ((st = pm::fill_u32(ctx, starts, 0u, dd.n_cells)) != PM_OK) return st;
        return pm::fill_u32(ctx, sizes, 0u, dd.n_cells);
    }
    void* work = nullptr;
    if ((st = pm::scratch_reserve(ctx, build_scratch_bytes(n), &work)) != PM_OK) return st;
    pm::Carve cv{static_cast<char*>(work)};
    cv.take<uint32_t>(64);
    uint32_t* keys = cv.take<uint32_t>(n);
    uint32_t* skeys = cv.take<uint32_t>(n);
    uint32_t* order = cv.take<uint32_t>(n);
    void* rwork = cv.take<char>(pm::radix_sort_scratch_bytes(n));
    float4* rec = cv.take<float4>((size_t)n * 3);
    const uint32_t blocks = pm::div_up(n, 256);
    const GridGeom G = geom_of(aabb, dd, layout);
    hipLaunchKernelGGL(k_grid_keys, dim3(blocks), dim3(256), 0, ctx->stream, d_photons9, n, stride, G, keys, rec);
    PM_CHECK_LAUNCH(ctx, "k_grid_keys");
    if ((st = pm::radix_sort_pairs_auto(ctx, keys, nullptr, skeys, order, n, bits_for(dd.n_cells), rwork)) != PM_OK)
        return st;
    if ((st = pm::fill_u32(ctx, sizes, 0u, dd.n_cells)) != PM_OK) return st;
    hipLaunchKernelGGL(k_cell_sizes, dim3(blocks), dim3(256), 0, ctx->stream, skeys, n, sizes);
    PM_CHECK_LAUNCH(ctx, "k_cell_sizes");
    if ((st = pm::scan_u32(ctx, sizes, starts, dd.n_cells, false, nullptr)) != PM_OK) return st;
    hipLaunchKernelGGL(k_grid_scatter, dim3(blocks), dim3(256), 0, ctx->stream, rec, n, order, photons, perm);
    PM_CHECK_LAUNCH(ctx, "k_grid_scatter");
    return PM_OK;
}

}  // namespace

namespace pm {

pm_status grid_rebuild_strided(pm_ctx* ctx, pm_grid* grid, const float* d_photons9, uint32_t n, size_t stride) {
    if (!ctx || !grid) return PM_ERR_INVALID;
    if (n > 0 && !d_photons9) return pm::fail(ctx, PM_ERR_INVALID, "pm_grid_rebuild: null photons");
    if (n > grid->n_cap)
        return pm::fail(ctx, PM_ERR_CAPACITY, "pm_grid_rebuild: %u photons exceed the grid's capacity of %u", n,
                        grid->n_cap);
    HostDims dd;
    dd.cell = grid->desc.cell_size;
    for (int a = 0; a < 3; ++a) dd.g[a] = grid->desc.dims[a];
    dd.n_cells = grid->desc.n_cells;
    float aabb[6];
    for (int a = 0; a < 6; ++a) aabb[a] = grid->desc.aabb[a];
    return enqueue_build(ctx, d_photons9, n, stride, aabb, dd, grid->desc.layout, grid);
}

pm_status grid_build_strided(pm_ctx* ctx, const float* d_photons9, uint32_t n, size_t stride, const float aabb[6],
                             float cell_size, int32_t desired, int32_t layout, pm_grid** out) {
    if (!ctx || !out) return PM_ERR_INVALID;
    *out = nullptr;
    if (!aabb || (n > 0 && !d_photons9)) return pm::fail(ctx, PM_ERR_INVALID, "pm_grid_build: null argument");
    if (layout != PM_LAYOUT_LINEAR && layout != PM_LAYOUT_MORTON)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_grid_build: bad layout %d", layout);
    if (n > 0x7FFFFFFFu) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_grid_build: n too large");
    if (!(cell_size > 0.f) && desired <= 0)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_grid_build: need cell_size > 0 or desired_per_cell > 0");

    if (n == 0) return build_empty(ctx, aabb, cell_size, layout, out);

    HostDims d;
    float cell = cell_size > 0.f ? cell_size : reference_cell(n, aabb, desired);
    if (!(cell > 0.f) || !make_dims(aabb, cell, layout, d))
        return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_grid_build: grid too large for cell %g", (double)cell);

    pm_status st;
    uint32_t* dmax = nullptr;
    {
        void* w = nullptr;
        if ((st = pm::scratch_reserve(ctx, build_scratch_bytes(n), &w)) != PM_OK) return st;
        pm::Carve cv{static_cast<char*>(w)};
        dmax = cv.take<uint32_t>(64);
    }

    // One build at dims d into a fresh allocation of capacity n.
    auto build_once = [&](const HostDims& dd, pm_grid*& out_g) -> pm_status {
        pm_grid* g = new pm_grid();
        hipError_t e = pm::block_take(ctx, grid_bytes(dd.n_cells, n), &g->mem, &g->mem_bytes);
        if (e != hipSuccess) {
            delete g;
            return pm::hip_fail(ctx, e, "hipMalloc(grid)");
        }
        g->owner = ctx;
        g->n_cap = n;
        pm_status bs = enqueue_build(ctx, d_photons9, n, stride, aabb, dd, layout, g);
        if (bs != PM_OK) {
            pm_grid_destroy(g);
            return bs;
        }
        out_g = g;
        return PM_OK;
    };

    pm_grid* g = nullptr;
    if ((st = build_once(d, g)) != PM_OK) return st;
    if (!(cell_size > 0.f)) {
        // analyze_density (PhotonGrid.cu:92-141): at the reference size, the
        // fullest cell (read from the built grid's sizes) decides whether the
        // cells shrink to the size for desired / 5; then the grid is built
        // again at that size.
        if ((st = pm::fill_u32(ctx, dmax, 0u, 1)) != PM_OK) return st;
        hipLaunchKernelGGL(k_max_u32, dim3(std::min<uint32_t>(pm::div_up(d.n_cells, 256), 4096)), dim3(256), 0,
                           ctx->stream, g->desc.d_cell_sizes, d.n_cells, dmax);
        PM_CHECK_LAUNCH(ctx, "k_max_u32");
        uint32_t mx = 0;
        PM_HIP(ctx, hipMemcpyAsync(&mx, dmax, 4, hipMemcpyDeviceToHost, ctx->stream));
        PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (mx > 5u * (uint32_t)desired) {
            pm_grid_destroy(g);
            g = nullptr;
            cell = reference_cell(n, aabb, std::max(1, desired / 5));
            if (!(cell > 0.f) || !make_dims(aabb, cell, layout, d))
                return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_grid_build: grid too large for cell %g", (double)cell);
            if ((st = build_once(d, g)) != PM_OK) return st;
        }
    }
    *out = g;
    return PM_OK;
}

}  // namespace pm

extern "C" pm_status pm_grid_rebuild(pm_ctx* ctx, pm_grid* grid, const float* d_photons9, uint32_t n) {
    return pm::grid_rebuild_strided(ctx, grid, d_photons9, n, n);
}

extern "C" pm_status pm_grid_build(pm_ctx* ctx, const float* d_photons9, uint32_t n, const float aabb[6],
                                   float cell_size, int32_t desired, int32_t layout, pm_grid** out) {
    return pm::grid_build_strided(ctx, d_photons9, n, n, aabb, cell_size, desired, layout, out);
}

extern "C" pm_status pm_grid_describe(const pm_grid* grid, pm_grid_desc* out) {
    if (!grid || !out) return PM_ERR_INVALID;
    *out = grid->desc;
    return PM_OK;
}

extern "C" pm_status pm_grid_copy_to_host(pm_ctx* ctx, const pm_grid* grid, uint32_t* h_starts, uint32_t* h_sizes,
                                          float* h_photons9, uint32_t* h_perm) {
    if (!ctx || !grid) return PM_ERR_INVALID;
    const pm_grid_desc& D = grid->desc;
    if (h_starts)
        PM_HIP(ctx, hipMemcpyAsync(h_starts, D.d_cell_starts, (size_t)D.n_cells * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (h_sizes)
        PM_HIP(ctx, hipMemcpyAsync(h_sizes, D.d_cell_sizes, (size_t)D.n_cells * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (h_photons9)
        PM_HIP(ctx, hipMemcpyAsync(h_photons9, D.d_photons, (size_t)D.n_photons * 36, hipMemcpyDeviceToHost, ctx->stream));
    if (h_perm)
        PM_HIP(ctx, hipMemcpyAsync(h_perm, D.d_perm, (size_t)D.n_photons * 4, hipMemcpyDeviceToHost, ctx->stream));
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PM_OK;
}

extern "C" pm_status pm_grid_destroy(pm_grid* grid) {
    if (!grid) return PM_OK;
    // the block goes back to the building context for its later builds (in
    // its stream's order); past the cache's size, or with the context gone,
    // it is freed
    if (grid->mem && !pm::block_give(grid->owner, grid->mem, grid->mem_bytes)) (void)hipFree(grid->mem);
    delete grid;
    return PM_OK;
}
