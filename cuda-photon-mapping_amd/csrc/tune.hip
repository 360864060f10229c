// Gather autotuner: the grid cell size and the fused gather's variant
// chosen by timing them on the caller's data (the reference sizes its
// launches by fixed rules, CudaUtils.cu:42-68, Defines.cuh:5-72). Choices
// are remembered per context under (k, photon density class) and applied
// to gathers flagged PM_KNN_TUNED and to the grids of tuned server jobs.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>

#include "pm_internal.hpp"

namespace {

// the bits a tuning chooses; the others (sorting, batch queue) stay the caller's
constexpr uint32_t kTuneBits =
    PM_KNN_NO_XCD_REMAP | PM_KNN_TWO_PASS | PM_KNN_LARGE_BUFFER | PM_KNN_XCD_CONTIGUOUS | PM_KNN_INNER_BOX;

// round(2 log2(n r^3 / V)): photons expected in an r-cube, in half octaves
int density_class(uint32_t n, const float aabb[6], float r) {
    double v = 1.0;
    for (int a = 0; a < 3; ++a) v *= std::max((double)std::fabs(aabb[a + 3] - aabb[a]), 1e-30);
    const double d = (double)n * (double)r * (double)r * (double)r / v;
    if (!(d > 0.0) || !std::isfinite(d)) return INT_MIN / 2;
    return (int)std::lround(2.0 * std::log2(d));
}

bool find(const pm_ctx* ctx, int k, int cls, pm_gather_tuning& out) {
    const pm::TuneEntry* best = nullptr;
    int bd = INT_MAX;
    for (const auto& e : ctx->tunes) {
        if (e.k != k) continue;
        const int d = std::abs(e.cls - cls);
        if (d < bd) {
            bd = d;
            best = &e;
        }
    }
    if (!best) return false;
    out = best->t;
    out.found = 1;
    return true;
}

// Mean duration of `reps` calls of fn after one untimed call, on the stream.
template <class F>
pm_status time_ms(pm_ctx* ctx, hipEvent_t e0, hipEvent_t e1, int reps, float& ms, F&& fn) {
    pm_status st = fn();
    if (st != PM_OK) return st;
    PM_HIP(ctx, hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < reps; ++i)
        if ((st = fn()) != PM_OK) return st;
    PM_HIP(ctx, hipEventRecord(e1, ctx->stream));
    PM_HIP(ctx, hipEventSynchronize(e1));
    PM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    ms /= (float)reps;
    return PM_OK;
}

}  // namespace

namespace pm {

uint32_t tuned_flags(const pm_ctx* ctx, const pm_grid* grid, float radius, int32_t k, uint32_t flags) {
    if (!(flags & PM_KNN_TUNED)) return flags;
    flags &= ~(uint32_t)PM_KNN_TUNED;
    pm_gather_tuning t{};
    if (!grid || !find(ctx, k, density_class(grid->desc.n_photons, grid->desc.aabb, radius), t)) return flags;
    return (flags & ~kTuneBits) | t.flags;
}

float tuned_cell(const pm_ctx* ctx, uint32_t n, const float aabb[6], float radius, int32_t k, float fallback) {
    pm_gather_tuning t{};
    if (!find(ctx, k, density_class(n, aabb, radius), t)) return fallback;
    return t.cell_over_r * radius;
}

}  // namespace pm

extern "C" pm_status pm_tuning_lookup(const pm_ctx* ctx, uint32_t n, const float aabb[6], float radius, int32_t k,
                                      pm_gather_tuning* out) {
    if (!ctx || !aabb || !out) return PM_ERR_INVALID;
    pm_gather_tuning t{};
    if (!find(ctx, k, density_class(n, aabb, radius), t)) {
        t = pm_gather_tuning{};
        t.cell_over_r = 1.0f;
        t.found = 0;
    }
    *out = t;
    return PM_OK;
}

extern "C" pm_status pm_tune_gather(pm_ctx* ctx, const float* d_photons9, uint32_t n, const float aabb[6],
                                    const float* d_queries3, const float* d_normals3, uint32_t nq, float radius,
                                    int32_t k, pm_gather_tuning* out) {
    if (!ctx) return PM_ERR_INVALID;
    if (!out || !aabb || !d_photons9 || n == 0 || !d_queries3 || !d_normals3 || nq == 0)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_tune_gather: needs photons, queries and normals");
    if (!(radius > 0.f) || !std::isfinite(radius) || k < 1 || k > 2048)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_tune_gather: radius must be > 0 and k in [1, 2048]");
    static const float kFracs[] = {0.25f, 1.0f / 3.0f, 0.4f, 0.5f, 2.0f / 3.0f, 1.0f};
    // the variants of the one-walk gather; larger k has one kernel per range
    static const uint32_t kVariants[] = {0u, PM_KNN_XCD_CONTIGUOUS, PM_KNN_LARGE_BUFFER, PM_KNN_TWO_PASS,
                                         PM_KNN_INNER_BOX};
    const int n_var = k <= 256 ? 5 : 1;
    float* rgb = nullptr;
    PM_HIP(ctx, hipMalloc(&rgb, (size_t)nq * 12));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    pm_status st = PM_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        st = pm::fail(ctx, PM_ERR_HIP, "pm_tune_gather: events");
    pm_gather_tuning best{}, def{};
    best.ms = def.ms = -1.f;
    ctx->tune_table.assign(6 * 5, -1.f);
    for (int fi = 0; st == PM_OK && fi < 6; ++fi) {
        const float frac = kFracs[fi];
        pm_grid* g = nullptr;
        st = pm_grid_build(ctx, d_photons9, n, aabb, radius * frac, 1000, PM_LAYOUT_LINEAR, &g);
        if (st == PM_ERR_UNSUPPORTED) {  // too many cells at this size: skip it
            st = PM_OK;
            continue;
        }
        if (st != PM_OK) break;
        // (a grid of far more cells than photons only costs memory and scans)
        if (g->desc.n_cells > std::max<uint64_t>(8ull * n, 1ull << 24) && frac < 1.0f) {
            PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            pm_grid_destroy(g);
            continue;
        }
        float build_ms = 0.f;
        st = time_ms(ctx, e0, e1, 2, build_ms, [&] { return pm_grid_rebuild(ctx, g, d_photons9, n); });
        for (int vi = 0; st == PM_OK && vi < n_var; ++vi) {
            const uint32_t f = kVariants[vi] | PM_KNN_SORT_QUERIES;
            float ms = 0.f;
            st = time_ms(ctx, e0, e1, 2, ms, [&] {
                return pm_grid_gather_radiance(ctx, g, d_queries3, d_normals3, nq, radius, k, f, rgb, nullptr);
            });
            if (st != PM_OK) break;
            const float tot = build_ms + ms;
            ctx->tune_table[(size_t)fi * 5 + vi] = tot;
            if (best.ms < 0.f || tot < best.ms) {
                best.cell_over_r = frac;
                best.flags = kVariants[vi];
                best.ms = tot;
            }
            if (fi == 5 && vi == 0) def.ms = tot;
        }
        (void)hipStreamSynchronize(ctx->stream);
        pm_grid_destroy(g);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(rgb);
    if (st != PM_OK) return st;
    if (best.ms < 0.f) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_tune_gather: no configuration could be built");
    best.ms_default = def.ms;
    best.found = 1;
    const int cls = density_class(n, aabb, radius);
    auto it = std::find_if(ctx->tunes.begin(), ctx->tunes.end(),
                           [&](const pm::TuneEntry& e) { return e.k == k && e.cls == cls; });
    if (it != ctx->tunes.end())
        it->t = best;
    else
        ctx->tunes.push_back(pm::TuneEntry{k, cls, best});
    *out = best;
    return PM_OK;
}

extern "C" pm_status pm_tuning_table(const pm_ctx* ctx, float* h_ms, uint32_t cap) {
    if (!ctx || (!h_ms && cap)) return PM_ERR_INVALID;
    for (uint32_t i = 0; i < cap; ++i) h_ms[i] = i < ctx->tune_table.size() ? ctx->tune_table[i] : -1.f;
    return PM_OK;
}
