// Per-frame pipeline enqueued without host synchronisation or allocation:
// photon trace -> both grids rebuilt in place -> eye rays -> both gathers ->
// per-pixel fold. The reference drives one frame per host call, with a
// synchronous copy or launch after every stage (Drawer.cu:18-57,
// Initializer.cu:142-177); here a frame is a fixed sequence of launches on
// the context's stream whose counts stay on the device, so a caller can
// capture it into a HIP graph once and replay it every frame.
//
// pm_pipeline_create renders one planning frame eagerly (pm_trace_photons,
// pm_grid_build, the frame's phases) to size everything: the photon maps and
// grids, the trace's path schedule (the planning frame's paths plus a
// quarter), the query block (the planning frame's hit count: the ray trees
// depend only on the scene and the camera), and the context's scratch,
// which one pipelined frame then grows to its final size.
// pm_pipeline_frame only enqueues. The maps, grids and image equal what the
// eager calls produce for the same seed whenever the scheduled paths fill
// the maps, which pm_pipeline_status reports.
#include <algorithm>
#include <new>

#include "pm_internal.hpp"

struct pm_pipeline {
    pm_ctx* ctx = nullptr;
    pm_pipeline_desc d{};
    void* maps = nullptr;  // global and caustic photon maps, photons9 each
    float* g9 = nullptr;
    float* c9 = nullptr;
    pm_grid* gg = nullptr;
    pm_grid* gc = nullptr;
    pm::TracePlan* trace = nullptr;
    pm::FramePlan* frame = nullptr;
    uint64_t paths = 0;
};

extern "C" pm_status pm_pipeline_destroy(pm_pipeline* pl) {
    if (!pl) return PM_OK;
    if (pl->ctx) (void)hipStreamSynchronize(pl->ctx->stream);
    pm::frame_plan_destroy(pl->frame);
    pm::trace_plan_destroy(pl->trace);
    pm_grid_destroy(pl->gg);
    pm_grid_destroy(pl->gc);
    if (pl->maps) (void)hipFree(pl->maps);
    delete pl;
    return PM_OK;
}

extern "C" pm_status pm_pipeline_create(pm_ctx* ctx, const pm_pipeline_desc* desc, pm_pipeline** out) {
    if (!ctx || !desc || !out) return PM_ERR_INVALID;
    *out = nullptr;
    const pm_pipeline_desc& d = *desc;
    if (!d.scene || d.n_global == 0) return pm::fail(ctx, PM_ERR_INVALID, "pm_pipeline_create: incomplete description");
    if (!(d.cell_global > 0.f) || (d.n_caustic > 0 && !(d.cell_caustic > 0.f)))
        return pm::fail(ctx, PM_ERR_INVALID, "pm_pipeline_create: grid cell sizes must be > 0");
    pm_pipeline* pl = new (std::nothrow) pm_pipeline();
    if (!pl) return pm::fail(ctx, PM_ERR_NOMEM, "pm_pipeline_create: host allocation");
    pl->ctx = ctx;
    pl->d = d;
    auto bail = [&](pm_status st) {
        pm_pipeline_destroy(pl);
        return st;
    };
    const size_t gbytes = (size_t)d.n_global * 36, cbytes = (size_t)std::max(1u, d.n_caustic) * 36;
    hipError_t e = hipMalloc(&pl->maps, pm::carve_size({gbytes, cbytes}));
    if (e != hipSuccess) return bail(pm::hip_fail(ctx, e, "hipMalloc(pipeline maps)"));
    pm::Carve cv{static_cast<char*>(pl->maps)};
    pl->g9 = cv.take<float>((size_t)d.n_global * 9);
    pl->c9 = cv.take<float>((size_t)std::max(1u, d.n_caustic) * 9);

    // planning frame: the eager trace and grid builds
    pm_status st;
    pm_trace_stats ts{};
    if ((st = pm_trace_photons(ctx, d.scene, d.plan_seed, d.n_global, d.n_caustic, d.max_bounces, d.max_paths, pl->g9,
                               d.n_caustic ? pl->c9 : nullptr, &ts)) != PM_OK)
        return bail(st);
    if (ts.n_global < d.n_global || ts.n_caustic < d.n_caustic)
        return bail(pm::fail(ctx, PM_ERR_CAPACITY,
                             "pm_pipeline_create: %llu paths filled only %u / %u global and %u / %u caustic photons",
                             (unsigned long long)ts.paths, ts.n_global, d.n_global, ts.n_caustic, d.n_caustic));
    if ((st = pm_grid_build(ctx, pl->g9, d.n_global, d.aabb, d.cell_global, 1000, PM_LAYOUT_LINEAR, &pl->gg)) != PM_OK)
        return bail(st);
    if (d.n_caustic > 0 &&
        (st = pm_grid_build(ctx, pl->c9, d.n_caustic, d.aabb, d.cell_caustic, 1000, PM_LAYOUT_LINEAR, &pl->gc)) != PM_OK)
        return bail(st);
    // the planning frame's render: its hit count sizes the query block (the
    // ray trees depend only on the scene and the camera)
    pm_frame_stats fs{};
    if ((st = pm_render_frame(ctx, d.scene, pl->gg, pl->gc, &d.camera, &d.params, nullptr, nullptr, &fs)) != PM_OK)
        return bail(st);
    // the path schedule: the planning frame's paths and a quarter more
    pl->paths = std::min<uint64_t>(d.max_paths, ts.paths + ts.paths / 4 + 4096);
    if ((st = pm::trace_plan_create(ctx, d.scene, d.n_global, d.n_caustic, d.max_bounces, pl->paths, &pl->trace)) !=
        PM_OK)
        return bail(st);
    if ((st = pm::frame_plan_create(ctx, d.scene, &d.camera, &d.params, fs.hits, &pl->frame)) != PM_OK)
        return bail(st);
    // one pipelined frame grows the context's scratch and scan state to what
    // every later frame uses
    if ((st = pm_pipeline_frame(pl, d.plan_seed, nullptr, nullptr)) != PM_OK) return bail(st);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return bail(pm::fail(ctx, PM_ERR_HIP, "pm_pipeline_create"));
    *out = pl;
    return PM_OK;
}

extern "C" pm_status pm_pipeline_frame(pm_pipeline* pl, uint64_t seed, float* d_rgb3, uint8_t* d_rgb8) {
    if (!pl) return PM_ERR_INVALID;
    pm_ctx* ctx = pl->ctx;
    const pm_pipeline_desc& d = pl->d;
    pm_status st;
    if ((st = pm::trace_plan_enqueue(ctx, pl->trace, seed, pl->g9, pl->c9)) != PM_OK) return st;
    if ((st = pm_grid_rebuild(ctx, pl->gg, pl->g9, d.n_global)) != PM_OK) return st;
    if (pl->gc && (st = pm_grid_rebuild(ctx, pl->gc, pl->c9, d.n_caustic)) != PM_OK) return st;
    return pm::frame_plan_enqueue(ctx, pl->frame, pl->gg, pl->gc, d_rgb3, d_rgb8);
}

extern "C" pm_status pm_pipeline_status(pm_pipeline* pl, pm_pipeline_info* out) {
    if (!pl || !out) return PM_ERR_INVALID;
    pm_ctx* ctx = pl->ctx;
    uint64_t h[3] = {0, 0, 0};
    PM_HIP(ctx, hipMemcpyAsync(h, pm::trace_plan_counts(pl->trace), 16, hipMemcpyDeviceToHost, ctx->stream));
    PM_HIP(ctx, hipMemcpyAsync(h + 2, pm::frame_plan_hits(pl->frame), 8, hipMemcpyDeviceToHost, ctx->stream));
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out->scheduled_paths = pl->paths;
    out->n_global = (uint32_t)std::min<uint64_t>(h[0], pl->d.n_global);
    out->n_caustic = (uint32_t)std::min<uint64_t>(h[1], pl->d.n_caustic);
    out->complete =
        h[0] >= pl->d.n_global && h[1] >= pl->d.n_caustic && h[2] <= pm::frame_plan_capacity(pl->frame) ? 1 : 0;
    out->hits = h[2];
    return PM_OK;
}

extern "C" pm_status pm_pipeline_grids(const pm_pipeline* pl, const pm_grid** global_grid,
                                       const pm_grid** caustic_grid) {
    if (!pl) return PM_ERR_INVALID;
    if (global_grid) *global_grid = pl->gg;
    if (caustic_grid) *caustic_grid = pl->gc;
    return PM_OK;
}
