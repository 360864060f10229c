// Serving loop: continuous batching of render jobs on one context.
//
// The reference renders one scene, one frame at a time (Drawer.cu:18-101).
// Here a server holds many jobs (a resident scene, photon budgets, a camera,
// a number of frames). Each step produces the next frame of every active
// job: its photon pass is traced (re-seeded per frame), both grids are
// rebuilt, its eye rays are traced into its own frame state, and then the
// gathers of ALL active frames run as one pm_gather_batch before each frame
// is finished. Jobs waiting in the queue are admitted at the start of a step
// as soon as their device footprint fits the memory budget, so arrivals join
// work already in flight (continuous batching); a finished job releases its
// memory at once. A job's frames are bit-identical alone and in any mix: every
// stage is deterministic and its inputs depend only on the job.
#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <new>
#include <vector>

#include "pm_internal.hpp"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Job {
    uint64_t id = 0;
    pm_job_desc d{};
    size_t footprint = 0;
    uint32_t done = 0;
    double ready_at = 0;  // when the next frame became eligible
    float* g9 = nullptr;  // photon maps (device)
    float* c9 = nullptr;
    pm_frame* frame = nullptr;
    pm_grid* gg = nullptr;
    pm_grid* gc = nullptr;
};

size_t job_footprint(const pm_job_desc& d) {
    const size_t n = (size_t)d.n_global + d.n_caustic;
    const size_t npix = (size_t)std::max(d.camera.width, 0) * (size_t)std::max(d.camera.height, 0);
    const size_t hmax = ((size_t)1 << std::max(1, std::min(d.params.max_depth, 6))) - 1;
    // maps + grid copies (photon block, perm) + frame (hit slots, query block)
    return n * 36 + n * 40 + npix * hmax * (40 + 60) + npix * 16 + ((size_t)64 << 20);
}

}  // namespace

struct pm_server {
    pm_ctx* ctx = nullptr;
    size_t budget = 0, in_use = 0;
    uint64_t next_id = 1;
    std::deque<std::unique_ptr<Job>> queue;
    std::vector<std::unique_ptr<Job>> active;
    std::vector<uint64_t> finished;   // ids
    std::vector<double> latencies;    // seconds, one per completed frame
    uint64_t queries = 0;
};

namespace {

void release(pm_server* S, Job& j) {
    if (j.gg) pm_grid_destroy(j.gg);
    if (j.gc) pm_grid_destroy(j.gc);
    j.gg = j.gc = nullptr;
    if (j.frame) pm_frame_destroy(j.frame);
    j.frame = nullptr;
    if (j.g9) (void)hipFree(j.g9);
    if (j.c9) (void)hipFree(j.c9);
    j.g9 = j.c9 = nullptr;
    S->in_use -= std::min(S->in_use, j.footprint);
}

pm_status admit(pm_server* S, Job& j) {
    pm_ctx* ctx = S->ctx;
    hipError_t e = hipMalloc(&j.g9, (size_t)std::max(1u, j.d.n_global) * 36);
    if (e == hipSuccess) e = hipMalloc(&j.c9, (size_t)std::max(1u, j.d.n_caustic) * 36);
    if (e != hipSuccess) {
        release(S, j);
        return pm::hip_fail(ctx, e, "pm_server: photon maps");
    }
    pm_status st = pm_frame_create(ctx, &j.frame);
    if (st != PM_OK) {
        release(S, j);
        return st;
    }
    S->in_use += j.footprint;
    return PM_OK;
}

}  // namespace

extern "C" pm_status pm_server_create(pm_ctx* ctx, size_t budget_bytes, pm_server** out) {
    if (!ctx || !out) return PM_ERR_INVALID;
    *out = nullptr;
    pm_server* S = new (std::nothrow) pm_server();
    if (!S) return pm::fail(ctx, PM_ERR_NOMEM, "pm_server_create: host allocation");
    S->ctx = ctx;
    if (budget_bytes == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
        budget_bytes = free_b / 10 * 8;
    }
    S->budget = budget_bytes;
    *out = S;
    return PM_OK;
}

extern "C" pm_status pm_server_destroy(pm_server* S) {
    if (!S) return PM_OK;
    (void)hipStreamSynchronize(S->ctx->stream);
    for (auto& j : S->active) release(S, *j);
    delete S;
    return PM_OK;
}

extern "C" pm_status pm_server_submit(pm_server* S, const pm_job_desc* desc, uint64_t* job_id) {
    if (!S || !desc) return PM_ERR_INVALID;
    pm_ctx* ctx = S->ctx;
    const pm_job_desc& d = *desc;
    if (!d.scene || d.n_global == 0 || d.frames == 0 || !d.d_rgb8)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_server_submit: incomplete job");
    if (d.params.flags & PM_RENDER_REFERENCE_CAUSTIC)
        return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_server_submit: reference-caustic mode renders one frame at a time");
    auto j = std::make_unique<Job>();
    j->id = S->next_id++;
    j->d = d;
    j->footprint = job_footprint(d);
    if (j->footprint > S->budget) return pm::fail(ctx, PM_ERR_CAPACITY, "pm_server_submit: job exceeds the budget");
    j->ready_at = now_s();
    if (job_id) *job_id = j->id;
    S->queue.push_back(std::move(j));
    return PM_OK;
}

// One step: admit what fits, then the next frame of every active job with
// all gathers in one batch. *frames_done = frames completed by this step.
extern "C" pm_status pm_server_step(pm_server* S, uint32_t* frames_done) {
    if (!S) return PM_ERR_INVALID;
    pm_ctx* ctx = S->ctx;
    if (frames_done) *frames_done = 0;
    pm_status st;
    while (!S->queue.empty() && S->in_use + S->queue.front()->footprint <= S->budget) {
        std::unique_ptr<Job> j = std::move(S->queue.front());
        S->queue.pop_front();
        if ((st = admit(S, *j)) != PM_OK) return st;
        S->active.push_back(std::move(j));
    }
    if (S->active.empty()) return PM_OK;
    std::vector<pm_gather_job> jobs;
    jobs.reserve(2 * S->active.size());
    for (auto& jp : S->active) {
        Job& j = *jp;
        const pm_job_desc& d = j.d;
        const uint64_t seed = d.seed + 0x9E3779B97F4A7C15ull * (j.done + 1);
        pm_trace_stats ts{};
        if ((st = pm_trace_photons(ctx, d.scene, seed, d.n_global, d.n_caustic, d.max_bounces, d.max_paths, j.g9,
                                   d.n_caustic ? j.c9 : nullptr, &ts)) != PM_OK)
            return st;
        if (j.gg) pm_grid_destroy(j.gg);
        if (j.gc) pm_grid_destroy(j.gc);
        j.gg = j.gc = nullptr;
        if ((st = pm_grid_build(ctx, j.g9, ts.n_global, d.aabb, d.cell_size, 1000, PM_LAYOUT_LINEAR,
                                &j.gg)) != PM_OK)
            return st;
        if (ts.n_caustic > 0 &&
            (st = pm_grid_build(ctx, j.c9, ts.n_caustic, d.aabb, d.cell_size, 1000, PM_LAYOUT_LINEAR, &j.gc)) != PM_OK)
            return st;
        uint32_t nq = 0;
        if ((st = pm_frame_rays(ctx, j.frame, d.scene, &d.camera, &d.params, &nq)) != PM_OK) return st;
        pm_gather_job fj[2];
        uint32_t nj = 0;
        if ((st = pm_frame_gather_jobs(j.frame, j.gg, j.gc, fj, &nj)) != PM_OK) return st;
        for (uint32_t i = 0; i < nj; ++i) jobs.push_back(fj[i]);
        S->queries += (uint64_t)nq * nj;
    }
    if ((st = pm_gather_batch(ctx, jobs.data(), (uint32_t)jobs.size(), PM_KNN_SORT_QUERIES)) != PM_OK) return st;
    for (auto& jp : S->active) {
        Job& j = *jp;
        const size_t npix = (size_t)j.d.camera.width * (size_t)j.d.camera.height;
        if ((st = pm_frame_finish(ctx, j.frame, nullptr, j.d.d_rgb8 + (size_t)j.done * npix * 3, nullptr)) != PM_OK)
            return st;
    }
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double t = now_s();
    uint32_t n_done = 0;
    std::vector<std::unique_ptr<Job>> keep;
    for (auto& jp : S->active) {
        Job& j = *jp;
        S->latencies.push_back(t - j.ready_at);
        j.ready_at = t;
        ++j.done;
        ++n_done;
        if (j.done == j.d.frames) {
            release(S, j);
            S->finished.push_back(j.id);
        } else {
            keep.push_back(std::move(jp));
        }
    }
    S->active = std::move(keep);
    if (frames_done) *frames_done = n_done;
    return PM_OK;
}

extern "C" pm_status pm_server_stats(const pm_server* S, pm_server_info* out) {
    if (!S || !out) return PM_ERR_INVALID;
    out->queued = (uint32_t)S->queue.size();
    out->active = (uint32_t)S->active.size();
    out->finished = (uint32_t)S->finished.size();
    out->frames = (uint64_t)S->latencies.size();
    out->queries = S->queries;
    out->bytes_in_use = S->in_use;
    return PM_OK;
}

// Frame latencies (seconds) recorded so far, oldest first.
extern "C" pm_status pm_server_latencies(const pm_server* S, double* out, uint64_t cap, uint64_t* n) {
    if (!S || !n) return PM_ERR_INVALID;
    *n = S->latencies.size();
    if (out)
        for (uint64_t i = 0; i < std::min<uint64_t>(cap, S->latencies.size()); ++i) out[i] = S->latencies[i];
    return PM_OK;
}
