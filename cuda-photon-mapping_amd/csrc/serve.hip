// Serving loop: continuous batching of render jobs.
//
// The reference renders one scene, one frame at a time (Drawer.cu:18-101).
// Here a server holds many jobs (a resident scene, photon budgets, a camera,
// a number of frames). Each step produces the next frame of every active
// job: its photon pass is traced (re-seeded per frame), both grids are
// rebuilt in place, its eye rays are traced into its own frame state, and
// then the gathers of ALL active frames run as one pm_gather_batch before
// each frame is finished. With worker contexts (pm_server_set_workers) the
// jobs' trace / grid / eye-ray stages run concurrently, one host thread and
// one HIP stream per worker, so the host round trips of one job's trace
// overlap other jobs' kernels; the shared gather launch waits on every
// worker's stream. Jobs waiting in the queue are admitted at the start of a step
// as soon as their device footprint fits the memory budget, so arrivals join
// work already in flight (continuous batching); a finished job releases its
// memory at once. A job's frames are bit-identical alone and in any mix: every
// stage is deterministic and its inputs depend only on the job.
#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <new>
#include <thread>
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
    float* g9 = nullptr;  // photon maps (device; blocks of the server context's cache)
    float* c9 = nullptr;
    size_t g9_bytes = 0, c9_bytes = 0;
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
    // worker contexts for the per-job stages, each on a stream of its own
    std::vector<pm_ctx*> workers;
    std::vector<hipStream_t> wstreams;
    std::vector<hipEvent_t> wevents;
};

namespace {

void release(pm_server* S, Job& j) {
    if (j.gg) pm_grid_destroy(j.gg);
    if (j.gc) pm_grid_destroy(j.gc);
    j.gg = j.gc = nullptr;
    if (j.frame) pm_frame_destroy(j.frame);
    j.frame = nullptr;
    // (the server's later jobs reuse the maps' blocks: their stages are
    // ordered after this job's on the server's stream, which every worker
    // stream waits for at the start of a step)
    if (j.g9 && !pm::block_give(S->ctx, j.g9, j.g9_bytes)) (void)hipFree(j.g9);
    if (j.c9 && !pm::block_give(S->ctx, j.c9, j.c9_bytes)) (void)hipFree(j.c9);
    j.g9 = j.c9 = nullptr;
    S->in_use -= std::min(S->in_use, j.footprint);
}

pm_status admit(pm_server* S, Job& j) {
    pm_ctx* ctx = S->ctx;
    void* p = nullptr;
    hipError_t e = pm::block_take(ctx, (size_t)std::max(1u, j.d.n_global) * 36, &p, &j.g9_bytes);
    j.g9 = e == hipSuccess ? static_cast<float*>(p) : nullptr;
    if (e == hipSuccess) {
        p = nullptr;
        e = pm::block_take(ctx, (size_t)std::max(1u, j.d.n_caustic) * 36, &p, &j.c9_bytes);
        j.c9 = e == hipSuccess ? static_cast<float*>(p) : nullptr;
    }
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

namespace {

void drop_workers(pm_server* S) {
    for (hipStream_t w : S->wstreams) (void)hipStreamSynchronize(w);
    for (pm_ctx* c : S->workers) pm_ctx_destroy(c);
    for (hipEvent_t e : S->wevents) (void)hipEventDestroy(e);
    for (hipStream_t w : S->wstreams) (void)hipStreamDestroy(w);
    S->workers.clear();
    S->wevents.clear();
    S->wstreams.clear();
}

}  // namespace


extern "C" pm_status pm_server_set_workers(pm_server* S, int32_t n_workers) {
    if (!S) return PM_ERR_INVALID;
    pm_ctx* ctx = S->ctx;
    if (n_workers < 0 || n_workers > 64) return pm::fail(ctx, PM_ERR_INVALID, "pm_server_set_workers: 0..64 workers");
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_workers(S);
    for (int32_t w = 0; w < n_workers; ++w) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        pm_ctx* wc = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            drop_workers(S);
            return pm::fail(ctx, PM_ERR_HIP, "pm_server_set_workers: stream");
        }
        S->wstreams.push_back(st);
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            drop_workers(S);
            return pm::fail(ctx, PM_ERR_HIP, "pm_server_set_workers: event");
        }
        S->wevents.push_back(ev);
        pm_status rc = pm_ctx_create(ctx->device, st, &wc);
        if (rc != PM_OK) {
            drop_workers(S);
            return pm::fail(ctx, rc, "pm_server_set_workers: worker context");
        }
        S->workers.push_back(wc);
    }
    return PM_OK;
}

extern "C" pm_status pm_server_destroy(pm_server* S) {
    if (!S) return PM_OK;
    (void)hipStreamSynchronize(S->ctx->stream);
    drop_workers(S);
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

namespace {

// Grid of this frame's n photons: rebuilt in place when the job's grid holds
// them, else built afresh (cell: the job's, or its tuned size).
// The map holds n photons at the plane stride of its capacity.
pm_status frame_grid(pm_ctx* ctx, pm_grid*& g, const float* ph9, uint32_t n, size_t stride, const pm_job_desc& d,
                     float r, int32_t k) {
    const float cell =
        (d.params.gather_flags & PM_KNN_TUNED) ? pm::tuned_cell(ctx, n, d.aabb, r, k, d.cell_size) : d.cell_size;
    if (g && g->desc.cell_size == cell && n <= g->n_cap) return pm::grid_rebuild_strided(ctx, g, ph9, n, stride);
    if (g) {
        PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        pm_grid_destroy(g);
        g = nullptr;
    }
    return pm::grid_build_strided(ctx, ph9, n, stride, d.aabb, cell, 1000, PM_LAYOUT_LINEAR, &g);
}

// One job's stages before the gathers, on context `ctx`: its next frame's
// photon pass, both grids, its eye rays; its gather jobs into out[0..n).
pm_status prepare(pm_ctx* ctx, Job& j, pm_gather_job* out, uint32_t& n, uint32_t& nq) {
    const pm_job_desc& d = j.d;
    const uint64_t seed = d.seed + 0x9E3779B97F4A7C15ull * (j.done + 1);
    pm_trace_stats ts{};
    pm_status st;
    if ((st = pm_trace_photons(ctx, d.scene, seed, d.n_global, d.n_caustic, d.max_bounces, d.max_paths, j.g9,
                               d.n_caustic ? j.c9 : nullptr, &ts)) != PM_OK)
        return st;
    if ((st = frame_grid(ctx, j.gg, j.g9, ts.n_global, d.n_global, d, d.params.r_global, d.params.k_global)) !=
        PM_OK)
        return st;
    if (ts.n_caustic > 0) {
        if ((st = frame_grid(ctx, j.gc, j.c9, ts.n_caustic, d.n_caustic, d, d.params.r_caustic,
                             d.params.k_caustic)) != PM_OK)
            return st;
    } else if (j.gc) {
        PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        pm_grid_destroy(j.gc);
        j.gc = nullptr;
    }
    nq = 0;
    if ((st = pm_frame_rays(ctx, j.frame, d.scene, &d.camera, &d.params, &nq)) != PM_OK) return st;
    return pm_frame_gather_jobs(j.frame, j.gg, j.gc, out, &n);
}

}  // namespace

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
    // per job: trace, grids (rebuilt in place), eye rays -> its gather jobs
    const size_t na = S->active.size();
    std::vector<pm_gather_job> fj(2 * na);
    std::vector<uint32_t> nj(na, 0), nqs(na, 0);
    const size_t W = S->workers.size();
    if (W == 0) {
        for (size_t i = 0; i < na; ++i)
            if ((st = prepare(ctx, *S->active[i], &fj[2 * i], nj[i], nqs[i])) != PM_OK) return st;
    } else {
        // the workers' streams start after the main stream's earlier work
        // (this step's queued frames read nothing of it, but a caller's
        // work on the main stream may have produced the scenes)
        PM_HIP(ctx, hipEventRecord(S->wevents[0], ctx->stream));
        for (size_t w = 0; w < W; ++w) {
            PM_HIP(ctx, hipStreamWaitEvent(S->wstreams[w], S->wevents[0], 0));
            S->workers[w]->tunes = ctx->tunes;  // the tuned grid sizes
        }
        std::vector<pm_status> res(W, PM_OK);
        std::vector<std::thread> th;
        th.reserve(W);
        for (size_t w = 0; w < W; ++w)
            th.emplace_back([&, w] {
                for (size_t i = w; i < na && res[w] == PM_OK; i += W)
                    res[w] = prepare(S->workers[w], *S->active[i], &fj[2 * i], nj[i], nqs[i]);
            });
        for (auto& t : th) t.join();
        for (size_t w = 0; w < W; ++w)
            if (res[w] != PM_OK) {
                ctx->err = S->workers[w]->err;
                return res[w];
            }
        for (size_t w = 0; w < W; ++w) {
            PM_HIP(ctx, hipEventRecord(S->wevents[w], S->wstreams[w]));
            PM_HIP(ctx, hipStreamWaitEvent(ctx->stream, S->wevents[w], 0));
        }
    }
    std::vector<pm_gather_job> jobs;
    jobs.reserve(2 * na);
    for (size_t i = 0; i < na; ++i) {
        for (uint32_t t = 0; t < nj[i]; ++t) jobs.push_back(fj[2 * i + t]);
        S->queries += (uint64_t)nqs[i] * nj[i];
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
