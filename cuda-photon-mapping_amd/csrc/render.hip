// Eye-ray frame: Raytracer::render_gpu / render_trace (Raytracer.cuh:140-279)
// with RaytracePlanner's deferred refraction rays (RaytracePlanner.cuh:30-69).
//
// The reference runs each pixel's whole ray tree in one thread, calling the
// kNN gather from inside it through a global-memory heap per thread
// (Drawer.cu:87-101). Here the frame is a wavefront:
//   1. k_render_rays: one lane per pixel (8x8 pixel tiles per wave) walks the
//      ray tree — primary ray, mirror reflections, a LIFO of deferred
//      refraction rays — and records every hit (point, normal, model, ray
//      coefficient, flags) in per-pixel slots, in the reference's visit order;
//   2. the hits are compacted (wave64 scan) into one query block;
//   3. the global and caustic gathers run as two batched, cell-ordered
//      k-nearest + radiance launches (the gather kernel of knn.hip);
//   4. k_render_combine: one lane per pixel folds its hits in order —
//      4·Kd·opaque·L_global (outside objects, Kd != 0) + emission +
//      opaque·L_caustic (when the caustic set is full), clamped to 1, times the
//      ray coefficient — then HDR, clamp and the uchar3 conversion.
// This is exact because the ray tree depends only on geometry and materials,
// never on the photon estimates.
//
// Media: the reference's MediumManager (MediumManager.cuh:30-88) reads inner
// stack entries it never wrote, so its refraction indices are undefined; here
// refraction follows the photon tracer's medium rule (PhotonTracing.cu:127-151)
// with one stack of (ior, id) per deferred ray, depth (max_depth + 1) / 2.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "pm_geom.hpp"

namespace {

using namespace pmg;

constexpr int kMaxDepth = 6;
constexpr int kMaxMedium = (kMaxDepth + 1) / 2;

struct Cam {
    float pos[3], fwd[3], right[3], up[3];
    float vw, vh, focal;
    int32_t w, h, pad;
};

struct RaysParams {
    const float* tris;
    const float* nrms;
    const DevModel* models;
    int n_models;
    const float4* bvh;
    float def_ior;
    Cam cam;
    int max_depth;
    int hmax;
    // per-pixel hit slots [hmax][npix]
    float* hp;   // 3 planes: hit point
    float* hn;   // 3 planes: normal
    float* hc;   // 3 planes: ray coefficient
    int* hm;     // model index | flags << 16 (bit 0: outside objects)
    uint32_t* count;  // [npix]
    uint32_t npix;
};

struct Plan {
    V3 o, d;
    V3 coef;
    int depth;
    int ms;
    float mn[kMaxMedium];
    int mid[kMaxMedium];
};

__global__ __launch_bounds__(256) void k_render_rays(RaysParams P) {
    __shared__ uint2 s_stack[kLdsStack * kStackLanes];
    const BvhStack S{s_stack + threadIdx.x};
    // 8x8 pixel tile per wave: coherent rays share model AABB outcomes
    const uint32_t lane = threadIdx.x & 63, wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int W = P.cam.w, H = P.cam.h;
    const int tiles_x = (W + 7) / 8;
    const int tx = (int)(wave % (uint32_t)tiles_x), ty = (int)(wave / (uint32_t)tiles_x);
    const int x = tx * 8 + (int)(lane & 7), y = ty * 8 + (int)(lane >> 3);
    if (ty * 8 >= H) return;  // wave-uniform
    const bool active = x < W && y < H;
    const uint32_t pix = active ? (uint32_t)(y * W + x) : 0u;

    // Camera::generate_ray_direction (Camera.cuh:67-81); Ray ctor normalises
    const Cam& c = P.cam;
    const float u = ((float)x + 0.5f) / W, v = ((float)y + 0.5f) / H;
    const V3 pos = v3(c.pos[0], c.pos[1], c.pos[2]);
    const V3 horiz = mul(v3(c.right[0], c.right[1], c.right[2]), c.vw);
    const V3 vert = mul(v3(c.up[0], c.up[1], c.up[2]), c.vh);
    const V3 ll = sub(sub(mul(add(pos, v3(c.fwd[0], c.fwd[1], c.fwd[2])), c.focal), mul(horiz, 0.5f)), mul(vert, 0.5f));
    Ray ray;
    ray.o = pos;
    ray.d = normalize_div(sub(add(add(ll, mul(horiz, u)), mul(vert, v)), pos));

    Plan plans[kMaxDepth];
    int nplans = 0;
    int new_depth = 0, cur_depth = -1;
    V3 coef = v3(1, 1, 1);
    bool in_obj = false;
    int ms = 0;
    float mn[kMaxMedium];
    int mid[kMaxMedium];
    uint32_t nh = 0;
    const int max_medium = (P.max_depth + 1) / 2;
    bool alive = active;
    // render_trace's loop (Raytracer.cuh:147-236); the wave runs until every
    // lane's tree is done (each lane's loop is bounded by 2^max_depth rays)
    for (int guard = 0; guard < (1 << kMaxDepth) + 4; ++guard) {
        if (alive && !(cur_depth < P.max_depth && cur_depth != new_depth)) {
            if (nplans == 0) {
                alive = false;
            } else {
                const Plan& p = plans[--nplans];
                ray.o = p.o;
                ray.d = p.d;
                new_depth = p.depth;
                coef = p.coef;
                ms = p.ms;
                in_obj = ms > 0;
                for (int s = 0; s < kMaxMedium; ++s) {
                    mn[s] = p.mn[s];
                    mid[s] = p.mid[s];
                }
            }
        }
        if (!__builtin_amdgcn_ballot_w64(alive)) break;
        if (alive) cur_depth = new_depth;
        float th = 0.f;
        V3 n = v3(0, 0, 0);
        int hmod = -1;
        // all lanes join the (wave-uniform) triangle loops; dead lanes shoot a
        // null ray that hits nothing
        Ray rr = ray;
        if (!alive) rr.d = v3(0, 0, 0);
        bool hit = scene_hit(P.models, P.n_models, P.tris, P.nrms, P.bvh, rr, th, n, hmod, S) && alive;
        if (!hit) continue;
        // Raytracer::find_intersection: point = dir*t + origin; normal turned
        // toward the ray only inside objects
        V3 ip = add(mul(ray.d, th), ray.o);
        if (in_obj && dot(ray.d, n) > 0) n = mul(n, -1.f);
        const DevModel& mat = P.models[hmod];
        if (nh < (uint32_t)P.hmax) {
            size_t s = (size_t)nh * P.npix + pix, plane = (size_t)P.hmax * P.npix;
            P.hp[s] = ip.x;
            P.hp[plane + s] = ip.y;
            P.hp[2 * plane + s] = ip.z;
            P.hn[s] = n.x;
            P.hn[plane + s] = n.y;
            P.hn[2 * plane + s] = n.z;
            P.hc[s] = coef.x;
            P.hc[plane + s] = coef.y;
            P.hc[2 * plane + s] = coef.z;
            P.hm[s] = hmod | ((in_obj ? 0 : 1) << 16);
            ++nh;
        }
        if (mat.opaque < 1.f && cur_depth != P.max_depth - 1) {
            float n1, n2;
            int nms = ms;
            float nmn[kMaxMedium];
            int nmid[kMaxMedium];
            for (int s = 0; s < kMaxMedium; ++s) {
                nmn[s] = mn[s];
                nmid[s] = mid[s];
            }
            if (ms > 0 && mid[ms - 1] == mat.id) {
                n1 = mn[ms - 1];
                n2 = ms >= 2 ? mn[ms - 2] : P.def_ior;
                nms = ms - 1;
            } else {
                n1 = ms > 0 ? mn[ms - 1] : P.def_ior;
                n2 = mat.ior;
                if (ms < max_medium) {
                    nmn[ms] = mat.ior;
                    nmid[ms] = mat.id;
                    nms = ms + 1;
                }
            }
            // Ray::refract (Ray.cuh:85-96)
            float eta = n1 / n2;
            float c1 = -dot(ray.d, n);
            float w = eta * c1;
            float c2m = (w - eta) * (w + eta);
            if (!(c2m < -1.f) && nplans < P.max_depth) {
                Plan& p = plans[nplans++];
                p.d = add(mul(ray.d, eta), mul(n, w - sqrtf(1.f + c2m)));
                p.o = add(ip, mul(neg(n), 0.01f));
                p.depth = cur_depth + 1;
                p.coef = mul(coef, 1.f - mat.opaque);
                p.ms = nms;
                for (int s = 0; s < kMaxMedium; ++s) {
                    p.mn[s] = nmn[s];
                    p.mid[s] = nmid[s];
                }
            }
        }
        if (!in_obj && !(mat.ks[0] == 0.f && mat.ks[1] == 0.f && mat.ks[2] == 0.f) && cur_depth != P.max_depth - 1) {
            // Ray::reflect (Ray.cuh:64-71) and the Phong-style coefficient
            float dnd = dot(ray.d, n);
            Ray nr;
            nr.d = normalize_div(sub(ray.d, mul(mul(n, 2.f), dnd)));
            nr.o = add(ip, mul(n, 0.01f));
            float cc = powf(fmaxf(dot(nr.d, neg(ray.d)), 0.0f), mat.shininess);
            coef = mulv(coef, mul(mul(v3(mat.ks[0], mat.ks[1], mat.ks[2]), mat.opaque), cc));
            ray = nr;
            new_depth = cur_depth + 1;
        }
    }
    if (active) P.count[pix] = nh;
}

// Slot (pixel, h) -> compacted query q = off[pixel] + h: query point,
// normal, and the inverse map for the combine pass.
__global__ __launch_bounds__(256) void k_render_compact(const float* __restrict__ hp, const float* __restrict__ hn,
                                                       const uint32_t* __restrict__ count,
                                                       const uint32_t* __restrict__ off, uint32_t npix, int hmax,
                                                       uint32_t nq, float* __restrict__ q3, float* __restrict__ n3) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    uint32_t c = count[pix], o = off[pix];
    size_t plane = (size_t)hmax * npix;
    for (uint32_t h = 0; h < c; ++h) {
        size_t s = (size_t)h * npix + pix;
        uint32_t q = o + h;
        if (q >= nq) break;  // (a query block sized below the hit count keeps its first nq)
        q3[q] = hp[s];
        q3[nq + q] = hp[plane + s];
        q3[2 * (size_t)nq + q] = hp[2 * plane + s];
        n3[q] = hn[s];
        n3[nq + q] = hn[plane + s];
        n3[2 * (size_t)nq + q] = hn[2 * plane + s];
    }
}

__global__ __launch_bounds__(256) void k_render_combine(const DevModel* __restrict__ models,
                                                       const float* __restrict__ hc, const int* __restrict__ hm,
                                                       const uint32_t* __restrict__ count,
                                                       const uint32_t* __restrict__ off, uint32_t npix, int hmax,
                                                       uint32_t nq, const float* __restrict__ lg,
                                                       const float* __restrict__ lc,
                                                       const uint32_t* __restrict__ cc_count, int k_caustic,
                                                       int have_caustic, float* __restrict__ rgb3,
                                                       uint8_t* __restrict__ rgb8) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    uint32_t c = count[pix], o = off[pix];
    size_t plane = (size_t)hmax * npix;
    V3 acc = v3(0, 0, 0);
    for (uint32_t h = 0; h < c; ++h) {
        size_t s = (size_t)h * npix + pix;
        uint32_t q = o + h;
        if (q >= nq) break;
        int code = hm[s];
        const DevModel& mat = models[code & 0xFFFF];
        bool outside = (code >> 16) & 1;
        V3 cur = v3(0, 0, 0);
        if (outside && !(mat.kd[0] == 0.f && mat.kd[1] == 0.f && mat.kd[2] == 0.f)) {
            V3 di_op = mul(v3(mat.kd[0], mat.kd[1], mat.kd[2]), mat.opaque);
            V3 re = v3(lg[q], lg[nq + q], lg[2 * (size_t)nq + q]);
            cur = add(cur, mul(mulv(re, di_op), 4.f));
        }
        cur = add(cur, v3(mat.ke[0], mat.ke[1], mat.ke[2]));
        if (have_caustic && cc_count[q] == (uint32_t)k_caustic) {
            V3 t = v3(lc[q], lc[nq + q], lc[2 * (size_t)nq + q]);
            cur = add(cur, mul(t, mat.opaque));
        }
        cur = v3(fminf(cur.x, 1.f), fminf(cur.y, 1.f), fminf(cur.z, 1.f));
        acc = add(acc, mulv(cur, v3(hc[s], hc[plane + s], hc[2 * plane + s])));
    }
    // HDR (Raytracer.cuh:95-107), clamp_min(1), uchar3 (:238-239, :277)
    const float t = 1.0f / 0.6f;  // pow(exposure, -1) with exposure = 0.6f
    const float rgam = 1.0f / 2.2f;
    const float o3[3] = {acc.x, acc.y, acc.z};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float e = expf(-t * o3[ch]);
        float mapped = fminf(powf(1.0f - e, rgam), 1.f);
        if (rgb3) rgb3[(size_t)ch * npix + pix] = mapped;
        if (rgb8) rgb8[3 * (size_t)pix + ch] = (uint8_t)(mapped * 255);
    }
}

}  // namespace

// Per-frame device state: the ray trees' hit slots, the compacted query
// block and the two radiance planes, in one allocation that grows as needed.
struct pm_frame {
    pm_ctx* owner = nullptr;  // the context whose block cache `mem` returns to
    void* mem = nullptr;
    size_t cap = 0;
    uint32_t W = 0, H = 0, npix = 0, nq = 0;
    int hmax = 0, nm = 0;
    pm_render_params R{};
    bool rays_done = false, caustic = false;
    DevModel* d_models = nullptr;
    float *hp = nullptr, *hn = nullptr, *hc = nullptr;
    int* hm = nullptr;
    uint32_t *count = nullptr, *off = nullptr;
    uint64_t* total = nullptr;
    float *q3 = nullptr, *n3 = nullptr, *lg = nullptr, *lc = nullptr;
    uint32_t* ccount = nullptr;
    uint32_t* kidx = nullptr;
    float* kdist = nullptr;
};

namespace {

// Grow the frame's allocation to `need` bytes, keeping its first `keep`.
// Blocks come from and go back to the context's block cache: a frame's old
// block is reused by later work on the same stream, so only a frame that
// moves to another context synchronises and frees it.
pm_status frame_grow(pm_ctx* ctx, pm_frame* F, size_t need, size_t keep) {
    if (need <= F->cap) return PM_OK;
    void* fresh = nullptr;
    size_t got = 0;
    const size_t want = need + need / 8;
    hipError_t e = pm::block_take(ctx, want, &fresh, &got);
    if (e != hipSuccess) return pm::hip_fail(ctx, e, "hipMalloc(frame)");
    if (F->mem) {
        if (keep) PM_HIP(ctx, hipMemcpyAsync(fresh, F->mem, keep, hipMemcpyDeviceToDevice, ctx->stream));
        if (F->owner != ctx || !pm::block_give(ctx, F->mem, F->cap)) {
            // (the frame's last context may be gone: wait for the device)
            PM_HIP(ctx, F->owner == ctx ? hipStreamSynchronize(ctx->stream) : hipDeviceSynchronize());
            PM_HIP(ctx, hipFree(F->mem));
        }
    }
    F->mem = fresh;
    F->cap = got;
    F->owner = ctx;
    return PM_OK;
}

size_t rays_bytes(const pm_frame* F) {
    const size_t slot = (size_t)F->hmax * F->npix;
    return pm::carve_size({(size_t)F->nm * sizeof(DevModel), (size_t)F->nm * 32, slot * 12, slot * 12, slot * 12,
                           slot * 4, (size_t)F->npix * 4, (size_t)F->npix * 4, 64});
}

void carve_rays(pm_frame* F, pm::Carve& cv) {
    const size_t slot = (size_t)F->hmax * F->npix;
    F->d_models = cv.take<DevModel>(F->nm);
    cv.take<char>((size_t)F->nm * 32);
    F->hp = cv.take<float>(slot * 3);
    F->hn = cv.take<float>(slot * 3);
    F->hc = cv.take<float>(slot * 3);
    F->hm = cv.take<int>(slot);
    F->count = cv.take<uint32_t>(F->npix);
    F->off = cv.take<uint32_t>(F->npix);
    F->total = cv.take<uint64_t>(1);
}

bool quirk_mode(const pm_render_params& R) { return (R.flags & PM_RENDER_REFERENCE_CAUSTIC) != 0; }

}  // namespace

extern "C" pm_status pm_frame_create(pm_ctx* ctx, pm_frame** out) {
    if (!ctx || !out) return PM_ERR_INVALID;
    *out = new (std::nothrow) pm_frame();
    return *out ? PM_OK : pm::fail(ctx, PM_ERR_NOMEM, "pm_frame_create: host allocation");
}

extern "C" pm_status pm_frame_destroy(pm_frame* F) {
    if (!F) return PM_OK;
    if (F->mem && !pm::block_give(F->owner, F->mem, F->cap)) (void)hipFree(F->mem);
    delete F;
    return PM_OK;
}

// Phase 1: every pixel's ray tree (hits in the reference's visit order) and
// the compacted query block; returns the number of hits (queries per map).
extern "C" pm_status pm_frame_rays(pm_ctx* ctx, pm_frame* F, const pm_scene* scene, const pm_camera* camera,
                                   const pm_render_params* params, uint32_t* n_queries) {
    if (!ctx || !F || !scene || !camera || !params) return PM_ERR_INVALID;
    F->rays_done = false;
    if (!scene->d_tris9 || !scene->d_normals9 || !scene->h_models || scene->n_models <= 0 || scene->n_models > 0xFFFF)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_frame_rays: empty scene");
    const pm_render_params& R = *params;
    if (R.k_global < 1 || R.k_global > 2048 || R.k_caustic < 1 || R.k_caustic > 2048)
        return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_frame_rays: k outside [1, 2048]");
    if (R.max_depth < 1 || R.max_depth > kMaxDepth)
        return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_frame_rays: max_depth outside [1, %d]", kMaxDepth);
    if (camera->width <= 0 || camera->height <= 0) return pm::fail(ctx, PM_ERR_INVALID, "pm_frame_rays: empty canvas");
    const uint64_t npix64 = (uint64_t)camera->width * (uint64_t)camera->height;
    if (npix64 > (1ull << 26)) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_frame_rays: canvas too large");
    F->W = (uint32_t)camera->width;
    F->H = (uint32_t)camera->height;
    F->npix = (uint32_t)npix64;
    F->hmax = (1 << R.max_depth) - 1;
    F->nm = scene->n_models;
    F->R = R;
    pm_status st = frame_grow(ctx, F, rays_bytes(F), 0);
    if (st != PM_OK) return st;
    pm::Carve cv{static_cast<char*>(F->mem)};
    carve_rays(F, cv);
    pm::Carve ct{static_cast<char*>(F->mem)};  // the model-AABB scratch right after the model table
    ct.take<DevModel>(F->nm);
    void* tmp = ct.take<char>((size_t)F->nm * 32);
    std::vector<DevModel> host_models;
    if ((st = upload_models(ctx, scene, F->d_models, tmp, host_models)) != PM_OK) return st;

    RaysParams P{};
    P.tris = scene->d_tris9;
    P.nrms = scene->d_normals9;
    P.models = F->d_models;
    P.n_models = F->nm;
    P.bvh = bvh_nodes(scene);
    P.def_ior = scene->default_refr_index;
    std::memcpy(&P.cam, camera, sizeof(Cam));
    P.max_depth = R.max_depth;
    P.hmax = F->hmax;
    P.hp = F->hp;
    P.hn = F->hn;
    P.hc = F->hc;
    P.hm = F->hm;
    P.count = F->count;
    P.npix = F->npix;
    const uint32_t tiles = ((F->W + 7) / 8) * ((F->H + 7) / 8);
    hipLaunchKernelGGL(k_render_rays, dim3(pm::div_up(tiles, 4)), dim3(256), 0, ctx->stream, P);
    PM_CHECK_LAUNCH(ctx, "k_render_rays");
    if ((st = pm::scan_u32(ctx, F->count, F->off, F->npix, false, F->total)) != PM_OK) return st;
    uint64_t nq64 = 0;
    PM_HIP(ctx, hipMemcpyAsync(&nq64, F->total, 8, hipMemcpyDeviceToHost, ctx->stream));
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t nq = (uint32_t)nq64;
    F->nq = nq;

    // query block + radiance planes (+ the reference-caustic index/distance rows)
    const size_t a = rays_bytes(F);
    const size_t kc = quirk_mode(R) ? (size_t)R.k_caustic : 0;
    const size_t qbytes = pm::carve_size({(size_t)nq * 12, (size_t)nq * 12, (size_t)nq * 12, (size_t)nq * 12,
                                          (size_t)nq * 4, (size_t)nq * 4 * kc, (size_t)nq * 4 * kc});
    if ((st = frame_grow(ctx, F, a + qbytes, a)) != PM_OK) return st;
    pm::Carve c2{static_cast<char*>(F->mem)};
    carve_rays(F, c2);
    F->q3 = c2.take<float>((size_t)nq * 3);
    F->n3 = c2.take<float>((size_t)nq * 3);
    F->lg = c2.take<float>((size_t)nq * 3);
    F->lc = c2.take<float>((size_t)nq * 3);
    F->ccount = c2.take<uint32_t>(nq);
    F->kidx = c2.take<uint32_t>((size_t)nq * kc);
    F->kdist = c2.take<float>((size_t)nq * kc);
    if (nq > 0) {
        hipLaunchKernelGGL(k_render_compact, dim3(pm::div_up(F->npix, 256)), dim3(256), 0, ctx->stream, F->hp, F->hn,
                           F->count, F->off, F->npix, F->hmax, nq, F->q3, F->n3);
        PM_CHECK_LAUNCH(ctx, "k_render_compact");
    }
    F->caustic = false;
    F->rays_done = true;
    if (n_queries) *n_queries = nq;
    return PM_OK;
}

// Phase 2 description: the frame's global and caustic gathers as batch jobs
// (queries and normals of the frame, radiance into the frame's planes).
extern "C" pm_status pm_frame_gather_jobs(pm_frame* F, const pm_grid* global_grid, const pm_grid* caustic_grid,
                                          pm_gather_job* jobs, uint32_t* n_jobs) {
    if (!F || !global_grid || !jobs || !n_jobs) return PM_ERR_INVALID;
    if (!F->rays_done || quirk_mode(F->R)) return PM_ERR_INVALID;
    pm_gather_job g{};
    g.grid = global_grid;
    g.d_queries3 = F->q3;
    g.nq = F->nq;
    g.radius = F->R.r_global;
    g.k = F->R.k_global;
    g.d_normals3 = F->n3;
    g.d_rgb3 = F->lg;
    g.flags = F->R.gather_flags;
    jobs[0] = g;
    *n_jobs = 1;
    F->caustic = caustic_grid != nullptr;
    if (caustic_grid) {
        pm_gather_job c = g;
        c.grid = caustic_grid;
        c.radius = F->R.r_caustic;
        c.k = F->R.k_caustic;
        c.d_rgb3 = F->lc;
        c.d_count = F->ccount;
        jobs[1] = c;
        *n_jobs = 2;
    }
    return PM_OK;
}

// Phase 3: fold every pixel's hits with the estimates, HDR, uchar3.
extern "C" pm_status pm_frame_finish(pm_ctx* ctx, pm_frame* F, float* d_rgb3, uint8_t* d_rgb8,
                                     pm_frame_stats* stats) {
    if (!ctx || !F) return PM_ERR_INVALID;
    if (!F->rays_done) return pm::fail(ctx, PM_ERR_INVALID, "pm_frame_finish: no rays traced");
    hipLaunchKernelGGL(k_render_combine, dim3(pm::div_up(F->npix, 256)), dim3(256), 0, ctx->stream, F->d_models,
                       F->hc, F->hm, F->count, F->off, F->npix, F->hmax, F->nq, F->lg, F->lc, F->ccount,
                       F->R.k_caustic, F->caustic ? 1 : 0, d_rgb3, d_rgb8);
    PM_CHECK_LAUNCH(ctx, "k_render_combine");
    if (stats) {
        stats->hits = F->nq;
        stats->pixels = F->npix;
    }
    return PM_OK;
}

extern "C" pm_status pm_frame_copy_queries(pm_ctx* ctx, const pm_frame* F, float* d_queries3, float* d_normals3,
                                           uint32_t cap, uint32_t* nq) {
    if (!ctx || !F || !nq) return PM_ERR_INVALID;
    if (!F->rays_done) return pm::fail(ctx, PM_ERR_INVALID, "pm_frame_copy_queries: no rays traced");
    *nq = F->nq;
    const uint32_t m = std::min(F->nq, cap);
    if (m == 0) return PM_OK;
    if (!d_queries3 || !d_normals3) return pm::fail(ctx, PM_ERR_INVALID, "pm_frame_copy_queries: null buffer");
    for (int p = 0; p < 3; ++p) {
        PM_HIP(ctx, hipMemcpyAsync(d_queries3 + (size_t)p * m, F->q3 + (size_t)p * F->nq, (size_t)m * 4,
                                   hipMemcpyDeviceToDevice, ctx->stream));
        PM_HIP(ctx, hipMemcpyAsync(d_normals3 + (size_t)p * m, F->n3 + (size_t)p * F->nq, (size_t)m * 4,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    }
    return PM_OK;
}

extern "C" pm_status pm_render_frame(pm_ctx* ctx, const pm_scene* scene, const pm_grid* global_grid,
                                     const pm_grid* caustic_grid, const pm_camera* camera,
                                     const pm_render_params* params, float* d_rgb3, uint8_t* d_rgb8,
                                     pm_frame_stats* stats) {
    if (!ctx || !scene || !camera || !params) return PM_ERR_INVALID;
    if (!global_grid) return pm::fail(ctx, PM_ERR_INVALID, "pm_render_frame: global grid required");
    if (!ctx->frame_state && pm_frame_create(ctx, &ctx->frame_state) != PM_OK) return PM_ERR_NOMEM;
    pm_frame* F = ctx->frame_state;
    pm_status st = pm_frame_rays(ctx, F, scene, camera, params, nullptr);
    if (st != PM_OK) return st;
    const pm_render_params& R = *params;
    const uint32_t gflags = R.gather_flags | PM_KNN_SORT_QUERIES;
    const uint32_t nq = F->nq;
    F->caustic = caustic_grid != nullptr;
    if (nq > 0 && !quirk_mode(R) && R.k_global <= 256 && (!caustic_grid || R.k_caustic <= 256)) {
        // both gathers in one cell-ordered launch (one tail instead of two)
        pm_gather_job jobs[2];
        uint32_t nj = 0;
        if ((st = pm_frame_gather_jobs(F, global_grid, caustic_grid, jobs, &nj)) != PM_OK) return st;
        if ((st = pm_gather_batch(ctx, jobs, nj, gflags)) != PM_OK) return st;
    } else if (nq > 0) {
        if ((st = pm_grid_gather_radiance(ctx, global_grid, F->q3, F->n3, nq, R.r_global, R.k_global, gflags, F->lg,
                                          nullptr)) != PM_OK)
            return st;
        if (caustic_grid) {
            if (!quirk_mode(R)) {
                if ((st = pm_grid_gather_radiance(ctx, caustic_grid, F->q3, F->n3, nq, R.r_caustic, R.k_caustic,
                                                  gflags, F->lc, F->ccount)) != PM_OK)
                    return st;
            } else {
                // Raytracer.cuh:203-204: caustic indices, global map arrays
                if ((st = pm_grid_knn(ctx, caustic_grid, F->q3, nq, R.r_caustic, R.k_caustic, gflags, F->kidx,
                                      F->kdist, F->ccount)) != PM_OK)
                    return st;
                if ((st = pm_radiance(ctx, global_grid, F->kidx, F->kdist, F->ccount, F->n3, nq, R.k_caustic, F->lc)) !=
                    PM_OK)
                    return st;
            }
        }
    }
    return pm_frame_finish(ctx, F, d_rgb3, d_rgb8, stats);
}

// ---------------------------------------------------------------------------
// Planned frames: pm_render_frame's phases enqueued without reading the hit
// count back. The ray trees depend only on the scene and the camera, so a
// frame's hit count is the same every frame: the query block is sized by the
// planning frame's count (`hits`), the count stays on the device, and the
// gathers skip positions past it. Same kernels, same frame.
namespace pm {

struct FramePlan {
    void* mem = nullptr;
    uint32_t W = 0, H = 0, npix = 0, cap = 0;
    int hmax = 0, nm = 0;
    pm_render_params R{};
    RaysParams P{};
    DevModel* d_models = nullptr;
    float *hc = nullptr, *q3 = nullptr, *n3 = nullptr, *lg = nullptr, *lc = nullptr;
    int* hm = nullptr;
    uint32_t *count = nullptr, *off = nullptr, *ccount = nullptr;
    uint64_t* total = nullptr;
};

pm_status frame_plan_create(pm_ctx* ctx, const pm_scene* scene, const pm_camera* camera,
                            const pm_render_params* params, uint64_t hits, FramePlan** out) {
    *out = nullptr;
    if (!scene->d_tris9 || !scene->d_normals9 || !scene->h_models || scene->n_models <= 0 || scene->n_models > 0xFFFF)
        return fail(ctx, PM_ERR_INVALID, "frame plan: empty scene");
    const pm_render_params& R = *params;
    if (R.k_global < 1 || R.k_global > 2048 || R.k_caustic < 1 || R.k_caustic > 2048)
        return fail(ctx, PM_ERR_UNSUPPORTED, "frame plan: k outside [1, 2048]");
    if (R.max_depth < 1 || R.max_depth > kMaxDepth)
        return fail(ctx, PM_ERR_UNSUPPORTED, "frame plan: max_depth outside [1, %d]", kMaxDepth);
    if (quirk_mode(R)) return fail(ctx, PM_ERR_UNSUPPORTED, "frame plan: the reference-caustic mode is eager only");
    if (camera->width <= 0 || camera->height <= 0) return fail(ctx, PM_ERR_INVALID, "frame plan: empty canvas");
    const uint64_t npix64 = (uint64_t)camera->width * (uint64_t)camera->height;
    const uint64_t cap64 = npix64 * (uint64_t)((1 << R.max_depth) - 1);
    if (npix64 > (1ull << 26) || cap64 > (1ull << 30)) return fail(ctx, PM_ERR_UNSUPPORTED, "frame plan: canvas too large");
    FramePlan* F = new FramePlan();
    F->W = (uint32_t)camera->width;
    F->H = (uint32_t)camera->height;
    F->npix = (uint32_t)npix64;
    F->hmax = (1 << R.max_depth) - 1;
    F->cap = (uint32_t)std::min<uint64_t>(cap64, std::max<uint64_t>(hits, 1));
    F->nm = scene->n_models;
    F->R = R;
    const size_t cap = F->cap, npix = F->npix, slots = (size_t)cap64;
    const size_t bytes = carve_size({(size_t)F->nm * sizeof(DevModel), (size_t)F->nm * 32, slots * 12, slots * 12,
                                     slots * 12, slots * 4, npix * 4, npix * 4, 64, cap * 12, cap * 12, cap * 12,
                                     cap * 12, cap * 4});
    hipError_t e = hipMalloc(&F->mem, bytes);
    if (e != hipSuccess) {
        delete F;
        return hip_fail(ctx, e, "hipMalloc(frame plan)");
    }
    Carve cv{static_cast<char*>(F->mem)};
    F->d_models = cv.take<DevModel>(F->nm);
    void* tmp = cv.take<char>((size_t)F->nm * 32);
    float* hp = cv.take<float>(slots * 3);
    float* hn = cv.take<float>(slots * 3);
    F->hc = cv.take<float>(slots * 3);
    F->hm = cv.take<int>(slots);
    F->count = cv.take<uint32_t>(npix);
    F->off = cv.take<uint32_t>(npix);
    F->total = cv.take<uint64_t>(1);
    F->q3 = cv.take<float>(cap * 3);
    F->n3 = cv.take<float>(cap * 3);
    F->lg = cv.take<float>(cap * 3);
    F->lc = cv.take<float>(cap * 3);
    F->ccount = cv.take<uint32_t>(cap);
    std::vector<DevModel> host_models;
    pm_status st = upload_models(ctx, scene, F->d_models, tmp, host_models);
    if (st == PM_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = fail(ctx, PM_ERR_HIP, "frame plan: upload");
    if (st != PM_OK) {
        frame_plan_destroy(F);
        return st;
    }
    RaysParams& P = F->P;
    P.tris = scene->d_tris9;
    P.nrms = scene->d_normals9;
    P.models = F->d_models;
    P.n_models = F->nm;
    P.bvh = bvh_nodes(scene);
    P.def_ior = scene->default_refr_index;
    std::memcpy(&P.cam, camera, sizeof(Cam));
    P.max_depth = R.max_depth;
    P.hmax = F->hmax;
    P.hp = hp;
    P.hn = hn;
    P.hc = F->hc;
    P.hm = F->hm;
    P.count = F->count;
    P.npix = F->npix;
    *out = F;
    return PM_OK;
}

void frame_plan_destroy(FramePlan* F) {
    if (!F) return;
    if (F->mem) (void)hipFree(F->mem);
    delete F;
}

const uint64_t* frame_plan_hits(const FramePlan* F) { return F->total; }
uint32_t frame_plan_capacity(const FramePlan* F) { return F->cap; }

pm_status frame_plan_enqueue(pm_ctx* ctx, FramePlan* F, const pm_grid* global_grid, const pm_grid* caustic_grid,
                             float* d_rgb3, uint8_t* d_rgb8) {
    const pm_render_params& R = F->R;
    const uint32_t tiles = ((F->W + 7) / 8) * ((F->H + 7) / 8);
    hipLaunchKernelGGL(k_render_rays, dim3(div_up(tiles, 4)), dim3(256), 0, ctx->stream, F->P);
    PM_CHECK_LAUNCH(ctx, "k_render_rays (plan)");
    pm_status st = scan_u32(ctx, F->count, F->off, F->npix, false, F->total);
    if (st != PM_OK) return st;
    hipLaunchKernelGGL(k_render_compact, dim3(div_up(F->npix, 256)), dim3(256), 0, ctx->stream, F->P.hp, F->P.hn,
                       F->count, F->off, F->npix, F->hmax, F->cap, F->q3, F->n3);
    PM_CHECK_LAUNCH(ctx, "k_render_compact (plan)");
    const uint32_t* nq = reinterpret_cast<const uint32_t*>(F->total);  // (the low word: hits < 2^32)
    const uint32_t gflags = R.gather_flags | PM_KNN_SORT_QUERIES;
    if ((st = gather_radiance_counted(ctx, global_grid, F->q3, F->n3, F->cap, nq, R.r_global, R.k_global, gflags, F->lg,
                                      nullptr)) != PM_OK)
        return st;
    if (caustic_grid &&
        (st = gather_radiance_counted(ctx, caustic_grid, F->q3, F->n3, F->cap, nq, R.r_caustic, R.k_caustic, gflags,
                                      F->lc, F->ccount)) != PM_OK)
        return st;
    hipLaunchKernelGGL(k_render_combine, dim3(div_up(F->npix, 256)), dim3(256), 0, ctx->stream, F->d_models, F->hc,
                       F->hm, F->count, F->off, F->npix, F->hmax, F->cap, F->lg, F->lc, F->ccount, R.k_caustic,
                       caustic_grid ? 1 : 0, d_rgb3, d_rgb8);
    PM_CHECK_LAUNCH(ctx, "k_render_combine (plan)");
    return PM_OK;
}

}  // namespace pm
