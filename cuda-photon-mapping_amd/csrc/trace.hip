// Photon emission and bounce tracing: trace_photons / emit / trace
// (PhotonTracing.cu:37-55, 178-422).
//
// One lane traces one light path. Paths are numbered; path p draws from its
// own Philox stream (seed, 0, p), so the photons a path stores do not depend
// on which lane, wave or launch traced it. The global and caustic maps take
// the first n photons of each kind in (path, event) order: a batch of paths
// stages its photons per path, two wave64 scans over the per-path counts give
// each photon its map slot, and a compaction kernel writes the SoA blocks.
// The reference instead appends with atomicAdd until both maps are full
// (PhotonArray.cuh:31-46), which makes the stored set launch-dependent.
//
// The medium stack of PhotonTracing.cu:127-176 (photon_trace_stack,
// PhotonTraceStack.cuh:37-116) is a per-lane array of max_bounces entries;
// its strided global-memory layout is gone.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pm_geom.hpp"

namespace {

using namespace pmg;

constexpr int kMaxBounces = 16;
enum : uint8_t { DIF = 0, SPEC = 1, ABSORB = 2, REFR = 3, NONE = 4 };

struct CaEntry {
    float n1, n2, ca;
};

struct TraceParams {
    const float* tris;
    const float* nrms;
    const DevModel* models;
    int n_models;
    const float4* bvh;
    int light;
    float pp[3];
    float def_ior;
    const CaEntry* ca;
    int n_ca;
    uint64_t seed;
    int max_b;
    uint64_t path_base;
    uint32_t n_paths;
    float* stage;     // [9][stage_depth][n_paths]
    uint8_t* kind;    // [stage_depth][n_paths]: 0 global, 1 caustic
    int stage_depth;  // photon slots per path: max_b, or 1 once the global map is full
    uint32_t* cnt_g;  // [n_paths]
    uint32_t* cnt_c;  // [n_paths]
    int caustic_only;  // the global map is already full: stop a path once it cannot store caustics
    // models whose first event can be specular or refraction (Ks != 0 or
    // opaque < 1), as a bit mask; valid when n_models <= 64
    uint64_t spec_mask;
    int spec_mask_valid;
    uint32_t n_segs;  // caustic phase: emission blocks (queue segments) of this batch
    // Planned traces (a fixed batch schedule enqueued without host
    // synchronisation): the maps' fill counts so far, on the device, decide
    // each batch's phase; null for the host-driven loop.
    const uint64_t* got;
    uint32_t need_g, need_c;
};

// Phase of a batch: 0 global (both maps still filling), 1 caustic (global
// map full), 2 done (both full). Host-driven batches carry theirs in
// caustic_only; planned batches read the device fill counts written by the
// previous batches.
__device__ __forceinline__ int batch_phase(const TraceParams& P) {
    if (!P.got) return P.caustic_only;
    const uint64_t g = P.got[0], c = P.got[1];
    if (g >= P.need_g && c >= P.need_c) return 2;
    return g >= P.need_g ? 1 : 0;
}

// photon_trace_stack (PhotonTraceStack.cuh:37-116) as a register shift
// stack of 16-bit model indices, the top entry in the low bits: an entry's
// refractive index is its model's ior and its id the model's id, and only
// the top two entries are ever read, so no per-lane array is indexed.
struct Medium {
    uint64_t w[4];
    int size;
    __device__ __forceinline__ void clear() {
        w[0] = w[1] = w[2] = w[3] = 0;
        size = 0;
    }
    __device__ __forceinline__ uint32_t top() const { return (uint32_t)w[0] & 0xFFFFu; }
    __device__ __forceinline__ uint32_t second() const { return (uint32_t)(w[0] >> 16) & 0xFFFFu; }
    __device__ __forceinline__ void push(uint32_t m) {
        w[3] = (w[3] << 16) | (w[2] >> 48);
        w[2] = (w[2] << 16) | (w[1] >> 48);
        w[1] = (w[1] << 16) | (w[0] >> 48);
        w[0] = (w[0] << 16) | m;
        ++size;
    }
    __device__ __forceinline__ void pop() {
        w[0] = (w[0] >> 16) | (w[1] << 48);
        w[1] = (w[1] >> 16) | (w[2] << 48);
        w[2] = (w[2] >> 16) | (w[3] << 48);
        w[3] >>= 16;
        --size;
    }
};

__device__ __forceinline__ bool ca_lookup(const TraceParams& P, float n1, float n2, float& ca) {
    for (int i = 0; i < P.n_ca; ++i)
        if (P.ca[i].n1 == n1 && P.ca[i].n2 == n2) {
            ca = P.ca[i].ca;
            return true;
        }
    return false;
}

// Emission (PhotonTracing.cu:345-370) with gpu_get_random_point_with_normal
// (Model.cuh:202-243): roundf triangle pick, barycentrics x = u,
// y = u' (1 - x), origin offset 1e-4 along the (unnormalised) normal.
__device__ __forceinline__ Ray emit(const TraceParams& P, Stream& rng) {
    const DevModel& L = P.models[P.light];
    const uint32_t lt = L.t1 - L.t0;
    int tsel = (int)roundf(0 + rng.next() * (float)((3 * (int)lt - 3) / 3));
    uint32_t t = L.t0 + (uint32_t)tsel;
    float ux = rng.next();
    float uy = 0.f + rng.next() * ((1.f - ux) - 0.f);
    float uz = 1.f - ux - uy;
    V3 point = add(add(mul(load_v(P.tris, t, 0), ux), mul(load_v(P.tris, t, 1), uy)), mul(load_v(P.tris, t, 2), uz));
    V3 lnrm = add(add(mul(load_v(P.nrms, t, 0), ux), mul(load_v(P.nrms, t, 1), uy)), mul(load_v(P.nrms, t, 2), uz));
    Ray ray;
    ray.o = add(point, mul(lnrm, 0.0001f));
    float prob, dt;
    do {
        ray.d.x = -1.f + rng.next() * (1.f - -1.f);
        ray.d.y = -1.f + rng.next() * (1.f - -1.f);
        ray.d.z = -1.f + rng.next() * (1.f - -1.f);
        ray.d = normalize_rcp(ray.d);
        prob = rng.next();
        dt = dot(ray.d, lnrm);
        if (dt < 0) {
            dt = -dt;
            ray.d = mul(ray.d, -1.f);
        }
    } while (prob > dt);
    return ray;
}

// With the global map full, a path whose first ray reaches no specular or
// refractive model's AABB (the test scene_hit gates every model with) first
// hits a diffuse surface or nothing: it can store no caustic photon.
__device__ __forceinline__ bool may_store_caustic(const TraceParams& P, const Ray& ray) {
    bool any = false;
    for (uint64_t msk = P.spec_mask; msk; msk &= msk - 1) any |= box_hit(P.models[__builtin_ctzll(msk)].box, ray);
    return any;
}

// A caustic-phase path that survived emission: its ray and generator state.
struct PathState {
    float o[3], d[3];
    uint32_t i, blk, pos;
    uint32_t buf[4];
};

// The state of one light path between bounces.
struct Live {
    Ray ray;
    Stream rng;
    Medium ms;  // photon_trace_stack: (refractive index, model id) entries
    int diffuse_surfs;
    int b;  // bounces done
    uint32_t i;  // path slot in the batch (staging column)
    uint32_t ng, nc, ev;
    uint8_t last;
    __device__ __forceinline__ void start(uint32_t slot, const Stream& r, const Ray& ry) {
        ray = ry;
        rng = r;
        ms.clear();
        diffuse_surfs = 0;
        b = 0;
        i = slot;
        ng = nc = ev = 0;
        last = NONE;
    }
};

// One bounce of a path (trace / destiny, PhotonTracing.cu:127-326), staging
// its photon at path slot L.i. Returns false once the path has ended; the
// caller then stores the path's photon counts.
__device__ __forceinline__ bool bounce_step(const TraceParams& P, Live& L, BvhStack S) {
    const V3 pp = v3(P.pp[0], P.pp[1], P.pp[2]);
    Ray& ray = L.ray;
    Stream& rng = L.rng;
    Medium& ms = L.ms;
    const uint32_t i = L.i;
    {
        // trace (PhotonTracing.cu:278-326)
        float th;
        V3 n;
        int hm;
        if (!scene_hit(P.models, P.n_models, P.tris, P.nrms, P.bvh, ray, th, n, hm, S)) return false;
        V3 ip = add(mul(ray.d, th), ray.o);
        if (dot(ray.d, n) > 0) n = mul(n, -1.f);
        const DevModel& mo = P.models[hm];
        const float cosNL = dot(neg(ray.d), n);
        const float ior = mo.ior;
        const int mid = mo.id;
        uint8_t dest = 0xFF;
        // refract_destity (PhotonTracing.cu:220-243) with get_refraction_pair
        if (mo.opaque != 1.f) {
            float n1, n2;
            if (ms.size == 0) {
                n1 = P.def_ior;
                n2 = ior;
            } else if (P.models[ms.top()].id == mid) {
                n1 = P.models[ms.top()].ior;
                n2 = ms.size == 1 ? P.def_ior : P.models[ms.second()].ior;
            } else {
                n1 = P.models[ms.top()].ior;
                n2 = ior;
            }
            float ca;
            bool tir = ca_lookup(P, n1, n2, ca) && cosNL <= ca;
            if (!tir) {
                float f = (n1 - n2) / (n1 + n2);
                float f0 = f * f;
                float x = 1.f - cosNL;
                float x2 = x * x;
                float x5 = (x2 * x2) * x;
                float refr_p = 1.f - (f0 + (1.f - f0) * x5);
                float e = rng.next();
                if (!(e > refr_p)) dest = REFR;
            }
        }
        // destiny: Russian roulette with e in [0, 2) (PhotonTracing.cu:245-276)
        if (dest == 0xFF) {
            const V3 kd = v3(mo.kd[0], mo.kd[1], mo.kd[2]), ks = v3(mo.ks[0], mo.ks[1], mo.ks[2]);
            float max_lp = maxc(pp);
            float e = 0.f + rng.next() * (2.f - 0.f);
            float pd = maxc(mulv(kd, pp)) / max_lp;
            if (e <= pd) {
                dest = DIF;
            } else {
                float ps = maxc(mulv(ks, pp)) / max_lp;
                dest = (e <= pd + ps) ? SPEC : ABSORB;
            }
        }
        Ray nr{v3(0, 0, 0), v3(0, 0, 0)};
        bool end = false;
        if (dest == REFR) {
            // get_update_refraction_pair (PhotonTracing.cu:127-151)
            float n1, n2;
            if (ms.size != 0 && P.models[ms.top()].id == mid) {
                n1 = P.models[ms.top()].ior;
                n2 = ms.size == 1 ? P.def_ior : P.models[ms.second()].ior;
                ms.pop();
            } else {
                n1 = ms.size == 0 ? P.def_ior : P.models[ms.top()].ior;
                n2 = ior;
                if (ms.size < P.max_b) ms.push((uint32_t)hm);
            }
            // Ray::refract (Ray.cuh:85-96)
            float eta = n1 / n2;
            float c1 = -dot(ray.d, n);
            float w = eta * c1;
            float c2m = (w - eta) * (w + eta);
            if (c2m < -1.f) {
                end = true;
            } else {
                nr.d = add(mul(ray.d, eta), mul(n, w - sqrtf(1.f + c2m)));
                nr.o = add(ip, mul(neg(n), 0.01f));
            }
        } else {
            bool ks_one = fabsf(mo.ks[0] - 1.f) < 1e-3f && fabsf(mo.ks[1] - 1.f) < 1e-3f && fabsf(mo.ks[2] - 1.f) < 1e-3f;
            if (dest == DIF || (dest == ABSORB && !ks_one)) {
                // PathOperator::response (PathOperator.cuh:28-37)
                const bool caustic = (L.last == SPEC || L.last == REFR) && L.diffuse_surfs == 0;
                if ((caustic || !P.caustic_only) && L.ev < (uint32_t)P.stage_depth) {
                    const float v[9] = {ip.x, ip.y, ip.z, pp.x, pp.y, pp.z, ray.d.x, ray.d.y, ray.d.z};
#pragma unroll
                    for (int c = 0; c < 9; ++c) P.stage[((size_t)c * P.stage_depth + L.ev) * P.n_paths + i] = v[c];
                    P.kind[(size_t)L.ev * P.n_paths + i] = caustic ? 1 : 0;
                    ++L.ev;
                    if (caustic)
                        ++L.nc;
                    else
                        ++L.ng;
                }
            } else if (dest == SPEC) {
                // Ray::reflect (Ray.cuh:64-71)
                float dnd = dot(ray.d, n);
                nr.d = normalize_div(sub(ray.d, mul(mul(n, 2.f), dnd)));
                nr.o = add(ip, mul(n, 0.01f));
            }
            if (dest == DIF) {
                // Ray::reflect_spherical (Ray.cuh:16-56)
                V3 d;
                do {
                    float x1, x2, s1, s2;
                    do {
                        x1 = -1.f + rng.next() * (1.f - -1.f);
                        x2 = -1.f + rng.next() * (1.f - -1.f);
                        s1 = x1 * x1;
                        s2 = x2 * x2;
                    } while (s1 + s2 >= 1);
                    float f = sqrtf(1.f - s1 - s2);
                    d = v3(2.f * x1 * f, 2.f * x2 * f, 1.f - 2.f * (s1 + s2));
                } while (dot(d, n) < 0);
                nr.d = normalize_div(d);
                nr.o = add(ip, mul(nr.d, 0.0001f));
            }
        }
        if (dest == DIF) ++L.diffuse_surfs;
        L.last = dest;
        if (dest == ABSORB || end) return false;
        // after a diffuse event no later photon of this path is caustic
        // (PathOperator.cuh:28-37); with the global map full it stores nothing more
        if (P.caustic_only && L.diffuse_surfs > 0) return false;
        ray = nr;
    }
    return ++L.b < P.max_b;
}

// Caustic phase, step 1: emission only. Most paths end here; each block of
// kEmitBlock paths packs its survivors, in path order, at the front of its
// own queue segment and records how many there are, so the bounce loop
// (k_trace_continue) runs in dense waves without any global atomic.
constexpr int kEmitBlock = 1024;

__global__ __launch_bounds__(kEmitBlock) void k_trace_emit(TraceParams P, PathState* __restrict__ queue,
                                                          uint32_t* __restrict__ seg_len) {
    __shared__ uint32_t s_wave[kEmitBlock / 64];
    const uint32_t i = blockIdx.x * kEmitBlock + threadIdx.x;
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    if (P.got && batch_phase(P) != 1) {
        // a planned batch outside the caustic phase: no survivors here
        if (threadIdx.x == 0) seg_len[blockIdx.x] = 0;
        return;
    }
    bool keep = false;
    Ray ray;
    Stream rng;
    if (i < P.n_paths) {
        rng.init(P.seed, 0u, (uint32_t)(P.path_base + i));
        ray = emit(P, rng);
        P.cnt_c[i] = 0;  // (the global counts are not read in this phase)
        keep = may_store_caustic(P, ray);
    }
    const uint64_t m = __ballot(keep);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kEmitBlock / 64; ++w) {
        const uint32_t c = s_wave[w];
        before += (uint32_t)w < wave ? c : 0u;
        total += c;
    }
    if (threadIdx.x == 0) seg_len[blockIdx.x] = total;
    if (!keep) return;
    const uint32_t slot = blockIdx.x * kEmitBlock + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    PathState st;
    st.o[0] = ray.o.x, st.o[1] = ray.o.y, st.o[2] = ray.o.z;
    st.d[0] = ray.d.x, st.d[1] = ray.d.y, st.d[2] = ray.d.z;
    st.i = i;
    st.blk = rng.blk;
    st.pos = rng.pos;
    for (int c = 0; c < 4; ++c) st.buf[c] = rng.buf[c];
    queue[slot] = st;
}

// Path r of the caustic-phase survivor list (in path order): entry
// r - seg_off[s] of queue segment s, the last segment whose exclusive offset
// is <= r.
__device__ __forceinline__ void queued_path(const TraceParams& P, const PathState* __restrict__ queue,
                                            const uint32_t* __restrict__ seg_off, uint32_t r, Live& L) {
    uint32_t lo = 0, hi = P.n_segs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (seg_off[mid] <= r)
            lo = mid;
        else
            hi = mid - 1;
    }
    const PathState st = queue[lo * kEmitBlock + (r - seg_off[lo])];
    Stream rng;
    rng.init(P.seed, 0u, (uint32_t)(P.path_base + st.i));
    rng.blk = st.blk;
    rng.pos = st.pos;
    for (int c = 0; c < 4; ++c) rng.buf[c] = st.buf[c];
    Ray ray;
    ray.o = v3(st.o[0], st.o[1], st.o[2]);
    ray.d = v3(st.d[0], st.d[1], st.d[2]);
    L.start(st.i, rng, ray);
}

// Persistent path tracing. Block b owns a contiguous range of the batch's
// paths (global phase: path slots; caustic phase: survivor list entries).
// Each lane traces one path a bounce at a time and, when it ends, takes the
// block's next path from an LDS counter, so a wave keeps its lanes busy
// instead of waiting on its longest path. Which lane traces a path does not
// change the photons it stores.
template <bool kQueued>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_trace_pool(TraceParams P, const PathState* __restrict__ queue,
                                                   const uint32_t* __restrict__ seg_off,
                                                   const uint64_t* __restrict__ n_queued) {
    __shared__ uint2 s_stack[kLdsStack * kStackLanes];
    __shared__ uint32_t s_next;
    if (P.got) {
        // planned batch: the queued kernels trace the caustic phase when the
        // scene allows it, this one everything else; nothing once both maps
        // are full
        const int ph = batch_phase(P);
        if (ph == 2 || (!kQueued && ph == 1 && P.spec_mask_valid) || (kQueued && ph != 1)) return;
        P.caustic_only = ph;
    }
    const BvhStack S{s_stack + threadIdx.x};
    const uint32_t total = kQueued ? (uint32_t)*n_queued : P.n_paths;
    const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = min(blockIdx.x * per, total), hi = min(lo + per, total);
    if (threadIdx.x == 0) s_next = lo;
    __syncthreads();
    const uint32_t lane = __lane_id();
    Live L;
    bool alive = false, done = false;
    for (;;) {
        const bool need = !alive && !done;
        const uint64_t m = __ballot(need);
        if (m) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(m);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&s_next, (uint32_t)__popcll(m));
            base = __shfl(base, (int)leader, 64);
            if (need) {
                const uint32_t idx = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (idx < hi) {
                    if (kQueued) {
                        queued_path(P, queue, seg_off, idx, L);
                    } else {
                        Stream rng;
                        rng.init(P.seed, 0u, (uint32_t)(P.path_base + idx));
                        const Ray ray = emit(P, rng);
                        L.start(idx, rng, ray);
                    }
                    alive = true;
                } else {
                    done = true;
                }
            }
        }
        // every lane is now tracing a path or out of work for good
        if (!__ballot(alive)) break;
        if (alive && !bounce_step(P, L, S)) {
            P.cnt_g[L.i] = L.ng;
            P.cnt_c[L.i] = L.nc;
            alive = false;
        }
    }
}

// Write each staged photon to its map slot: base + exclusive-scan offset of
// its path + its rank among the path's photons of the same kind.
__global__ __launch_bounds__(256) void k_trace_compact(const float* __restrict__ stage, const uint8_t* __restrict__ kind,
                                                      const uint32_t* __restrict__ cnt_g,
                                                      const uint32_t* __restrict__ cnt_c,
                                                      const uint32_t* __restrict__ off_g,
                                                      const uint32_t* __restrict__ off_c, uint32_t n_paths, int depth,
                                                      uint64_t base_g, uint64_t base_c, float* __restrict__ out_g,
                                                      uint32_t cap_g, float* __restrict__ out_c, uint32_t cap_c,
                                                      const uint64_t* __restrict__ got) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_paths) return;
    if (got) {
        // planned batch: the bases are the device fill counts; a caustic-phase
        // batch has no global counts
        base_g = got[0];
        base_c = got[1];
        if (base_g >= cap_g && base_c >= cap_c) return;
        if (base_g >= cap_g) cnt_g = nullptr;
    }
    // cnt_g == nullptr: caustic phase, every staged photon is caustic
    const uint32_t ne = (cnt_g ? cnt_g[i] : 0u) + cnt_c[i];
    if (ne == 0) return;
    uint64_t sg = cnt_g ? base_g + off_g[i] : 0, sc = base_c + off_c[i];
    for (uint32_t e = 0; e < ne; ++e) {
        bool c = kind[(size_t)e * n_paths + i] != 0;
        uint64_t s = c ? sc++ : sg++;
        uint32_t cap = c ? cap_c : cap_g;
        float* out = c ? out_c : out_g;
        if (s < cap) {
#pragma unroll
            for (int p = 0; p < 9; ++p) out[(size_t)p * cap + s] = stage[((size_t)p * depth + e) * n_paths + i];
        }
    }
}

// Planned batches: add the batch's totals to the device fill counts (a
// caustic-phase batch stores no global photon).
__global__ void k_trace_advance(uint64_t* __restrict__ got, const uint64_t* __restrict__ tot, uint32_t need_g) {
    if (threadIdx.x != 0) return;
    const bool caustic_phase = got[0] >= need_g;
    got[0] += caustic_phase ? 0ull : tot[0];
    got[1] += tot[1];
}

}  // namespace

namespace {

// Checks of a trace request and the host tables it needs: each model's
// triangle range and ior, and compute_critical_angles (PhotonTracing.cu:71-125):
// cos(asin(n2/n1)) for every ordered pair with n2/n1 < 1, against each other
// and the default.
pm_status trace_tables(pm_ctx* ctx, const pm_scene* scene, uint32_t n_global, int32_t max_bounces, const char* who,
                       std::vector<DevModel>& hm, std::vector<CaEntry>& ca) {
    if (!scene->d_tris9 || !scene->d_normals9 || !scene->h_models || scene->n_models <= 0)
        return pm::fail(ctx, PM_ERR_INVALID, "%s: empty scene", who);
    if (scene->light_model < 0 || scene->light_model >= scene->n_models)
        return pm::fail(ctx, PM_ERR_INVALID, "%s: bad light model", who);
    if (max_bounces < 1 || max_bounces > kMaxBounces)
        return pm::fail(ctx, PM_ERR_UNSUPPORTED, "%s: max_bounces outside [1, %d]", who, kMaxBounces);
    if (scene->n_models > 65536) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "%s: more than 65536 models", who);
    if (n_global == 0) return pm::fail(ctx, PM_ERR_INVALID, "%s: n_global must be >= 1", who);
    const int nm = scene->n_models;
    hm.assign(nm, DevModel{});
    for (int i = 0; i < nm; ++i) {
        const pm_model& s = scene->h_models[i];
        if (s.tri_end > scene->n_tris || s.tri_begin > s.tri_end)
            return pm::fail(ctx, PM_ERR_INVALID, "%s: model %d triangle range", who, i);
        hm[i].t0 = s.tri_begin;
        hm[i].t1 = s.tri_end;
        hm[i].ior = s.refr_index;
    }
    if (hm[scene->light_model].t1 == hm[scene->light_model].t0)
        return pm::fail(ctx, PM_ERR_INVALID, "%s: light model has no triangles", who);
    ca.clear();
    auto add_pair = [&](float e1, float e2) {
        float eta = e2 / e1;
        if (!(eta < 1.f)) return;
        for (auto& c : ca)
            if (c.n1 == e1 && c.n2 == e2) return;
        ca.push_back(CaEntry{e1, e2, std::cos(std::asin(eta))});
    };
    for (int i = 0; i + 1 < nm; ++i)
        for (int j = i + 1; j < nm; ++j) {
            add_pair(hm[i].ior, hm[j].ior);
            add_pair(hm[j].ior, hm[i].ior);
        }
    for (int i = 0; i < nm; ++i) {
        add_pair(scene->default_refr_index, hm[i].ior);
        add_pair(hm[i].ior, scene->default_refr_index);
    }
    return PM_OK;
}

// The scene part of a trace's kernel parameters.
void trace_params_base(const pm_scene* scene, const DevModel* d_models, const CaEntry* d_ca, int n_ca,
                       uint32_t n_global, int B, TraceParams& P) {
    const int nm = scene->n_models;
    P = TraceParams{};
    P.tris = scene->d_tris9;
    P.nrms = scene->d_normals9;
    P.models = d_models;
    P.n_models = nm;
    P.bvh = bvh_nodes(scene);
    P.spec_mask = 0;
    P.spec_mask_valid = nm <= 64 ? 1 : 0;
    for (int i = 0; i < nm && i < 64; ++i) {
        const pm_model& m = scene->h_models[i];
        if (m.opaque != 1.f || m.ks[0] != 0.f || m.ks[1] != 0.f || m.ks[2] != 0.f) P.spec_mask |= 1ull << i;
    }
    P.light = scene->light_model;
    for (int c = 0; c < 3; ++c) P.pp[c] = scene->light_intensity[c] / (float)n_global;
    P.def_ior = scene->default_refr_index;
    P.ca = d_ca;
    P.n_ca = n_ca;
    P.max_b = B;
}

}  // namespace

extern "C" pm_status pm_trace_photons(pm_ctx* ctx, const pm_scene* scene, uint64_t seed, uint32_t n_global,
                                      uint32_t n_caustic, int32_t max_bounces, uint64_t max_paths,
                                      float* d_global9, float* d_caustic9, pm_trace_stats* stats) {
    if (!ctx || !scene) return PM_ERR_INVALID;
    std::vector<DevModel> hm;
    std::vector<CaEntry> ca;
    pm_status st0 = trace_tables(ctx, scene, n_global, max_bounces, "pm_trace_photons", hm, ca);
    if (st0 != PM_OK) return st0;
    if ((n_global && !d_global9) || (n_caustic && !d_caustic9))
        return pm::fail(ctx, PM_ERR_INVALID, "pm_trace_photons: null output map");
    const int nm = scene->n_models;

    // Global phase: batches of up to 4M paths, each staging up to
    // max_bounces photons. Caustic phase (global map full): a path stores at
    // most one photon (its first diffuse event, or an absorption, ends it),
    // so batches of up to 16M paths stage one slot each.
    constexpr uint32_t kGlobalBatch = 4u << 20, kCausticBatch = 16u << 20;
    const int B = max_bounces;
    const uint64_t want = (uint64_t)n_global + n_caustic;
    const uint32_t batch_g = (uint32_t)std::max<uint64_t>(
        std::min<uint64_t>(std::min<uint64_t>(std::max<uint64_t>(want, 65536), kGlobalBatch), max_paths), 1);
    // (16 paths per requested photon bound the queue of small traces; a
    // scene with a lower caustic yield just takes more batches)
    const uint32_t batch_c = (uint32_t)std::max<uint64_t>(
        std::min<uint64_t>(std::min<uint64_t>(kCausticBatch, std::max<uint64_t>(16 * want, 65536)), max_paths), 1);
    const uint32_t npath = std::max(batch_g, batch_c);
    const size_t slots = std::max((size_t)batch_g * B, (size_t)batch_c);
    const uint32_t nseg = pm::div_up(batch_c, kEmitBlock);
    size_t rs_bytes = pm::carve_size({(size_t)nm * sizeof(DevModel), ca.size() * sizeof(CaEntry) + 16,
                                      (size_t)nm * 32, slots * 36, slots, (size_t)npath * 4, (size_t)npath * 4,
                                      (size_t)npath * 4, (size_t)npath * 4, 3 * sizeof(uint64_t),
                                      (size_t)nseg * kEmitBlock * sizeof(PathState), (size_t)nseg * 4,
                                      (size_t)nseg * 4});
    void* work = nullptr;
    pm_status st = pm::scratch_reserve(ctx, rs_bytes, &work);
    if (st != PM_OK) return st;
    pm::Carve cv{static_cast<char*>(work)};
    DevModel* d_models = cv.take<DevModel>(nm);
    CaEntry* d_ca = cv.take<CaEntry>(ca.size() + 1);
    void* d_tmp = cv.take<char>(32 * (size_t)nm);
    float* stage = cv.take<float>(slots * 9);
    uint8_t* kind = cv.take<uint8_t>(slots);
    uint32_t* cnt_g = cv.take<uint32_t>(npath);
    uint32_t* cnt_c = cv.take<uint32_t>(npath);
    uint32_t* off_g = cv.take<uint32_t>(npath);
    uint32_t* off_c = cv.take<uint32_t>(npath);
    uint64_t* tot = cv.take<uint64_t>(3);
    PathState* queue = cv.take<PathState>((size_t)nseg * kEmitBlock);
    uint32_t* seg_len = cv.take<uint32_t>(nseg);
    uint32_t* seg_off = cv.take<uint32_t>(nseg);

    // model records with their AABBs (the live effect of create_kd_tree)
    if ((st = upload_models(ctx, scene, d_models, d_tmp, hm)) != PM_OK) return st;
    if (!ca.empty())
        PM_HIP(ctx, hipMemcpyAsync(d_ca, ca.data(), ca.size() * sizeof(CaEntry), hipMemcpyHostToDevice, ctx->stream));

    TraceParams P;
    trace_params_base(scene, d_models, d_ca, (int)ca.size(), n_global, B, P);
    P.seed = seed;
    P.stage = stage;
    P.kind = kind;
    P.cnt_g = cnt_g;
    P.cnt_c = cnt_c;

    // resident blocks of the persistent tracer: one full round over the CUs
    int per_cu = 0;
    PM_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_pool<false>, 256, 0));
    const uint32_t pool_blocks = (uint32_t)std::max(1, per_cu) * (uint32_t)ctx->cu_count;

    uint64_t paths = 0, got_g = 0, got_c = 0;
    uint32_t batch = batch_g;
    while ((got_g < n_global || got_c < n_caustic) && paths < max_paths) {
        // Once the global map is full, its later photons are dropped anyway,
        // so paths stop at their first diffuse event: the caustic photons
        // (and the order they are stored in) are unchanged.
        P.caustic_only = got_g >= n_global ? 1 : 0;
        P.stage_depth = P.caustic_only ? 1 : B;
        const uint32_t nb = (uint32_t)std::min<uint64_t>(std::min(batch, P.caustic_only ? batch_c : batch_g),
                                                         max_paths - paths);
        P.path_base = paths;
        P.n_paths = nb;
        // In that phase most paths end at emission; the survivors go through
        // a queue so that the bounce loop runs in dense waves.
        const bool queued = P.caustic_only && P.spec_mask_valid;
        if (queued) {
            P.n_segs = pm::div_up(nb, kEmitBlock);
            hipLaunchKernelGGL(k_trace_emit, dim3(P.n_segs), dim3(kEmitBlock), 0, ctx->stream, P, queue, seg_len);
            PM_CHECK_LAUNCH(ctx, "k_trace_emit");
            if ((st = pm::scan_u32(ctx, seg_len, seg_off, P.n_segs, false, tot + 2)) != PM_OK) return st;
            hipLaunchKernelGGL(k_trace_pool<true>, dim3(std::min(pm::div_up(nb, 256), pool_blocks)), dim3(256), 0,
                               ctx->stream, P, queue, seg_off, tot + 2);
            PM_CHECK_LAUNCH(ctx, "k_trace_pool<queued>");
        } else {
            hipLaunchKernelGGL(k_trace_pool<false>, dim3(std::min(pm::div_up(nb, 256), pool_blocks)), dim3(256), 0,
                               ctx->stream, P, nullptr, nullptr, nullptr);
            PM_CHECK_LAUNCH(ctx, "k_trace_pool");
        }
        // (in the caustic phase no path stores a global photon)
        if (!P.caustic_only && (st = pm::scan_u32(ctx, cnt_g, off_g, nb, false, tot)) != PM_OK) return st;
        if ((st = pm::scan_u32(ctx, cnt_c, off_c, nb, false, tot + 1)) != PM_OK) return st;
        hipLaunchKernelGGL(k_trace_compact, dim3(pm::div_up(nb, 256)), dim3(256), 0, ctx->stream, stage, kind,
                           P.caustic_only ? nullptr : cnt_g, cnt_c, off_g, off_c, nb, P.stage_depth, got_g, got_c,
                           d_global9, n_global, d_caustic9, n_caustic, nullptr);
        PM_CHECK_LAUNCH(ctx, "k_trace_compact");
        uint64_t h[2];
        PM_HIP(ctx, hipMemcpyAsync(h, tot, 16, hipMemcpyDeviceToHost, ctx->stream));
        PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        got_g += P.caustic_only ? 0 : h[0];
        got_c += h[1];
        paths += nb;
        // size the next batch for the map that is still short
        double rate_g = (double)got_g / (double)paths, rate_c = (double)got_c / (double)paths;
        double need = 0;
        if (got_g < n_global) need = std::max(need, (n_global - got_g) / std::max(rate_g, 1e-9));
        if (got_c < n_caustic) need = std::max(need, (n_caustic - got_c) / std::max(rate_c, 1e-9));
        // (the yield estimate's relative error is ~1/sqrt(photons so far),
        // well under the 2% margin; a short batch only costs one more)
        const uint32_t cap = got_g >= n_global ? batch_c : batch_g;
        batch = (uint32_t)std::min<double>(std::max<double>(need * 1.02 + 1024, 65536), (double)cap);
    }
    if (stats) {
        stats->paths = paths;
        stats->n_global = (uint32_t)std::min<uint64_t>(got_g, n_global);
        stats->n_caustic = (uint32_t)std::min<uint64_t>(got_c, n_caustic);
    }
    return PM_OK;
}

// ---------------------------------------------------------------------------
// Planned traces: the same photon maps as pm_trace_photons, from a fixed
// schedule of batches that is enqueued without reading anything back. Every
// batch launches both phases' kernels; each reads the maps' fill counts on
// the device and works only in its phase (a batch after both maps are full
// does nothing). The maps are the first n photons of each kind in (path,
// event) order whatever the batch sizes, so they equal pm_trace_photons'
// maps whenever the schedule's paths suffice to fill them.
namespace pm {

struct TracePlan {
    void* mem = nullptr;  // model table, critical-angle table, AABB scratch, fill counts
    TraceParams P{};
    uint64_t* got = nullptr;
    uint64_t paths = 0;
    uint32_t batch = 0, n_batches = 0, pool_blocks = 0;
    uint32_t n_global = 0, n_caustic = 0;
    int B = 0;
};

size_t trace_plan_scratch_bytes(const TracePlan* T) {
    const size_t slots = (size_t)T->batch * T->B;
    const uint32_t nseg = div_up(T->batch, kEmitBlock);
    const size_t per_path = (size_t)T->batch * 4;
    return carve_size({slots * 36, slots, per_path, per_path, per_path, per_path, 3 * sizeof(uint64_t),
                       (size_t)nseg * kEmitBlock * sizeof(PathState), (size_t)nseg * 4, (size_t)nseg * 4});
}

pm_status trace_plan_create(pm_ctx* ctx, const pm_scene* scene, uint32_t n_global, uint32_t n_caustic,
                            int32_t max_bounces, uint64_t paths, TracePlan** out) {
    *out = nullptr;
    std::vector<DevModel> hm;
    std::vector<CaEntry> ca;
    pm_status st = trace_tables(ctx, scene, n_global, max_bounces, "trace plan", hm, ca);
    if (st != PM_OK) return st;
    if (paths == 0) return fail(ctx, PM_ERR_INVALID, "trace plan: no paths");
    const int nm = scene->n_models;
    TracePlan* T = new TracePlan();
    T->n_global = n_global;
    T->n_caustic = n_caustic;
    T->B = max_bounces;
    T->paths = paths;
    T->batch = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(std::max<uint64_t>((uint64_t)n_global + n_caustic, 65536),
                                                               4u << 20),
                                           paths);
    T->n_batches = (uint32_t)((paths + T->batch - 1) / T->batch);
    const size_t bytes = carve_size({(size_t)nm * sizeof(DevModel), ca.size() * sizeof(CaEntry) + 16, (size_t)nm * 32, 16});
    hipError_t e = hipMalloc(&T->mem, bytes);
    if (e != hipSuccess) {
        delete T;
        return hip_fail(ctx, e, "hipMalloc(trace plan)");
    }
    Carve cv{static_cast<char*>(T->mem)};
    DevModel* d_models = cv.take<DevModel>(nm);
    CaEntry* d_ca = cv.take<CaEntry>(ca.size() + 1);
    void* d_tmp = cv.take<char>(32 * (size_t)nm);
    T->got = cv.take<uint64_t>(2);
    if ((st = upload_models(ctx, scene, d_models, d_tmp, hm)) != PM_OK ||
        (!ca.empty() && (st = hipMemcpyAsync(d_ca, ca.data(), ca.size() * sizeof(CaEntry), hipMemcpyHostToDevice,
                                             ctx->stream) == hipSuccess
                                 ? PM_OK
                                 : fail(ctx, PM_ERR_HIP, "trace plan: critical angles")) != PM_OK)) {
        trace_plan_destroy(T);
        return st;
    }
    // (the tables are read by every later frame: the host copies complete here)
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        trace_plan_destroy(T);
        return fail(ctx, PM_ERR_HIP, "trace plan: upload");
    }
    trace_params_base(scene, d_models, d_ca, (int)ca.size(), n_global, max_bounces, T->P);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_pool<false>, 256, 0) != hipSuccess) per_cu = 1;
    T->pool_blocks = (uint32_t)std::max(1, per_cu) * (uint32_t)ctx->cu_count;
    *out = T;
    return PM_OK;
}

void trace_plan_destroy(TracePlan* T) {
    if (!T) return;
    if (T->mem) (void)hipFree(T->mem);
    delete T;
}

const uint64_t* trace_plan_counts(const TracePlan* T) { return T->got; }

pm_status trace_plan_enqueue(pm_ctx* ctx, TracePlan* T, uint64_t seed, float* d_global9, float* d_caustic9) {
    void* work = nullptr;
    pm_status st = scratch_reserve(ctx, trace_plan_scratch_bytes(T), &work);
    if (st != PM_OK) return st;
    const size_t slots = (size_t)T->batch * T->B;
    const uint32_t nseg_max = div_up(T->batch, kEmitBlock);
    Carve cv{static_cast<char*>(work)};
    float* stage = cv.take<float>(slots * 9);
    uint8_t* kind = cv.take<uint8_t>(slots);
    uint32_t* cnt_g = cv.take<uint32_t>(T->batch);
    uint32_t* cnt_c = cv.take<uint32_t>(T->batch);
    uint32_t* off_g = cv.take<uint32_t>(T->batch);
    uint32_t* off_c = cv.take<uint32_t>(T->batch);
    uint64_t* tot = cv.take<uint64_t>(3);
    PathState* queue = cv.take<PathState>((size_t)nseg_max * kEmitBlock);
    uint32_t* seg_len = cv.take<uint32_t>(nseg_max);
    uint32_t* seg_off = cv.take<uint32_t>(nseg_max);
    if ((st = fill_u32(ctx, reinterpret_cast<uint32_t*>(T->got), 0u, 4)) != PM_OK) return st;
    TraceParams P = T->P;
    P.seed = seed;
    P.stage = stage;
    P.kind = kind;
    P.cnt_g = cnt_g;
    P.cnt_c = cnt_c;
    P.stage_depth = T->B;
    P.caustic_only = 0;
    P.got = T->got;
    P.need_g = T->n_global;
    P.need_c = T->n_caustic;
    for (uint32_t b = 0; b < T->n_batches; ++b) {
        const uint64_t base = (uint64_t)b * T->batch;
        const uint32_t nb = (uint32_t)std::min<uint64_t>(T->batch, T->paths - base);
        P.path_base = base;
        P.n_paths = nb;
        P.n_segs = div_up(nb, kEmitBlock);
        hipLaunchKernelGGL(k_trace_pool<false>, dim3(std::min(div_up(nb, 256), T->pool_blocks)), dim3(256), 0,
                           ctx->stream, P, nullptr, nullptr, nullptr);
        PM_CHECK_LAUNCH(ctx, "k_trace_pool (plan)");
        if (P.spec_mask_valid) {
            hipLaunchKernelGGL(k_trace_emit, dim3(P.n_segs), dim3(kEmitBlock), 0, ctx->stream, P, queue, seg_len);
            PM_CHECK_LAUNCH(ctx, "k_trace_emit (plan)");
            if ((st = scan_u32(ctx, seg_len, seg_off, P.n_segs, false, tot + 2)) != PM_OK) return st;
            hipLaunchKernelGGL(k_trace_pool<true>, dim3(std::min(div_up(nb, 256), T->pool_blocks)), dim3(256), 0,
                               ctx->stream, P, queue, seg_off, tot + 2);
            PM_CHECK_LAUNCH(ctx, "k_trace_pool<queued> (plan)");
        }
        if ((st = scan_u32(ctx, cnt_g, off_g, nb, false, tot)) != PM_OK) return st;
        if ((st = scan_u32(ctx, cnt_c, off_c, nb, false, tot + 1)) != PM_OK) return st;
        hipLaunchKernelGGL(k_trace_compact, dim3(div_up(nb, 256)), dim3(256), 0, ctx->stream, stage, kind, cnt_g, cnt_c,
                           off_g, off_c, nb, T->B, 0ull, 0ull, d_global9, T->n_global, d_caustic9, T->n_caustic, T->got);
        PM_CHECK_LAUNCH(ctx, "k_trace_compact (plan)");
        hipLaunchKernelGGL(k_trace_advance, dim3(1), dim3(64), 0, ctx->stream, T->got, tot, T->n_global);
        PM_CHECK_LAUNCH(ctx, "k_trace_advance");
    }
    return PM_OK;
}

}  // namespace pm
