// CPU oracle: eye-ray frame (Raytracer::render_cpu / render_trace).
// TEST INFRASTRUCTURE ONLY (see oracle_math.hpp).
//
// Restates Raytracer::render_trace (Raytracer.cuh:140-241), HDR (:95-107),
// render_cpu (:281-299), Camera::generate_ray_direction (Camera.cuh:67-81),
// Ray::reflect / refract (Ray.cuh:64-96) and the refraction plan stack of
// RaytracePlanner (RaytracePlanner.cuh:30-69). Definitions where the
// reference is undefined:
//   * media: the reference's MediumManager (MediumManager.cuh:30-88) reads
//     inner-stack entries it never wrote (uninitialised host memory on the
//     CPU path); refraction here uses the photon tracer's medium stack rule
//     (PhotonTracing.cu:127-151): entering a model pushes (ior, id), leaving
//     it pops, indices = (current medium, next medium), and a deferred ray is
//     "in object" while its stack is non-empty. Each deferred ray carries its
//     own stack (depth <= (max_depth + 1) / 2, Drawer.cu:74);
//   * the caustic estimate reads the caustic map's own power and direction
//     (the reference reads the global map's arrays with caustic indices,
//     Raytracer.cuh:203-204); flag bit 0 restores the reference's reads.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "oracle_math.hpp"

using namespace orc;

namespace {

struct RModel {
    uint32_t t0, t1;
    V3 kd, ks, ke;
    float ior, opaque, shininess;
    int id;
    float box[6];
};

struct RRay {
    V3 o, d;
};

static inline V3 tv(const float* t9, uint32_t t, int k) {
    const float* p = t9 + 9 * (size_t)t + 3 * k;
    return v3(p[0], p[1], p[2]);
}

static bool box_hit(const float* c, const RRay& r) {
    float tmin = 0.0f, tmax = INFINITY;
    const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
    for (int i = 0; i < 3; ++i) {
        bool sign = std::signbit(d[i]);
        float bmin = c[sign * 3 + i], bmax = c[!sign * 3 + i];
        float dmin = (bmin - o[i]) / d[i], dmax = (bmax - o[i]) / d[i];
        tmin = std::fmax(dmin, tmin);
        tmax = std::fmin(dmax, tmax);
    }
    return tmin <= tmax;
}

static bool tri_hit(const RRay& r, V3 v0, V3 v1, V3 v2, float& t, V3& uvw) {
    V3 N = cross(sub(v1, v0), sub(v2, v0));
    float denom = 1.0f / dot(N, N);
    float nd = dot(N, r.d);
    if (std::fabs(nd) < 1e-6f) return false;
    t = (dot(N, v0) - dot(N, r.o)) / nd;
    if (t < 0) return false;
    V3 P = add(r.o, mul(r.d, t));
    float w = dot(N, cross(sub(v1, v0), sub(P, v0)));
    if (w < 0) return false;
    float v = dot(N, cross(sub(v0, v2), sub(P, v2)));
    if (v < 0) return false;
    w *= denom;
    v *= denom;
    float u = 1.f - v - w;
    if (u < 0) return false;
    uvw = v3(u, v, w);
    return true;
}

// Raytracer::find_intersection (Raytracer.cuh:66-93)
static bool find_hit(const std::vector<RModel>& M, const float* t9, const float* n9, const RRay& r, bool reverse,
                     int& mi, V3& nrm, V3& ip) {
    float best = 0.f;
    int bm = -1;
    V3 bn = v3(0, 0, 0);
    for (int m = 0; m < (int)M.size(); ++m) {
        if (!box_hit(M[m].box, r)) continue;
        float mt = FLT_MAX;
        bool found = false;
        V3 uvw = v3(0, 0, 0);
        uint32_t ti = 0;
        for (uint32_t t = M[m].t0; t < M[m].t1; ++t) {
            float pt;
            V3 pu;
            if (tri_hit(r, tv(t9, t, 0), tv(t9, t, 1), tv(t9, t, 2), pt, pu) && pt < mt) {
                mt = pt;
                uvw = pu;
                ti = t;
                found = true;
            }
        }
        if (!found) continue;
        V3 n = normalize_rcp(add(add(mul(tv(n9, ti, 0), uvw.x), mul(tv(n9, ti, 1), uvw.y)), mul(tv(n9, ti, 2), uvw.z)));
        if (bm < 0 || best == 0.f || mt < best) {
            best = mt;
            bm = m;
            bn = n;
        }
    }
    if (bm < 0) return false;
    ip = add(mul(r.d, best), r.o);
    if (reverse && dot(r.d, bn) > 0) bn = mul(bn, -1.f);
    mi = bm;
    nrm = bn;
    return true;
}

struct Grid {
    const uint32_t* counts;
    const uint32_t* starts;
    const float* ph;
    uint32_t n;
    float aabb[6];
    float cell;
    int32_t dims[3];
    int32_t layout;
};

}  // namespace

extern "C" int orc_render(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
                          int32_t n_models, float default_ior, const uint32_t* g_counts, const uint32_t* g_starts,
                          const float* g_photons9, uint32_t g_n, const float* g_aabb, float g_cell,
                          const int32_t* g_dims, int32_t g_layout, const uint32_t* c_counts,
                          const uint32_t* c_starts, const float* c_photons9, uint32_t c_n, const float* c_aabb,
                          float c_cell, const int32_t* c_dims, int32_t c_layout, const void* camera_raw,
                          const int32_t* iparams, const float* fparams, int32_t threads, float* out_rgb3,
                          uint8_t* out_rgb8, uint64_t* out_stats) {
    // iparams: max_depth, k_global, k_caustic, flags; fparams: r_global, r_caustic
    struct Raw {
        uint32_t t0, t1;
        float kd[3], ks[3], ke[3], ior, opaque, shin;
        int32_t id, pad;
        float aabb[6];
    };
    struct Cam {
        float pos[3], fwd[3], right[3], up[3];
        float vw, vh, focal;
        int32_t w, h, pad;
    };
    const Raw* raw = static_cast<const Raw*>(models_raw);
    const Cam& cam = *static_cast<const Cam*>(camera_raw);
    std::vector<RModel> M(n_models);
    for (int i = 0; i < n_models; ++i) {
        RModel& m = M[i];
        m.t0 = raw[i].t0;
        m.t1 = raw[i].t1;
        if (m.t1 > n_tris || m.t0 > m.t1) return -1;
        m.kd = v3(raw[i].kd[0], raw[i].kd[1], raw[i].kd[2]);
        m.ks = v3(raw[i].ks[0], raw[i].ks[1], raw[i].ks[2]);
        m.ke = v3(raw[i].ke[0], raw[i].ke[1], raw[i].ke[2]);
        m.ior = raw[i].ior;
        m.opaque = raw[i].opaque;
        m.shininess = raw[i].shin;
        m.id = raw[i].id;
        uint32_t b[2] = {m.t0, m.t1};
        orc_model_aabb(tris9, b, 1, m.box);
    }
    const int max_depth = iparams[0], kg = iparams[1], kc = iparams[2], flags = iparams[3];
    const float rg = fparams[0], rcs = fparams[1];
    const int max_medium = (max_depth + 1) / 2;
    const int W = cam.w, H = cam.h;
    if (max_depth < 1 || max_depth > 6 || kg < 1 || kc < 1) return -1;
    const bool have_c = c_n > 0 && c_counts;
    uint64_t total_hits = 0;

#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total_hits)
#endif
    for (int64_t j = 0; j < H; ++j) {
        std::vector<uint32_t> gi(std::max(kg, kc)), ci(std::max(kg, kc)), cc(1);
        std::vector<float> gd(std::max(kg, kc)), cd(std::max(kg, kc));
        for (int i = 0; i < W; ++i) {
            // Camera::generate_ray_direction (Camera.cuh:67-81), Ray ctor normalises (Ray.cuh:12-15)
            float u = ((float)i + 0.5f) / W, v = ((float)j + 0.5f) / H;
            V3 pos = v3(cam.pos[0], cam.pos[1], cam.pos[2]);
            V3 fwd = v3(cam.fwd[0], cam.fwd[1], cam.fwd[2]);
            V3 horiz = mul(v3(cam.right[0], cam.right[1], cam.right[2]), cam.vw);
            V3 vert = mul(v3(cam.up[0], cam.up[1], cam.up[2]), cam.vh);
            V3 ll = sub(sub(mul(add(pos, fwd), cam.focal), mul(horiz, 0.5f)), mul(vert, 0.5f));
            V3 dir = sub(add(add(ll, mul(horiz, u)), mul(vert, v)), pos);
            RRay ray{pos, normalize_div(dir)};

            // render_trace (Raytracer.cuh:140-241)
            struct Plan {
                RRay ray;
                int depth;
                V3 coef;
                bool in_obj;
                int ms;
                float mn[8];
                int mid[8];
            };
            std::vector<Plan> plans;
            int new_depth = 0, cur_depth = -1;
            V3 acc = v3(0, 0, 0), coef = v3(1, 1, 1);
            bool in_obj = false;
            int ms = 0;
            float mn[8];
            int mid[8];
            for (;;) {
                if (!(cur_depth < max_depth && cur_depth != new_depth)) {
                    if (plans.empty()) break;
                    Plan p = plans.back();
                    plans.pop_back();
                    ray = p.ray;
                    new_depth = p.depth;
                    coef = p.coef;
                    in_obj = p.in_obj;
                    ms = p.ms;
                    for (int s = 0; s < ms; ++s) {
                        mn[s] = p.mn[s];
                        mid[s] = p.mid[s];
                    }
                }
                cur_depth = new_depth;
                int hm;
                V3 nrm, ip;
                if (!find_hit(M, tris9, normals9, ray, in_obj, hm, nrm, ip)) continue;
                ++total_hits;
                const RModel& mat = M[hm];
                V3 cur = v3(0, 0, 0);
                float q[3] = {ip.x, ip.y, ip.z}, nq[3] = {nrm.x, nrm.y, nrm.z};
                if (!in_obj && !(mat.kd.x == 0.f && mat.kd.y == 0.f && mat.kd.z == 0.f)) {
                    V3 di_op = mul(mat.kd, mat.opaque);
                    uint32_t cnt;
                    orc_grid_knn(g_counts, g_starts, g_photons9, g_n, g_aabb, g_cell, g_dims, g_layout, q, 1, rg, kg, 1,
                                 gi.data(), gd.data(), &cnt, nullptr, nullptr, nullptr);
                    float re[3];
                    orc_radiance(gi.data(), gd.data(), &cnt, 1, kg, g_photons9 + 3 * (size_t)g_n,
                                 g_photons9 + 6 * (size_t)g_n, g_n, nq, re, nullptr);
                    cur = add(cur, mul(mulv(v3(re[0], re[1], re[2]), di_op), 4.f));
                }
                cur = add(cur, mat.ke);
                if (have_c) {
                    uint32_t cnt;
                    orc_grid_knn(c_counts, c_starts, c_photons9, c_n, c_aabb, c_cell, c_dims, c_layout, q, 1, rcs, kc, 1,
                                 ci.data(), cd.data(), &cnt, nullptr, nullptr, nullptr);
                    if ((int)cnt == kc) {
                        const float* src = (flags & 1) ? g_photons9 : c_photons9;
                        uint32_t sn = (flags & 1) ? g_n : c_n;
                        float tc[3];
                        bool ok = true;
                        if (flags & 1)
                            for (uint32_t e = 0; e < cnt; ++e) ok = ok && ci[e] < g_n;
                        if (ok) {
                            orc_radiance(ci.data(), cd.data(), &cnt, 1, kc, src + 3 * (size_t)sn, src + 6 * (size_t)sn, sn,
                                         nq, tc, nullptr);
                            cur = add(cur, mul(v3(tc[0], tc[1], tc[2]), mat.opaque));
                        }
                    }
                }
                cur = v3(std::fmin(cur.x, 1.f), std::fmin(cur.y, 1.f), std::fmin(cur.z, 1.f));
                acc = add(acc, mulv(cur, coef));
                if (mat.opaque < 1.f && cur_depth != max_depth - 1) {
                    // refraction: media rule of PhotonTracing.cu:127-151
                    float n1, n2;
                    int nms = ms;
                    float nmn[8];
                    int nmid[8];
                    for (int s = 0; s < ms; ++s) {
                        nmn[s] = mn[s];
                        nmid[s] = mid[s];
                    }
                    if (ms > 0 && mid[ms - 1] == mat.id) {
                        n1 = mn[ms - 1];
                        n2 = ms >= 2 ? mn[ms - 2] : default_ior;
                        nms = ms - 1;
                    } else {
                        n1 = ms > 0 ? mn[ms - 1] : default_ior;
                        n2 = mat.ior;
                        if (ms < max_medium) {
                            nmn[ms] = mat.ior;
                            nmid[ms] = mat.id;
                            nms = ms + 1;
                        }
                    }
                    // Ray::refract (Ray.cuh:85-96)
                    float eta = n1 / n2;
                    float c1 = -dot(ray.d, nrm);
                    float w = eta * c1;
                    float c2m = (w - eta) * (w + eta);
                    if (!(c2m < -1.f) && (int)plans.size() < max_depth) {
                        Plan p;
                        p.ray.d = add(mul(ray.d, eta), mul(nrm, w - std::sqrt(1.f + c2m)));
                        p.ray.o = add(ip, mul(neg(nrm), 0.01f));
                        p.depth = cur_depth + 1;
                        p.coef = mul(coef, 1.f - mat.opaque);
                        p.in_obj = nms > 0;
                        p.ms = nms;
                        for (int s = 0; s < nms; ++s) {
                            p.mn[s] = nmn[s];
                            p.mid[s] = nmid[s];
                        }
                        plans.push_back(p);
                    }
                }
                if (!in_obj && !(mat.ks.x == 0.f && mat.ks.y == 0.f && mat.ks.z == 0.f) && cur_depth != max_depth - 1) {
                    float dnd = dot(ray.d, nrm);
                    RRay nr;
                    nr.d = normalize_div(sub(ray.d, mul(mul(nrm, 2.f), dnd)));
                    nr.o = add(ip, mul(nrm, 0.01f));
                    float c = std::pow(std::fmax(dot(nr.d, neg(ray.d)), 0.0f), mat.shininess);
                    coef = mulv(coef, mul(mul(mat.ks, mat.opaque), c));
                    ray = nr;
                    new_depth = cur_depth + 1;
                }
            }
            // HDR (Raytracer.cuh:95-107), clamp, uchar3 (:238-239, :292)
            const float t = 1.0f / 0.6f;  // pow(exposure, -1), exposure = 0.6f
            const float rgam = 1.0f / 2.2f;
            float o[3] = {acc.x, acc.y, acc.z};
            size_t id = (size_t)i + (size_t)j * W;
            for (int c = 0; c < 3; ++c) {
                float e = std::exp(-t * o[c]);
                float mapped = std::pow(1.0f - e, rgam);
                mapped = std::fmin(mapped, 1.f);
                if (out_rgb3) out_rgb3[(size_t)c * W * H + id] = mapped;
                if (out_rgb8) out_rgb8[3 * id + c] = (uint8_t)(mapped * 255);
            }
        }
    }
    if (out_stats) out_stats[0] = total_hits;
    return 0;
}
