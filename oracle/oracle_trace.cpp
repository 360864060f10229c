// CPU oracle: photon emission and bounce tracing (PhotonTracing.cu).
// TEST INFRASTRUCTURE ONLY (see oracle_math.hpp).
//
// Restates emit / trace / destiny / refract_destity / find_intersection
// (PhotonTracing.cu:178-375), Model::intersection and traingle_intersection
// (Model.cuh:69-120, 366-446), AABB::intersects_with_ray (AABB.cuh:103-111),
// Ray::reflect / refract / reflect_spherical (Ray.cuh:16-96), the photon
// medium stack (PhotonTracing.cu:127-176) and PathOperator
// (PathOperator.cuh:7-42), with these definitions where the reference is
// undefined or non-deterministic:
//   * random numbers: one Philox4x32-10 stream per path (orc::Stream);
//   * photon order: the global and caustic maps take the first n photons of
//     each kind in (path index, event) order;
//   * a path ends on a miss (trace() returns no value there) and when a
//     refraction fails (Ray::refract reports total internal reflection);
//   * Fresnel-Schlick powers are evaluated as products (x*x, (x2*x2)*x).
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

enum Path { DIF = 0, SPEC = 1, ABSORB = 2, REFR = 3, NONE = 4 };

struct Mat {
    V3 kd, ks, ke;
    float ior, opaque, shininess;
};

struct Model {
    uint32_t t0, t1;
    Mat m;
    int id;
    float box[6];
};

struct Ray {
    V3 o, d;
};

static inline V3 tri_v(const float* t9, uint32_t t, int k) {
    const float* p = t9 + 9 * (size_t)t + 3 * k;
    return v3(p[0], p[1], p[2]);
}

// AABB::intersects_with_ray (AABB.cuh:103-111), update_tmin_tmax (:115-127).
static bool box_hit(const float* c, const Ray& r) {
    float tmin = 0.0f, tmax = INFINITY;
    const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
    for (int i = 0; i < 3; ++i) {
        bool sign = std::signbit(d[i]);
        float bmin = c[sign * 3 + i];
        float bmax = c[!sign * 3 + i];
        float dmin = (bmin - o[i]) / d[i];
        float dmax = (bmax - o[i]) / d[i];
        tmin = std::fmax(dmin, tmin);
        tmax = std::fmin(dmax, tmax);
    }
    return tmin <= tmax;
}

// traingle_intersection (Model.cuh:69-120).
static bool tri_hit(const Ray& r, V3 v0, V3 v1, V3 v2, float& t, V3& uvw) {
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 N = cross(e1, e2);
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

struct Hit {
    float t;
    V3 n;
    int model;
};

// find_intersection (PhotonTracing.cu:178-213 / Raytracer.cuh:66-93) over
// Model::intersection (Model.cuh:366-446).
static bool scene_hit(const std::vector<Model>& M, const float* t9, const float* n9, const Ray& r, Hit& h) {
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
            if (tri_hit(r, tri_v(t9, t, 0), tri_v(t9, t, 1), tri_v(t9, t, 2), pt, pu) && pt < mt) {
                mt = pt;
                uvw = pu;
                ti = t;
                found = true;
            }
        }
        if (!found) continue;
        V3 n = add(add(mul(tri_v(n9, ti, 0), uvw.x), mul(tri_v(n9, ti, 1), uvw.y)), mul(tri_v(n9, ti, 2), uvw.z));
        n = normalize_rcp(n);
        if (bm < 0 || best == 0.f || mt < best) {
            best = mt;
            bm = m;
            bn = n;
        }
    }
    if (bm < 0) return false;
    h.t = best;
    h.n = bn;
    h.model = bm;
    return true;
}

struct CA {
    float n1, n2, ca;
};

// compute_critical_angles / calc_angle (PhotonTracing.cu:71-125).
static std::vector<CA> critical_angles(const std::vector<Model>& M, float def) {
    std::vector<CA> t;
    auto add_pair = [&](float e1, float e2) {
        float eta = e2 / e1;
        if (!(eta < 1.f)) return;
        for (auto& c : t)
            if (c.n1 == e1 && c.n2 == e2) return;
        t.push_back(CA{e1, e2, std::cos(std::asin(eta))});
    };
    for (size_t i = 0; i + 1 < M.size(); ++i)
        for (size_t j = i + 1; j < M.size(); ++j) {
            add_pair(M[i].m.ior, M[j].m.ior);
            add_pair(M[j].m.ior, M[i].m.ior);
        }
    for (auto& m : M) {
        add_pair(def, m.m.ior);
        add_pair(m.m.ior, def);
    }
    return t;
}

struct MediumStack {
    std::vector<float> n;
    std::vector<int> id;
};

struct Tracer {
    const std::vector<Model>& M;
    const float* t9;
    const float* n9;
    std::vector<CA> cat;
    float def;
    int cap;

    bool ca_lookup(float n1, float n2, float& ca) const {
        for (auto& c : cat)
            if (c.n1 == n1 && c.n2 == n2) {
                ca = c.ca;
                return true;
            }
        return false;
    }
    // get_refraction_pair (PhotonTracing.cu:153-176)
    void pair(const MediumStack& s, float nm, int id, float& a, float& b) const {
        size_t sz = s.n.size();
        if (sz == 0) {
            a = def;
            b = nm;
            return;
        }
        if (s.id[sz - 1] == id) {
            a = s.n[sz - 1];
            b = sz == 1 ? def : s.n[sz - 2];
            return;
        }
        a = s.n[sz - 1];
        b = nm;
    }
    // get_update_refraction_pair (PhotonTracing.cu:127-151)
    void pair_update(MediumStack& s, float nm, int id, float& a, float& b) const {
        size_t sz = s.n.size();
        if (sz != 0 && s.id[sz - 1] == id) {
            a = s.n[sz - 1];
            b = sz == 1 ? def : s.n[sz - 2];
            s.n.pop_back();
            s.id.pop_back();
            return;
        }
        if ((int)sz < cap) {
            s.n.push_back(nm);
            s.id.push_back(id);
        }
        if (sz == 0) {
            a = def;
            b = nm;
        } else {
            a = s.n[sz - 1];
            b = nm;
        }
    }
};

}  // namespace

static bool load_models(const void* models_raw, int32_t n_models, uint32_t n_tris, const float* tris9,
                        std::vector<Model>& M);

extern "C" int orc_set_threads(int32_t n) {
#ifdef _OPENMP
    static const int dflt = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : dflt);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// One Philox4x32-10 block, for the generator's known-answer tests.
extern "C" void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    orc::Philox::block(ctr, key[0], key[1], out);
}

extern "C" int orc_scene_hit(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
                             int32_t n_models, const float* rays6, uint32_t nr, float* out_t, int32_t* out_model,
                             float* out_n3) {
    std::vector<Model> M;
    if (!load_models(models_raw, n_models, n_tris, tris9, M)) return -1;
    for (uint32_t i = 0; i < nr; ++i) {
        Ray r;
        r.o = v3(rays6[i], rays6[nr + i], rays6[2 * (size_t)nr + i]);
        r.d = v3(rays6[3 * (size_t)nr + i], rays6[4 * (size_t)nr + i], rays6[5 * (size_t)nr + i]);
        Hit h;
        if (scene_hit(M, tris9, normals9, r, h)) {
            out_t[i] = h.t;
            out_model[i] = h.model;
            out_n3[i] = h.n.x;
            out_n3[nr + i] = h.n.y;
            out_n3[2 * (size_t)nr + i] = h.n.z;
        } else {
            out_t[i] = -1.f;
            out_model[i] = -1;
            out_n3[i] = out_n3[nr + i] = out_n3[2 * (size_t)nr + i] = 0.f;
        }
    }
    return 0;
}

// pm_model records (include/pm355.h) -> oracle models with their AABBs.
static bool load_models(const void* models_raw, int32_t n_models, uint32_t n_tris, const float* tris9,
                        std::vector<Model>& M) {
    struct Raw {
        uint32_t t0, t1;
        float kd[3], ks[3], ke[3], ior, opaque, shin;
        int32_t id, pad;
        float aabb[6];
    };
    const Raw* raw = static_cast<const Raw*>(models_raw);
    M.assign(n_models, Model{});
    for (int i = 0; i < n_models; ++i) {
        Model& m = M[i];
        m.t0 = raw[i].t0;
        m.t1 = raw[i].t1;
        if (m.t1 > n_tris || m.t0 > m.t1) return false;
        m.m.kd = v3(raw[i].kd[0], raw[i].kd[1], raw[i].kd[2]);
        m.m.ks = v3(raw[i].ks[0], raw[i].ks[1], raw[i].ks[2]);
        m.m.ke = v3(raw[i].ke[0], raw[i].ke[1], raw[i].ke[2]);
        m.m.ior = raw[i].ior;
        m.m.opaque = raw[i].opaque;
        m.m.shininess = raw[i].shin;
        m.id = raw[i].id;
        // create_kd_tree's live effect: the model AABB (KdTree.cu:395-464)
        uint32_t b[2] = {m.t0, m.t1};
        orc_model_aabb(tris9, b, 1, m.box);
    }
    return true;
}

namespace {

struct Stored {
    bool caustic;
    V3 pos, dir;
};

// One light path (emit + trace loop, PhotonTracing.cu:345-374): path p of
// the seed's numbering, photons handed to put(caustic, pos, dir) in event
// order, every bounce to log(event) when log is non-null.
template <class Put>
static void trace_path(const Tracer& T, const std::vector<Model>& M, const float* tris9, const float* normals9,
                       const Model& L, V3 pp, uint64_t seed, uint64_t p, int max_bounces, Put&& put,
                       std::vector<orc_trace_event>* log) {
    const uint32_t lt = L.t1 - L.t0;
    Stream rng;
    rng.init(seed, 0u, (uint32_t)p);
    // emit (PhotonTracing.cu:345-370) with gpu_get_random_point_with_normal (Model.cuh:202-243)
    int ti = (int)std::round(0 + rng.next() * (float)((3 * (int)lt - 3) / 3));
    uint32_t t = L.t0 + (uint32_t)ti;
    float ux = rng.next();
    float uy = 0.f + rng.next() * ((1.f - ux) - 0.f);
    float uz = 1.f - ux - uy;
    V3 point = add(add(mul(tri_v(tris9, t, 0), ux), mul(tri_v(tris9, t, 1), uy)), mul(tri_v(tris9, t, 2), uz));
    V3 lnrm = add(add(mul(tri_v(normals9, t, 0), ux), mul(tri_v(normals9, t, 1), uy)), mul(tri_v(normals9, t, 2), uz));
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
    MediumStack ms;
    int diffuse_surfs = 0;
    int last = NONE;
    for (int b = 0; b < max_bounces; ++b) {
        // trace (PhotonTracing.cu:278-326)
        Hit h;
        if (!scene_hit(M, tris9, normals9, ray, h)) break;
        V3 ip = add(mul(ray.d, h.t), ray.o);
        V3 n = h.n;
        if (dot(ray.d, n) > 0) n = mul(n, -1.f);
        const Model& mo = M[h.model];
        float cosNL = dot(neg(ray.d), n);
        // destiny / refract_destity (PhotonTracing.cu:220-276)
        int dest = -1, flags = 0, stored = 0;
        float q1 = 0.f, q2 = 0.f;
        if (mo.m.opaque != 1.f) {
            float n1, n2;
            T.pair(ms, mo.m.ior, mo.id, n1, n2);
            q1 = n1;
            q2 = n2;
            float ca;
            bool tir = T.ca_lookup(n1, n2, ca) && cosNL <= ca;
            if (tir) flags |= 1;
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
        if (dest < 0) {
            float max_lp = std::fmax(std::fmax(pp.x, pp.y), pp.z);
            float e = 0.f + rng.next() * (2.f - 0.f);
            V3 lpd = mulv(mo.m.kd, pp);
            float pd = std::fmax(std::fmax(lpd.x, lpd.y), lpd.z) / max_lp;
            if (e <= pd) {
                dest = DIF;
            } else {
                V3 lps = mulv(mo.m.ks, pp);
                float ps = std::fmax(std::fmax(lps.x, lps.y), lps.z) / max_lp;
                dest = (e <= pd + ps) ? SPEC : ABSORB;
            }
        }
        Ray nr{v3(0, 0, 0), v3(0, 0, 0)};
        bool end = false;
        if (dest == REFR) {
            float n1, n2;
            T.pair_update(ms, mo.m.ior, mo.id, n1, n2);
            // Ray::refract (Ray.cuh:85-96)
            float eta = n1 / n2;
            float c1 = -dot(ray.d, n);
            float w = eta * c1;
            float c2m = (w - eta) * (w + eta);
            if (c2m < -1.f) {
                end = true;
                flags |= 2;
            } else {
                nr.d = add(mul(ray.d, eta), mul(n, w - std::sqrt(1.f + c2m)));
                nr.o = add(ip, mul(neg(n), 0.01f));
            }
        } else {
            bool ks_one = std::fabs(mo.m.ks.x - 1.f) < 1e-3f && std::fabs(mo.m.ks.y - 1.f) < 1e-3f &&
                          std::fabs(mo.m.ks.z - 1.f) < 1e-3f;
            if (dest == DIF || (dest == ABSORB && !ks_one)) {
                // PathOperator::response (PathOperator.cuh:28-37)
                bool caustic = (last == SPEC || last == REFR) && diffuse_surfs == 0;
                put(caustic, ip, ray.d);
                stored = caustic ? 2 : 1;
            } else if (dest == SPEC) {
                // Ray::reflect (Ray.cuh:64-71)
                float dnd = dot(ray.d, n);
                V3 refl = sub(ray.d, mul(mul(n, 2.f), dnd));
                nr.d = normalize_div(refl);
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
                    float f = std::sqrt(1.f - s1 - s2);
                    d = v3(2.f * x1 * f, 2.f * x2 * f, 1.f - 2.f * (s1 + s2));
                } while (dot(d, n) < 0);
                nr.d = normalize_div(d);
                nr.o = add(ip, mul(nr.d, 0.0001f));
            }
        }
        if (log) {
            orc_trace_event ev;
            ev.path = (uint32_t)p;
            ev.bounce = b;
            ev.model = h.model;
            ev.dest = dest;
            ev.stored = stored;
            ev.flags = flags;
            const V3 vs[6] = {ray.o, ray.d, ip, n, nr.o, nr.d};
            float* dst[6] = {ev.o, ev.d, ev.p, ev.n, ev.next_o, ev.next_d};
            for (int k = 0; k < 6; ++k) {
                dst[k][0] = vs[k].x;
                dst[k][1] = vs[k].y;
                dst[k][2] = vs[k].z;
            }
            ev.n1 = q1;
            ev.n2 = q2;
            log->push_back(ev);
        }
        // PathOperator::inform
        if (dest == DIF) ++diffuse_surfs;
        last = dest;
        if (dest == ABSORB || end) break;
        ray = nr;
    }
}

struct TraceSetup {
    std::vector<Model> M;
    Tracer* T = nullptr;
    V3 pp;
    ~TraceSetup() { delete T; }
};

static int setup_trace(const float* tris9, uint32_t n_tris, const void* models_raw, int32_t n_models,
                       int32_t light_model, const float intensity[3], float default_ior, uint32_t n_global,
                       int32_t max_bounces, TraceSetup& S) {
    if (n_models <= 0 || light_model < 0 || light_model >= n_models || max_bounces < 1 || max_bounces > 16) return -1;
    if (!load_models(models_raw, n_models, n_tris, tris9, S.M)) return -1;
    if (S.M[light_model].t1 == S.M[light_model].t0) return -1;
    S.T = new Tracer{S.M, tris9, nullptr, critical_angles(S.M, default_ior), default_ior, max_bounces};
    S.pp = v3(intensity[0] / (float)n_global, intensity[1] / (float)n_global, intensity[2] / (float)n_global);
    return 0;
}

}  // namespace

// Paths are traced in chunks, in parallel within a chunk (each path owns its
// random stream, so a path's photons do not depend on the schedule), and
// merged in path order: the maps are the first n photons of each kind in
// (path, event) order, exactly as a sequential loop over the paths.
extern "C" int orc_trace_photons(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
                                 int32_t n_models, int32_t light_model, const float intensity[3], float default_ior,
                                 uint64_t seed, uint32_t n_global, uint32_t n_caustic, int32_t max_bounces,
                                 uint64_t max_paths, float* out_global9, float* out_caustic9, uint64_t* out_stats) {
    TraceSetup S;
    if (setup_trace(tris9, n_tris, models_raw, n_models, light_model, intensity, default_ior, n_global, max_bounces,
                    S) != 0)
        return -1;
    const Model& L = S.M[light_model];
    const V3 pp = S.pp;
    uint32_t ng = 0, nc = 0;
    uint64_t p = 0;
    auto store = [&](const Stored& s) {
        float* out = s.caustic ? out_caustic9 : out_global9;
        uint32_t cap = s.caustic ? n_caustic : n_global;
        uint32_t& k = s.caustic ? nc : ng;
        if (k < cap) {
            const float v[9] = {s.pos.x, s.pos.y, s.pos.z, pp.x, pp.y, pp.z, s.dir.x, s.dir.y, s.dir.z};
            for (int c = 0; c < 9; ++c) out[(size_t)c * cap + k] = v[c];
        }
        ++k;
    };
    const int B = max_bounces;
    uint64_t chunk = 2048;
    std::vector<Stored> buf;
    std::vector<uint8_t> cnt;
    while ((ng < n_global || nc < n_caustic) && p < max_paths) {
        const uint64_t c = std::min<uint64_t>(chunk, max_paths - p);
        buf.resize((size_t)c * B);
        cnt.assign((size_t)c, 0);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)c; ++i) {
            Stored* dst = &buf[(size_t)i * B];
            uint8_t& k = cnt[(size_t)i];
            trace_path(*S.T, S.M, tris9, normals9, L, pp, seed, p + (uint64_t)i, B,
                       [&](bool caustic, V3 pos, V3 dir) { dst[k++] = Stored{caustic, pos, dir}; }, nullptr);
        }
        for (uint64_t i = 0; i < c && (ng < n_global || nc < n_caustic); ++i, ++p)
            for (int e = 0; e < cnt[(size_t)i]; ++e) store(buf[(size_t)i * B + e]);
        chunk = std::min<uint64_t>(chunk * 2, 1u << 16);
    }
    if (out_stats) {
        out_stats[0] = p;
        out_stats[1] = ng < n_global ? ng : n_global;
        out_stats[2] = nc < n_caustic ? nc : n_caustic;
    }
    return 0;
}

extern "C" int orc_trace_events(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
                                int32_t n_models, int32_t light_model, const float intensity[3], float default_ior,
                                uint64_t seed, uint32_t n_global, int32_t max_bounces, uint64_t n_paths,
                                orc_trace_event* out, uint64_t cap, uint64_t* n_events) {
    TraceSetup S;
    if (setup_trace(tris9, n_tris, models_raw, n_models, light_model, intensity, default_ior, n_global, max_bounces,
                    S) != 0)
        return -1;
    const Model& L = S.M[light_model];
    const int64_t nth = 64;
    std::vector<std::vector<orc_trace_event>> logs((size_t)nth);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nth; ++s) {
        const uint64_t a = n_paths * (uint64_t)s / nth, b = n_paths * (uint64_t)(s + 1) / nth;
        for (uint64_t p = a; p < b; ++p)
            trace_path(*S.T, S.M, tris9, normals9, L, S.pp, seed, p, max_bounces, [](bool, V3, V3) {}, &logs[(size_t)s]);
    }
    uint64_t n = 0;
    for (auto& l : logs)
        for (auto& e : l) {
            if (n < cap) out[n] = e;
            ++n;
        }
    *n_events = n;
    return n <= cap ? 0 : -4;
}
