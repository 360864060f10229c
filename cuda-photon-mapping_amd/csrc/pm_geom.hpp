// Device geometry shared by photon tracing and the eye-ray renderer: FP32
// vector helpers in the reference's host evaluation order (the library is
// built with -ffp-contract=off, and division / sqrtf are IEEE correctly
// rounded), ray-triangle and ray-box tests, the scene hit, and the per-path
// Philox4x32-10 stream.
#pragma once
#include <cfloat>
#include <vector>

#include "pm_internal.hpp"

namespace pmg {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, float f) { return V3{a.x * f, a.y * f, a.z * f}; }
__device__ __forceinline__ V3 mulv(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// vec3::dot host branch (vec3.cuh:258-274)
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vec3::cross host branch (vec3.cuh:276-300)
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ bool is_null(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
// static vec3::normalize: division by the length (vec3.cuh:226-240)
__device__ __forceinline__ V3 normalize_div(V3 a) {
    if (is_null(a)) return a;
    float l = length(a);
    return V3{a.x / l, a.y / l, a.z / l};
}
// member vec3::normalize: multiply by 1/length (vec3.cuh:211-224, :156-169)
__device__ __forceinline__ V3 normalize_rcp(V3 a) {
    if (is_null(a)) return a;
    float u = 1.0f / length(a);
    return V3{a.x * u, a.y * u, a.z * u};
}
__device__ __forceinline__ float maxc(V3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }

struct Ray {
    V3 o, d;
};

// Per-model record on the device (material of Material.cuh:4-31, the
// triangle range of the model, its AABB and id). bvh is the model's root
// node in the scene's BVH node array, or -1 to test every triangle.
struct DevModel {
    uint32_t t0, t1;
    float kd[3], ks[3], ke[3];
    float ior, opaque, shininess;
    int32_t id, bvh;
    float box[6];
};

}  // namespace pmg

// Scene acceleration structure: per-model AABBs and triangle BVHs, built once
// per geometry and shared by photon tracing and eye-ray rendering.
struct pm_accel {
    uint32_t n_tris = 0;
    int32_t n_models = 0;
    std::vector<uint32_t> ranges;  // [n_models][2] triangle range the BVH was built for
    std::vector<float> boxes;      // [n_models][6] exact model AABB
    std::vector<int32_t> roots;    // [n_models] root node, -1 = no BVH (every triangle tested)
    uint32_t n_nodes = 0;
    float4* nodes = nullptr;  // [n_nodes][4]
};

namespace pmg {

__device__ __forceinline__ V3 load_v(const float* __restrict__ t9, uint32_t t, int k) {
    const float* p = t9 + 9 * (size_t)t + 3 * k;
    return v3(p[0], p[1], p[2]);
}

// AABB::intersects_with_ray (AABB.cuh:103-111, update_tmin_tmax :115-127)
__device__ __forceinline__ bool box_hit(const float* c, const Ray& r) {
    float tmin = 0.0f, tmax = __builtin_inff();
    const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        bool sign = __builtin_signbit(d[i]);
        float bmin = sign ? c[3 + i] : c[i];
        float bmax = sign ? c[i] : c[3 + i];
        float dmin = (bmin - o[i]) / d[i];
        float dmax = (bmax - o[i]) / d[i];
        tmin = fmaxf(dmin, tmin);
        tmax = fminf(dmax, tmax);
    }
    return tmin <= tmax;
}

// traingle_intersection (Model.cuh:69-120)
__device__ __forceinline__ bool tri_hit(const Ray& r, V3 v0, V3 v1, V3 v2, float& t, V3& uvw) {
    V3 N = cross(sub(v1, v0), sub(v2, v0));
    float denom = 1.0f / dot(N, N);
    float nd = dot(N, r.d);
    if (fabsf(nd) < 1e-6f) return false;
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

// ---- per-model triangle BVH ------------------------------------------------
// A node is 4 float4: child 0 box (lo.xyz, hi.xyz), child 1 box, then the two
// child codes as int bits (>= 0: node index; < 0: ~triangle index). Boxes
// are inflated by a margin far above the rounding error of tri_hit, so a
// triangle that tri_hit accepts always lies inside every box on its path.
constexpr int kBvhStack = 64;  // a node's key prefix grows down the tree: depth <= 64
// The top kLdsStack entries of each lane's traversal stack live in LDS
// (lane-interleaved, one 8-byte entry per lane per level, so a wave's
// accesses are conflict-free whatever its lanes' depths); deeper entries,
// rare in practice, fall back to a per-lane array.
constexpr int kLdsStack = 12;
constexpr int kStackLanes = 256;  // block size of every kernel that traces rays
struct BvhStack {
    uint2* lds;  // this lane's column of a [kLdsStack][kStackLanes] block array
};

__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz, V3 o, V3 inv,
                                     float tb, float& tn) {
    const float ax = (lx - o.x) * inv.x, bx = (hx - o.x) * inv.x;
    const float ay = (ly - o.y) * inv.y, by = (hy - o.y) * inv.y;
    const float az = (lz - o.z) * inv.z, bz = (hz - o.z) * inv.z;
    const float tmin = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.f));
    const float tmax = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tb));
    tn = tmin;
    return tmin <= tmax;
}

__device__ __forceinline__ V3 inv_dir(V3 d) {
    // a zero component becomes a tiny one of the same sign: its slab is then
    // all of t when the origin lies inside it and empty otherwise
    return v3(1.0f / (d.x != 0.f ? d.x : copysignf(1e-30f, d.x)), 1.0f / (d.y != 0.f ? d.y : copysignf(1e-30f, d.y)),
              1.0f / (d.z != 0.f ? d.z : copysignf(1e-30f, d.z)));
}

// Nearest triangle of one model through its BVH: the smallest t among the
// triangles tri_hit accepts, the lowest index on ties -- exactly what testing
// every triangle in index order returns. Triangles with t >= limit may be
// skipped (the caller discards them). A box is pruned only when its entry t
// exceeds the bound by a relative 2^-16, far above the slab test's rounding.
__device__ __forceinline__ bool bvh_hit(const float4* __restrict__ nodes, int root, const float* __restrict__ t9,
                                        const Ray& r, V3 inv, float limit, float& mt, uint32_t& ti, V3& uvw,
                                        BvhStack S) {
    int stack[kBvhStack - kLdsStack];
    float stn[kBvhStack - kLdsStack];
    int sp = 0;
    int node = root;
    bool found = false;
    mt = FLT_MAX;
    for (;;) {
        const float bound = fminf(limit, mt);
        const float tb = bound + bound * 0x1p-16f;
        if (node >= 0) {
            const float4* nd = nodes + 4 * (size_t)node;
            const float4 a = nd[0], b = nd[1], c = nd[2], e = nd[3];
            float t0, t1;
            const bool h0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, r.o, inv, tb, t0);
            const bool h1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, r.o, inv, tb, t1);
            const int c0 = __float_as_int(e.x), c1 = __float_as_int(e.y);
            if (h0 && h1) {
                const bool first0 = t0 <= t1;
                const int cn = first0 ? c1 : c0;
                const float ct = first0 ? t1 : t0;
                if (sp < kLdsStack)
                    S.lds[sp * kStackLanes] = make_uint2((uint32_t)cn, __float_as_uint(ct));
                else
                    stack[sp - kLdsStack] = cn, stn[sp - kLdsStack] = ct;
                ++sp;
                node = first0 ? c0 : c1;
                continue;
            }
            if (h0 || h1) {
                node = h0 ? c0 : c1;
                continue;
            }
        } else {
            const uint32_t t = ~(uint32_t)node;
            float pt;
            V3 pu;
            if (tri_hit(r, load_v(t9, t, 0), load_v(t9, t, 1), load_v(t9, t, 2), pt, pu) &&
                (pt < mt || (found && pt == mt && t < ti))) {
                mt = pt;
                ti = t;
                uvw = pu;
                found = true;
            }
        }
        // next deferred subtree that can still hold a closer hit
        const float bnd = fminf(limit, mt);
        const float tcut = bnd + bnd * 0x1p-16f;
        node = 0x7FFFFFFF;
        while (sp > 0) {
            --sp;
            int cn;
            float ct;
            if (sp < kLdsStack) {
                const uint2 e = S.lds[sp * kStackLanes];
                cn = (int)e.x;
                ct = __uint_as_float(e.y);
            } else {
                cn = stack[sp - kLdsStack];
                ct = stn[sp - kLdsStack];
            }
            if (ct <= tcut) {
                node = cn;
                break;
            }
        }
        if (node == 0x7FFFFFFF) break;
    }
    return found;
}

// Nearest hit over all models (find_intersection, PhotonTracing.cu:178-213 /
// Raytracer.cuh:66-93) with Model::intersection (Model.cuh:366-446): per
// model an AABB test, then the model's nearest triangle (the first of equal
// t in index order), normal = normalised barycentric blend of the vertex
// normals; a model's hit replaces the current one if the current t is 0 or
// the new t is smaller. A model with a BVH finds its nearest triangle by
// traversal (same result); the others test every triangle, with triangle
// data read at wave-uniform indices (scalar loads).
__device__ __forceinline__ bool scene_hit(const DevModel* __restrict__ M, int n_models,
                                          const float* __restrict__ t9, const float* __restrict__ n9,
                                          const float4* __restrict__ bvh, const Ray& r, float& t_out, V3& n_out,
                                          int& m_out, BvhStack S) {
    float best = 0.f;
    int bm = -1;
    V3 bn = v3(0, 0, 0);
    V3 inv = v3(0, 0, 0);
    bool have_inv = false;
    for (int m = 0; m < n_models; ++m) {
        const bool in_box = box_hit(M[m].box, r);
        if (!__builtin_amdgcn_ballot_w64(in_box)) continue;
        float mt = FLT_MAX;
        bool found = false;
        V3 uvw = v3(0, 0, 0);
        uint32_t ti = 0;
        const int root = M[m].bvh;
        if (root >= 0) {
            if (!have_inv) {
                inv = inv_dir(r.d);
                have_inv = true;
            }
            // a later model matters only below the current best (unless best is 0)
            const float limit = (bm >= 0 && best != 0.f) ? best : FLT_MAX;
            if (in_box) found = bvh_hit(bvh, root, t9, r, inv, limit, mt, ti, uvw, S);
        } else {
            const uint32_t t0 = M[m].t0, t1 = M[m].t1;
            for (uint32_t t = t0; t < t1; ++t) {
                float pt;
                V3 pu;
                if (in_box && tri_hit(r, load_v(t9, t, 0), load_v(t9, t, 1), load_v(t9, t, 2), pt, pu) && pt < mt) {
                    mt = pt;
                    uvw = pu;
                    ti = t;
                    found = true;
                }
            }
        }
        if (!found) continue;
        V3 n = add(add(mul(load_v(n9, ti, 0), uvw.x), mul(load_v(n9, ti, 1), uvw.y)), mul(load_v(n9, ti, 2), uvw.z));
        n = normalize_rcp(n);
        if (bm < 0 || best == 0.f || mt < best) {
            best = mt;
            bm = m;
            bn = n;
        }
    }
    if (bm < 0) return false;
    t_out = best;
    n_out = bn;
    m_out = bm;
    return true;
}

// Host side: copy the scene's model table to the device as DevModel records
// with their AABBs -- from the scene's acceleration structure when it has
// one, else computed on the device over each model's triangles (the live
// effect of create_kd_tree, KdTree.cu:395-464). d_tmp needs 32 * n_models
// bytes. On return host_out holds the same records.
inline pm_status upload_models(pm_ctx* ctx, const pm_scene* scene, DevModel* d_models, void* d_tmp,
                               std::vector<DevModel>& host_out) {
    const int nm = scene->n_models;
    const pm_accel* acc = scene->accel;
    std::vector<DevModel>& hm = host_out;
    hm.assign(nm, DevModel{});
    if (acc && (acc->n_models != nm || acc->n_tris != scene->n_tris))
        return pm::fail(ctx, PM_ERR_INVALID, "scene: acceleration structure built for another scene");
    for (int i = 0; i < nm; ++i) {
        const pm_model& s = scene->h_models[i];
        if (s.tri_end > scene->n_tris || s.tri_begin > s.tri_end)
            return pm::fail(ctx, PM_ERR_INVALID, "scene: model %d triangle range", i);
        if (acc && (acc->ranges[2 * i] != s.tri_begin || acc->ranges[2 * i + 1] != s.tri_end))
            return pm::fail(ctx, PM_ERR_INVALID, "scene: model %d differs from its acceleration structure", i);
        DevModel& d = hm[i];
        d.t0 = s.tri_begin;
        d.t1 = s.tri_end;
        for (int c = 0; c < 3; ++c) {
            d.kd[c] = s.kd[c];
            d.ks[c] = s.ks[c];
            d.ke[c] = s.ke[c];
        }
        d.ior = s.refr_index;
        d.opaque = s.opaque;
        d.shininess = s.shininess;
        d.id = s.id;
        d.bvh = acc ? acc->roots[i] : -1;
        if (acc)
            for (int c = 0; c < 6; ++c) d.box[c] = acc->boxes[6 * (size_t)i + c];
    }
    if (!acc) {
        uint32_t* d_tb = static_cast<uint32_t*>(d_tmp);
        float* d_box = reinterpret_cast<float*>(static_cast<char*>(d_tmp) + 8 * (size_t)nm);
        std::vector<uint32_t> tb(2 * (size_t)nm);
        for (int i = 0; i < nm; ++i) {
            tb[2 * i] = hm[i].t0;
            tb[2 * i + 1] = hm[i].t1;
        }
        PM_HIP(ctx, hipMemcpyAsync(d_tb, tb.data(), tb.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        pm_status st;
        for (int i = 0; i < nm; ++i)
            if ((st = pm::model_aabb_launch(ctx, scene->d_tris9, d_tb + 2 * i, 1, d_box + 6 * i)) != PM_OK) return st;
        std::vector<float> boxes((size_t)nm * 6);
        PM_HIP(ctx, hipMemcpyAsync(boxes.data(), d_box, boxes.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < nm; ++i)
            for (int c = 0; c < 6; ++c) hm[i].box[c] = boxes[(size_t)i * 6 + c];
    }
    PM_HIP(ctx, hipMemcpyAsync(d_models, hm.data(), nm * sizeof(DevModel), hipMemcpyHostToDevice, ctx->stream));
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PM_OK;
}

inline const float4* bvh_nodes(const pm_scene* scene) { return scene->accel ? scene->accel->nodes : nullptr; }

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11); one stream per
// (seed, stream id, path id), four uniforms per block, u in (0, 1].
struct Stream {
    uint32_t k0, k1, path, sid, blk, pos;
    uint32_t buf[4];
    __device__ __forceinline__ void init(uint64_t seed, uint32_t stream_id, uint32_t path_id) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        path = path_id;
        sid = stream_id;
        blk = 0;
        pos = 4;
    }
    __device__ __forceinline__ void refill() {
        uint32_t c0 = path, c1 = blk, c2 = sid, c3 = 0x7068u;
        uint32_t a = k0, b = k1;
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
            uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
            uint32_t n0 = hi1 ^ c1 ^ a, n2 = hi0 ^ c3 ^ b;
            c0 = n0;
            c1 = lo1;
            c2 = n2;
            c3 = lo0;
            a += 0x9E3779B9u;
            b += 0xBB67AE85u;
        }
        buf[0] = c0;
        buf[1] = c1;
        buf[2] = c2;
        buf[3] = c3;
        ++blk;
        pos = 0;
    }
    // buf is a shift register: the next word is always buf[0], so no lane
    // state is ever indexed dynamically
    __device__ __forceinline__ float next() {
        if (pos == 4) refill();
        const uint32_t x = buf[0];
        buf[0] = buf[1];
        buf[1] = buf[2];
        buf[2] = buf[3];
        ++pos;
        return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
    }
};

}  // namespace pmg
