// Scene acceleration: one triangle BVH per model, built on the device.
//
// The reference declares a GPU kd-tree over each model's triangles
// (KdTree.cuh:42-157, KdTree.cu:329-375, after Zhou et al. 2008) but never
// runs its build loop; every ray tests every triangle of every model whose
// AABB it hits (Model.cuh:381-429). This file builds what that loop was for,
// as a linear BVH (binary radix tree over Morton codes, Karras 2012):
//   1. 30-bit Morton code of each triangle's centroid in the model AABB;
//   2. stable radix sort of (code, triangle);
//   3. one thread per internal node finds its key range and split from the
//      common-prefix lengths of sorted neighbours (ties in the code broken
//      by sorted position, so every key is distinct);
//   4. one thread per leaf walks to the root writing child boxes; at each
//      node the second arriving child (a device-scope atomic counter)
//      merges both boxes and goes on.
// Leaf boxes are inflated by 2^-13 * max(1, |coordinate|max of the model),
// about three orders of magnitude above the rounding error of the triangle
// test, so the traversal in pm_geom.hpp never culls a triangle the brute
// force would hit.
#include <algorithm>
#include <cmath>
#include <new>

#include "pm_geom.hpp"

namespace {

constexpr uint32_t kMinTris = 16;  // smaller models are tested triangle by triangle

__device__ __forceinline__ uint32_t morton30(float x, float y, float z) {
    auto q = [](float u) {
        u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
        return (uint32_t)u;
    };
    return pmd::spread3(q(x)) | (pmd::spread3(q(y)) << 1) | (pmd::spread3(q(z)) << 2);
}

__global__ __launch_bounds__(256) void k_bvh_keys(const float* __restrict__ t9, uint32_t t0, uint32_t n, float lx,
                                                  float ly, float lz, float sx, float sy, float sz,
                                                  uint32_t* __restrict__ keys) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* p = t9 + 9 * (size_t)(t0 + i);
    float cx = (p[0] + p[3] + p[6]) * (1.0f / 3.0f);
    float cy = (p[1] + p[4] + p[7]) * (1.0f / 3.0f);
    float cz = (p[2] + p[5] + p[8]) * (1.0f / 3.0f);
    keys[i] = morton30((cx - lx) * sx, (cy - ly) * sy, (cz - lz) * sz);
}

// Common prefix length of sorted keys i and j (-1 outside [0, n)); equal
// keys continue the prefix with the positions.
__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Internal node i of n - 1 (Karras 2012, Algorithm 1). Children: internal
// nodes as global indices base + g, leaves as ~(t0 + sorted triangle).
__global__ __launch_bounds__(256) void k_bvh_nodes(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                   int n, uint32_t t0, int base, float4* __restrict__ nodes,
                                                   int* __restrict__ parent_int, int* __restrict__ parent_leaf) {
    int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        int t = (l + div - 1) / div;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int g = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    int c0, c1;
    if (lo == g) {
        c0 = ~(int)(t0 + vals[g]);
        parent_leaf[g] = i;
    } else {
        c0 = base + g;
        parent_int[g] = i;
    }
    if (hi == g + 1) {
        c1 = ~(int)(t0 + vals[g + 1]);
        parent_leaf[g + 1] = i;
    } else {
        c1 = base + g + 1;
        parent_int[g + 1] = i;
    }
    if (i == 0) parent_int[0] = -1;
    nodes[4 * (size_t)(base + i) + 3] = make_float4(__int_as_float(c0), __int_as_float(c1), 0.f, 0.f);
}

// Boxes pass between threads of different blocks (and XCDs): device-scope
// atomic stores and loads, ordered by the fences around the arrival counter.
__device__ __forceinline__ void put_box(float4* nd, int slot, const float* b) {
    uint32_t* f = reinterpret_cast<uint32_t*>(nd) + 6 * slot;
    for (int c = 0; c < 6; ++c)
        __hip_atomic_store(f + c, __float_as_uint(b[c]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float get_f(const float4* nd, int c) {
    const uint32_t* f = reinterpret_cast<const uint32_t*>(nd) + c;
    return __uint_as_float(__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// One thread per leaf: write its box into the parent; the second child to
// arrive at a node merges the node's two boxes and carries on upwards.
__global__ __launch_bounds__(256) void k_bvh_boxes(const float* __restrict__ t9, const uint32_t* __restrict__ vals,
                                                   int n, uint32_t t0, int base, float margin, float4* nodes,
                                                   const int* __restrict__ parent_int,
                                                   const int* __restrict__ parent_leaf, uint32_t* arrive) {
    int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= n) return;
    const uint32_t tri = t0 + vals[j];
    const float* p = t9 + 9 * (size_t)tri;
    float b[6];
    for (int c = 0; c < 3; ++c) {
        b[c] = fminf(fminf(p[c], p[3 + c]), p[6 + c]) - margin;
        b[3 + c] = fmaxf(fmaxf(p[c], p[3 + c]), p[6 + c]) + margin;
    }
    int node = parent_leaf[j];
    int child = ~(int)tri;
    for (;;) {
        float4* nd = nodes + 4 * (size_t)(base + node);
        const int c0 = __float_as_int(nd[3].x);
        put_box(nd, c0 == child ? 0 : 1, b);
        __threadfence();
        if (atomicAdd(&arrive[node], 1u) == 0u) return;  // the sibling finishes this node
        __threadfence();
        for (int c = 0; c < 3; ++c) {
            b[c] = fminf(get_f(nd, c), get_f(nd, 6 + c));
            b[3 + c] = fmaxf(get_f(nd, 3 + c), get_f(nd, 9 + c));
        }
        const int up = parent_int[node];
        if (up < 0) return;
        child = base + node;
        node = up;
    }
}

}  // namespace

extern "C" pm_status pm_accel_build(pm_ctx* ctx, const pm_scene* scene, pm_accel** out) {
    if (!ctx) return PM_ERR_INVALID;
    if (!out || !scene) return pm::fail(ctx, PM_ERR_INVALID, "pm_accel_build: null argument");
    *out = nullptr;
    if (!scene->d_tris9 || !scene->h_models || scene->n_models <= 0)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_accel_build: empty scene");
    if (scene->n_tris >= 0x7FFFFFFFu) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_accel_build: too many triangles");
    const int nm = scene->n_models;
    std::vector<uint32_t> ranges(2 * (size_t)nm);
    uint32_t max_n = 0;
    uint64_t n_nodes = 0;
    for (int i = 0; i < nm; ++i) {
        const pm_model& s = scene->h_models[i];
        if (s.tri_end > scene->n_tris || s.tri_begin > s.tri_end)
            return pm::fail(ctx, PM_ERR_INVALID, "pm_accel_build: model %d triangle range", i);
        ranges[2 * i] = s.tri_begin;
        ranges[2 * i + 1] = s.tri_end;
        uint32_t c = s.tri_end - s.tri_begin;
        if (c >= kMinTris) {
            n_nodes += c - 1;
            max_n = std::max(max_n, c);
        }
    }
    if (n_nodes >= 0x7FFFFFFFull) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_accel_build: too many nodes");

    const size_t rs = pm::radix_sort_scratch_bytes(max_n);
    const size_t bytes = pm::carve_size({(size_t)nm * 8, (size_t)nm * 24, (size_t)max_n * 4, (size_t)max_n * 4,
                                         (size_t)max_n * 4, (size_t)max_n * 4, (size_t)max_n * 4, (size_t)max_n * 4,
                                         rs});
    void* work = nullptr;
    pm_status st = pm::scratch_reserve(ctx, bytes, &work);
    if (st != PM_OK) return st;
    pm::Carve cv{static_cast<char*>(work)};
    uint32_t* d_tb = cv.take<uint32_t>(2 * (size_t)nm);
    float* d_box = cv.take<float>(6 * (size_t)nm);
    uint32_t* keys = cv.take<uint32_t>(max_n);
    uint32_t* keys_s = cv.take<uint32_t>(max_n);
    uint32_t* vals_s = cv.take<uint32_t>(max_n);
    int* parent_int = cv.take<int>(max_n);
    int* parent_leaf = cv.take<int>(max_n);
    uint32_t* arrive = cv.take<uint32_t>(max_n);
    void* rwork = cv.take<char>(rs);

    pm_accel* a = new (std::nothrow) pm_accel();
    if (!a) return pm::fail(ctx, PM_ERR_NOMEM, "pm_accel_build: host allocation");
    a->n_tris = scene->n_tris;
    a->n_models = nm;
    a->ranges = ranges;
    a->roots.assign(nm, -1);
    a->n_nodes = (uint32_t)n_nodes;
    auto bail = [&](pm_status s) {
        pm_accel_destroy(a);
        return s;
    };
    if (n_nodes) {
        hipError_t e = hipMalloc(&a->nodes, n_nodes * 64);
        if (e != hipSuccess) {
            a->nodes = nullptr;
            return bail(pm::fail(ctx, PM_ERR_NOMEM, "pm_accel_build: %llu nodes", (unsigned long long)n_nodes));
        }
    }
    // exact model AABBs (the reference's create_kd_tree result)
    hipError_t e = hipMemcpyAsync(d_tb, ranges.data(), ranges.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return bail(pm::hip_fail(ctx, e, "pm_accel_build: upload"));
    for (int i = 0; i < nm; ++i)
        if ((st = pm::model_aabb_launch(ctx, scene->d_tris9, d_tb + 2 * i, 1, d_box + 6 * i)) != PM_OK) return bail(st);
    a->boxes.resize(6 * (size_t)nm);
    e = hipMemcpyAsync(a->boxes.data(), d_box, a->boxes.size() * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return bail(pm::hip_fail(ctx, e, "pm_accel_build: model boxes"));

    int base = 0;
    for (int i = 0; i < nm; ++i) {
        const uint32_t t0 = ranges[2 * i], n = ranges[2 * i + 1] - t0;
        if (n < kMinTris) continue;
        const float* bx = &a->boxes[6 * (size_t)i];
        float sc[3], big = 1.0f;
        for (int c = 0; c < 3; ++c) {
            float ext = bx[3 + c] - bx[c];
            sc[c] = ext > 0.f ? 1.0f / ext : 0.f;
            big = std::max(big, std::max(std::fabs(bx[c]), std::fabs(bx[3 + c])));
        }
        const float margin = big * 0x1p-13f;
        hipLaunchKernelGGL(k_bvh_keys, dim3(pm::div_up(n, 256)), dim3(256), 0, ctx->stream, scene->d_tris9, t0, n,
                           bx[0], bx[1], bx[2], sc[0], sc[1], sc[2], keys);
        PM_CHECK_LAUNCH(ctx, "k_bvh_keys");
        if ((st = pm::radix_sort_pairs_auto(ctx, keys, nullptr, keys_s, vals_s, n, 30, rwork)) != PM_OK)
            return bail(st);
        if ((st = pm::fill_u32(ctx, arrive, 0u, n)) != PM_OK) return bail(st);
        hipLaunchKernelGGL(k_bvh_nodes, dim3(pm::div_up(n - 1, 256)), dim3(256), 0, ctx->stream, keys_s, vals_s,
                           (int)n, t0, base, a->nodes, parent_int, parent_leaf);
        PM_CHECK_LAUNCH(ctx, "k_bvh_nodes");
        hipLaunchKernelGGL(k_bvh_boxes, dim3(pm::div_up(n, 256)), dim3(256), 0, ctx->stream, scene->d_tris9, vals_s,
                           (int)n, t0, base, margin, a->nodes, parent_int, parent_leaf, arrive);
        PM_CHECK_LAUNCH(ctx, "k_bvh_boxes");
        a->roots[i] = base;
        base += (int)(n - 1);
    }
    *out = a;
    return PM_OK;
}

extern "C" pm_status pm_accel_destroy(pm_accel* accel) {
    if (!accel) return PM_OK;
    if (accel->nodes) (void)hipFree(accel->nodes);
    delete accel;
    return PM_OK;
}
