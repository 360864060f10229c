// k nearest photons on the photon kd-tree (cpm::PhotonMap query).
//
// The reference's PhotonMap::locate_q (PhotonMap.cuh:289-371) is a debug
// search (priority queue of nodes, a 10000 sentinel, printf) used only by its
// tests; this is a standard kd search with the same answer as the grid
// gather: the k smallest (d2, photon index) among photons within r.
//
// One wave64 per query. A breadth-first descent from the root gathers up to
// 64 subtree roots whose half-spaces meet the query ball; then every lane
// walks one subtree depth-first with a private stack, visiting one node per
// step. Each step's visited photons feed the wave's LDS key buffer exactly as
// in the grid gather (pm_select.hpp). A subtree is skipped only when the
// squared distance from the query to its splitting plane exceeds the radius
// threshold by a relative 1e-5, so no photon within r is ever missed.
#include <cmath>

#include <cstring>

#include "pm_select.hpp"

namespace {

using namespace pms;

constexpr int kWaves = 4;
constexpr int kCap = 512;
constexpr int kStack = 40;

struct KdParams {
    const float* px;
    const float* py;
    const float* pz;
    const uint32_t* node_idx;
    const int8_t* node_plane;
    uint32_t cap;
    const float* qx;
    const float* qy;
    const float* qz;
    uint32_t nq;
    float T;
    float prune;  // plane-distance^2 bound: T * (1 + 1e-5)
    uint32_t bT;
    int k;
    uint32_t* out_idx;
    float* out_dist;
    uint32_t* out_count;
};

constexpr int kLdsStack = 12;  // top of each lane's depth-first stack, lane-interleaved in LDS

struct KdLds {
    WaveLds<kCap> sel;
    uint32_t frontier[2][64];
    uint32_t stk[kLdsStack][64];
};

__device__ __forceinline__ float kd_dist2(float px, float py, float pz, float qx, float qy, float qz) {
    float dx = px - qx, dy = py - qy, dz = pz - qz;
    float xx = __fmul_rn(dx, dx), yy = __fmul_rn(dy, dy), zz = __fmul_rn(dz, dz);
    return __fadd_rn(__fadd_rn(xx, yy), zz);
}

// Visit node: its photon's key and which children may hold photons within r.
__device__ __forceinline__ void visit(const KdParams& P, uint32_t node, float q0, float q1, float q2, bool& valid,
                                      uint32_t& pid, float& d2, bool& go_l, bool& go_r) {
    valid = false;
    go_l = go_r = false;
    if (node >= P.cap) return;
    pid = P.node_idx[node];
    if (pid == 0xFFFFFFFFu) return;
    const float x = P.px[node], y = P.py[node], z = P.pz[node];
    d2 = kd_dist2(x, y, z, q0, q1, q2);
    valid = true;
    const int p = P.node_plane[node];
    const float diff = (p == 0 ? q0 : p == 1 ? q1 : q2) - (p == 0 ? x : p == 1 ? y : z);
    const float dd = diff * diff;
    // left subtree: coordinate <= median; right: >= median
    go_l = !(diff > 0.f && dd > P.prune);
    go_r = !(diff < 0.f && dd > P.prune);
    go_l = go_l && 2 * node + 1 < P.cap;
    go_r = go_r && 2 * node + 2 < P.cap;
}

__global__ __launch_bounds__(256) void k_kd_knn(KdParams P) {
    __shared__ KdLds lds[kWaves];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t q = blockIdx.x * kWaves + wave;
    if (q >= P.nq) return;  // wave-uniform
    KdLds& L = lds[wave];
    const float q0 = P.qx[q], q1 = P.qy[q], q2 = P.qz[q];
    const float T = P.T;
    const uint32_t k = (uint32_t)P.k;
    uint32_t cnt = 0, found = 0;
    uint64_t khi = ~0ull;

    auto feed = [&](bool valid, uint32_t pid, float d2) {
        bool in = valid && d2 <= T;
        found += (uint32_t)__popcll(wballot(in));
        uint64_t key = make_key(d2, pid);
        bool acc = in && key < khi;
        uint64_t bal = wballot(acc);
        uint32_t na = (uint32_t)__popcll(bal);
        if (cnt + na > (uint32_t)kCap) {
            wave_sync();
            cnt = buf_select<kCap>(L.sel, lane, cnt, k, kCap / 2, P.bT, khi);
            acc = acc && key < khi;
            bal = wballot(acc);
            na = (uint32_t)__popcll(bal);
        }
        if (acc) L.sel.buf[cnt + mbcnt(bal)] = key;
        cnt += na;
    };

    // Breadth-first until the frontier would exceed 64 subtree roots.
    uint32_t fsize = 1;
    int cur = 0;
    if (lane == 0) L.frontier[0][0] = 0;
    wave_sync();
    while (fsize > 0 && fsize <= 32) {
        uint32_t node = lane < fsize ? L.frontier[cur][lane] : P.cap;
        bool valid, gl, gr;
        uint32_t pid = 0;
        float d2 = 0.f;
        visit(P, node, q0, q1, q2, valid, pid, d2, gl, gr);
        feed(valid, pid, d2);
        uint64_t bl = wballot(gl), br = wballot(gr);
        uint32_t nl = (uint32_t)__popcll(bl);
        wave_sync();
        if (gl) L.frontier[cur ^ 1][mbcnt(bl)] = 2 * node + 1;
        if (gr) L.frontier[cur ^ 1][nl + mbcnt(br)] = 2 * node + 2;
        fsize = nl + (uint32_t)__popcll(br);
        cur ^= 1;
        wave_sync();
    }
    // Depth-first: lane l walks the subtree rooted at frontier[l]. The top
    // kLdsStack entries of its stack live in LDS (one column per lane, no
    // bank conflicts), deeper ones in a per-lane array.
    uint32_t deep[kStack - kLdsStack];
    int sp = 0;
    auto push = [&](uint32_t v) {
        if (sp < kLdsStack)
            L.stk[sp][lane] = v;
        else
            deep[sp - kLdsStack] = v;
        ++sp;
    };
    if (lane < fsize) push(L.frontier[cur][lane]);
    for (;;) {
        if (!__builtin_amdgcn_ballot_w64(sp > 0)) break;
        uint32_t node = P.cap;
        if (sp > 0) {
            --sp;
            node = sp < kLdsStack ? L.stk[sp][lane] : deep[sp - kLdsStack];
        }
        bool valid, gl, gr;
        uint32_t pid = 0;
        float d2 = 0.f;
        visit(P, node, q0, q1, q2, valid, pid, d2, gl, gr);
        feed(valid, pid, d2);
        // right child below the left one, so the left subtree is walked
        // first; the order does not change the result (pruning uses r only)
        if (gr && sp < kStack) push(2 * node + 2);
        if (gl && sp < kStack) push(2 * node + 1);
    }
    wave_sync();

    const uint32_t m = found < k ? found : k;
    cnt = buf_select<kCap>(L.sel, lane, cnt, k, 256u, P.bT, khi);
    wave_sync();
    uint32_t* irow = P.out_idx + (size_t)q * P.k;
    float* drow = P.out_dist + (size_t)q * P.k;
    if (cnt <= 64) {
        uint64_t v[1];
        sort_keys<1, kCap>(L.sel, lane, cnt, v);
        write_rows<1>(v, lane, m, P.k, irow, drow);
    } else if (cnt <= 128) {
        uint64_t v[2];
        sort_keys<2, kCap>(L.sel, lane, cnt, v);
        write_rows<2>(v, lane, m, P.k, irow, drow);
    } else {
        uint64_t v[4];
        sort_keys<4, kCap>(L.sel, lane, cnt, v);
        write_rows<4>(v, lane, m, P.k, irow, drow);
    }
    if (lane == 0) P.out_count[q] = m;
}

}  // namespace

extern "C" pm_status pm_kdtree_knn(pm_ctx* ctx, const pm_kdtree* tree, const float* d_queries3, uint32_t nq,
                                   float radius, int32_t k, uint32_t* d_idx, float* d_dist, uint32_t* d_count) {
    if (!ctx) return PM_ERR_INVALID;
    if (!tree) return pm::fail(ctx, PM_ERR_INVALID, "pm_kdtree_knn: null tree");
    if (!(radius >= 0.f) || !std::isfinite(radius)) return pm::fail(ctx, PM_ERR_INVALID, "pm_kdtree_knn: bad radius");
    if (k < 1 || k > 256) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_kdtree_knn: k=%d outside [1, 256]", k);
    if (tree->levels > kStack) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_kdtree_knn: tree too deep");
    if (nq == 0) return PM_OK;
    if (!d_queries3 || !d_idx || !d_dist || !d_count) return pm::fail(ctx, PM_ERR_INVALID, "pm_kdtree_knn: null buffer");
    KdParams P{};
    P.px = tree->photons;
    P.py = tree->photons + tree->cap;
    P.pz = tree->photons + 2 * (size_t)tree->cap;
    P.node_idx = tree->node_idx;
    P.node_plane = tree->node_plane;
    P.cap = tree->cap;
    P.qx = d_queries3;
    P.qy = d_queries3 + nq;
    P.qz = d_queries3 + 2 * (size_t)nq;
    P.nq = nq;
    P.T = radius_threshold(radius);
    P.prune = P.T * (1.0f + 1e-5f) + 1e-30f;
    uint32_t tb;
    std::memcpy(&tb, &P.T, 4);
    P.bT = tb >> 22;
    P.k = k;
    P.out_idx = d_idx;
    P.out_dist = d_dist;
    P.out_count = d_count;
    hipLaunchKernelGGL(k_kd_knn, dim3(pm::div_up(nq, kWaves)), dim3(256), 0, ctx->stream, P);
    PM_CHECK_LAUNCH(ctx, "k_kd_knn");
    return PM_OK;
}
