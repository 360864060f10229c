// k nearest photons on the photon kd-tree (cpm::PhotonMap query).
//
// The reference's PhotonMap::locate_q (PhotonMap.cuh:289-371) is a debug
// search (priority queue of nodes, a 10000 sentinel, printf) used only by its
// tests; this is a standard kd search with the same answer as the grid
// gather: the k smallest (d2, photon index) among photons within r.
//
// Each query first searches the ball of 0.4 r (0.16 of the squared
// threshold): when that ball holds at least k photons, they are the k
// nearest overall (every photon outside it is farther than every photon
// inside), and the query is done having visited a fraction of the nodes.
// Otherwise the count c it found sizes a second ball, of squared radius
// 1.5 (k / c) times the first (the growth of a surface density), searched
// from an empty buffer when it is below 0.7 of the full threshold, and the
// full ball follows if that one also holds fewer than k. The last search is
// always the full one, so the kept set is that of the full search. Within a
// search, a subtree is also skipped once the key buffer's acceptance bound
// (the narrowed bin edge of the selection) lies below its plane distance.
//
// One wave64 per query, walking the tree with one LIFO stack of node ids in
// LDS shared by the wave: each step the 64 lanes pop the top 64 entries,
// visit one node each and push the children whose half-spaces meet the
// query ball, so the lanes stay busy however unevenly the ball splits the
// tree. The visited photons feed the wave's LDS key buffer exactly as in the
// grid gather (pm_select.hpp); the k smallest keys do not depend on the
// visiting order. A subtree is skipped only when the squared distance from
// the query to its splitting plane exceeds the radius threshold by a
// relative 1e-5, so no photon within r is ever missed.
#include <cmath>

#include <cstring>

#include "pm_select.hpp"

namespace {

using namespace pms;

constexpr int kWaves = 4;
constexpr uint32_t kMaxLevels = 40;

struct KdParams {
    // per node {x, y, z, photon index | split plane << 30}; w = 0xFFFFFFFF
    // for an empty slot
    const float4* nodes;
    uint32_t cap;
    const float* qx;
    const float* qy;
    const float* qz;
    uint32_t nq;
    float T;
    float prune;     // plane-distance^2 bound: T * (1 + 1e-5)
    float T_in;  // the first search's threshold, 0.16 T
    uint32_t bT;
    int k;
    uint32_t* out_idx;
    float* out_dist;
    uint32_t* out_count;
};

// The wave's node stack. A step pops at most 64 entries and pushes at most
// two children per popped node, so it grows the stack by at most 64; when
// fewer than 128 slots are free the wave pops one node per step instead (a
// depth-first walk from any state grows the stack by at most the tree's
// depth, at most kMaxLevels = 40 < 64 by the host's check), so the stack
// never overflows.
constexpr uint32_t kStackCap = 1024;

// CAP: the key buffer (256 keys for k <= 64, which narrow to at most 128;
// 512 otherwise): the smaller buffer leaves LDS for more waves per CU.
template <int CAP>
struct KdLds {
    WaveLds<CAP> sel;
    uint32_t stk[kStackCap];
};

__device__ __forceinline__ float kd_dist2(float px, float py, float pz, float qx, float qy, float qz) {
    float dx = px - qx, dy = py - qy, dz = pz - qz;
    float xx = __fmul_rn(dx, dx), yy = __fmul_rn(dy, dy), zz = __fmul_rn(dz, dz);
    return __fadd_rn(__fadd_rn(xx, yy), zz);
}

// Visit node: its photon's key and which children may hold photons within r.
__device__ __forceinline__ void visit(const KdParams& P, float prune, uint32_t node, float q0, float q1, float q2,
                                      bool& valid, uint32_t& pid, float& d2, bool& go_l, bool& go_r) {
    valid = false;
    go_l = go_r = false;
    if (node >= P.cap) return;
    const float4 rec = P.nodes[node];
    const uint32_t w = __float_as_uint(rec.w);
    if (w == 0xFFFFFFFFu) return;
    pid = w & 0x3FFFFFFFu;
    const float x = rec.x, y = rec.y, z = rec.z;
    d2 = kd_dist2(x, y, z, q0, q1, q2);
    valid = true;
    const int p = (int)(w >> 30);
    const float diff = (p == 0 ? q0 : p == 1 ? q1 : q2) - (p == 0 ? x : p == 1 ? y : z);
    const float dd = diff * diff;
    // left subtree: coordinate <= median; right: >= median
    go_l = !(diff > 0.f && dd > prune);
    go_r = !(diff < 0.f && dd > prune);
    go_l = go_l && 2 * node + 1 < P.cap;
    go_r = go_r && 2 * node + 2 < P.cap;
}

template <int kCap>
__global__ __launch_bounds__(256) void k_kd_knn(KdParams P) {
    __shared__ KdLds<kCap> lds[kWaves];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t q = blockIdx.x * kWaves + wave;
    if (q >= P.nq) return;  // wave-uniform
    KdLds<kCap>& L = lds[wave];
    const float q0 = P.qx[q], q1 = P.qy[q], q2 = P.qz[q];
    const uint32_t k = (uint32_t)P.k;
    uint32_t cnt = 0, found = 0;
    uint64_t khi = ~0ull;
    float T = P.T_in;

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

    for (int pass = 0;; ++pass) {
        if (pass > 0) {
            // (wave-uniform) the last ball holds the k nearest, or it was the full one
            if (found >= k || !(T < P.T)) break;
            float next = P.T;
            if (pass == 1) {
                const float est = T * 1.5f * (float)k / (float)(found > 0 ? found : 1u);
                if (est < 0.7f * P.T) next = est;
            }
            T = next;
            cnt = 0;
            found = 0;
            khi = ~0ull;
        }
        const float prune0 = T < P.T ? T * (1.0f + 1e-5f) + 1e-30f : P.prune;
        float prune = prune0;
        uint64_t khi_seen = khi;
        uint32_t top = 1;  // (wave-uniform)
        if (lane == 0) L.stk[0] = 0;
        wave_sync();
        while (top > 0) {
            const uint32_t n = kStackCap - top >= 128 ? (top < 64 ? top : 64u) : 1u;
            uint32_t node = P.cap;
            if (lane < n) node = L.stk[top - n + lane];
            wave_sync();  // every pop is read before the pushes reuse the slots
            top -= n;
            bool valid, gl, gr;
            uint32_t pid = 0;
            float d2 = 0.f;
            visit(P, prune, node, q0, q1, q2, valid, pid, d2, gl, gr);
            feed(valid, pid, d2);
            if (khi != khi_seen) {
                // keys at or above khi are never kept: a photon is only
                // worth visiting when its d2 is at most khi's d2 bits
                khi_seen = khi;
                const float b = __uint_as_float((uint32_t)(khi >> 32)) * (1.0f + 1e-5f);
                prune = b < prune0 ? b : prune0;
            }
            // right children below the left ones: left subtrees are popped first
            const uint64_t br = wballot(gr), bl = wballot(gl);
            const uint32_t nr = (uint32_t)__popcll(br);
            if (gr) L.stk[top + mbcnt(br)] = 2 * node + 2;
            if (gl) L.stk[top + nr + mbcnt(bl)] = 2 * node + 1;
            top = __builtin_amdgcn_readfirstlane(top + nr + (uint32_t)__popcll(bl));
            wave_sync();
        }
        wave_sync();
    }

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
    if (tree->levels > kMaxLevels) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_kdtree_knn: tree too deep");
    if (nq == 0) return PM_OK;
    if (!d_queries3 || !d_idx || !d_dist || !d_count) return pm::fail(ctx, PM_ERR_INVALID, "pm_kdtree_knn: null buffer");
    KdParams P{};
    P.nodes = tree->nodes;
    P.cap = tree->cap;
    P.qx = d_queries3;
    P.qy = d_queries3 + nq;
    P.qz = d_queries3 + 2 * (size_t)nq;
    P.nq = nq;
    P.T = radius_threshold(radius);
    P.prune = P.T * (1.0f + 1e-5f) + 1e-30f;
    P.T_in = P.T * 0.16f;
    uint32_t tb;
    std::memcpy(&tb, &P.T, 4);
    P.bT = tb >> 22;
    P.k = k;
    P.out_idx = d_idx;
    P.out_dist = d_dist;
    P.out_count = d_count;
    if (k <= 64)
        hipLaunchKernelGGL(k_kd_knn<256>, dim3(pm::div_up(nq, kWaves)), dim3(256), 0, ctx->stream, P);
    else
        hipLaunchKernelGGL(k_kd_knn<512>, dim3(pm::div_up(nq, kWaves)), dim3(256), 0, ctx->stream, P);
    PM_CHECK_LAUNCH(ctx, "k_kd_knn");
    return PM_OK;
}
