// Photon kd-tree build (cpm::PhotonMap, PhotonMap.cuh:189-283) and the
// per-model AABB of create_kd_tree (KdTree.cu:395-464).
//
// Photon kd-tree, level by level instead of the reference's sequential LIFO
// of quicksorts:
//   * three presorted orders, one per axis, each a stable radix sort of the
//     coordinate (so ties are ordered by photon index, the total order the
//     tree is defined on); a photon's rank on an axis is its position in
//     that sort, and its record in an axis's order holds its ranks on the two
//     other axes (8 bytes): comparing ranks compares (coordinate, index), and
//     in the split axis's own order the side is the position;
//   * at every level each node's segment [from, to) has its median at
//     mid = from + (to - from)/2 in the order of the node's split axis;
//     the other two orders are stable-partitioned around that median in one
//     pass (a segmented scan with decoupled look-back), so every order stays
//     sorted inside every child segment; subtrees of at most 1024 photons
//     are finished in LDS by one workgroup each;
//   * a child's split plane is the argmax extent of its points, read from
//     the first and last record of its segment in each axis order.
// The implicit heap (children 2i+1, 2i+2) is stored with 2^L - 1 slots so no
// node is dropped (the reference's Tree drops indices >= n, Tree.cuh:10-13).
#include <cfloat>
#include <cmath>

#include "pm_internal.hpp"

namespace {

// a photon in LDS (the subtrees finished by one workgroup)
struct Rec {
    float c[3];
    uint32_t id;
};

// A photon's record in the order of axis a: its ranks on the two other axes,
// other_axis(a, 0) < other_axis(a, 1).
struct Rk {
    uint32_t r[2];
};

__host__ __device__ __forceinline__ int other_axis(int a, int i) {
    return i == 0 ? (a == 0 ? 1 : 0) : (a == 2 ? 1 : 2);
}

// The level phase's arrays: the three orders (A[a][j]: the record at
// position j of axis a's order), each axis's sorted photon indices
// (order[a n + r]: the photon of rank r) and ranks (rank4[4 id + a]), and
// the input planes (coordinates by photon index).
struct Orders {
    const Rk* A[3];
    const uint32_t* order;
    const uint32_t* rank4;
    const float* ph9;
    uint32_t n;
};

// the photon at position j of axis a's order
__device__ __forceinline__ uint32_t id_at(const Orders& O, int a, uint32_t j) {
    return O.order[(size_t)other_axis(a, 0) * O.n + O.A[a][j].r[0]];
}

constexpr uint32_t kFinal = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t orderable(float f) {
    f = f + 0.0f;  // -0 -> +0 so that equal coordinates tie on the index
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_axis_keys(const float* __restrict__ p, uint32_t n, uint32_t* __restrict__ keys) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = orderable(p[i]);
}

// Each photon as one 48-byte record (three float4: pos + power.r, power.gb +
// dir.xy, dir.z): the gathers by sort order and by node read one record
// instead of up to nine scattered 4-byte values.
__global__ __launch_bounds__(256) void k_photon_records(const float* __restrict__ in9, uint32_t n,
                                                       float4* __restrict__ rec) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) v[p] = in9[(size_t)p * n + i];
    rec[3 * (size_t)i + 0] = make_float4(v[0], v[1], v[2], v[3]);
    rec[3 * (size_t)i + 1] = make_float4(v[4], v[5], v[6], v[7]);
    rec[3 * (size_t)i + 2] = make_float4(v[8], 0.f, 0.f, 0.f);
}

// rank4[order[a n + r]].a = r (blockIdx.y = axis a): each photon's three
// ranks in one 16-byte entry, so a record is one gather
__global__ __launch_bounds__(256) void k_ranks(const uint32_t* __restrict__ order, uint32_t n,
                                              uint32_t* __restrict__ rank4) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
    if (r >= n) return;
    rank4[4 * (size_t)order[(size_t)a * n + r] + a] = r;
}

// the records of axis a's order (blockIdx.y = a): ranks on the other axes
__global__ __launch_bounds__(256) void k_rank_recs(const uint32_t* __restrict__ order,
                                                  const uint4* __restrict__ rank4, uint32_t n,
                                                  Rk* __restrict__ out0, Rk* __restrict__ out1,
                                                  Rk* __restrict__ out2) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
    if (j >= n) return;
    const uint4 rk = rank4[order[(size_t)a * n + j]];
    Rk r;
    r.r[0] = a == 0 ? rk.y : rk.x;
    r.r[1] = a == 2 ? rk.y : rk.z;
    (a == 0 ? out0 : a == 1 ? out1 : out2)[j] = r;
}

// PhotonMap::find_largest_plane (PhotonMap.cuh:161-178) from the extreme
// photons of a non-empty segment in each axis order.
__device__ __forceinline__ int8_t plane_of(const Orders& O, uint32_t f, uint32_t t) {
    float d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float* c = O.ph9 + (size_t)a * O.n;
        d[a] = fabsf(c[id_at(O, a, t - 1)] - c[id_at(O, a, f)]);
    }
    int8_t p = 0;
    float mx = d[0];
    if (d[1] > mx) {
        mx = d[1];
        p = 1;
    }
    if (d[2] > mx) p = 2;
    return p;
}

// Per node of a level: its split plane (the argmax extent of its segment,
// PhotonMap.cuh:252-273: an empty node keeps +-FLT_MAX corners -> plane 0)
// and, when `medians`, its median, the median's key and the children's
// segments (med = {the median's rank on the plane, its index}). root_n > 0:
// the level is the root's, whose segment is [0, root_n). lookback_counter
// (nullable): the level partition before this launch has completed, so its
// look-back epoch is advanced here.
__global__ __launch_bounds__(256) void k_level_nodes(Orders O, uint32_t first, uint32_t count,
                                                    uint32_t cap, uint32_t root_n, bool medians, uint2* __restrict__ seg,
                                                    int8_t* __restrict__ plane, uint32_t* __restrict__ node_idx,
                                                    int8_t* __restrict__ node_plane, uint2* __restrict__ med,
                                                    uint32_t* __restrict__ lookback_counter) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && lookback_counter) pm::lookback_advance(lookback_counter, pm::lookback_epoch(lookback_counter));
    if (t >= count) return;
    uint32_t node = first + t;
    if (root_n > 0) seg[node] = make_uint2(0, root_n);
    const uint2 s = root_n > 0 ? make_uint2(0, root_n) : seg[node];
    const int8_t p = s.y > s.x ? plane_of(O, s.x, s.y) : (int8_t)0;
    plane[node] = p;
    if (!medians) return;
    uint32_t l = 2 * node + 1, r = 2 * node + 2;
    if (s.y <= s.x) {  // empty
        if (l < cap) seg[l] = make_uint2(0, 0);
        if (r < cap) seg[r] = make_uint2(0, 0);
        return;
    }
    uint32_t mid = s.x + (s.y - s.x) / 2;
    const uint32_t id = id_at(O, p, mid);
    node_idx[node] = id;
    node_plane[node] = p;
    med[node] = make_uint2(O.rank4[4 * (size_t)id + p], id);
    if (s.y - s.x == 1) {
        if (l < cap) seg[l] = make_uint2(0, 0);
        if (r < cap) seg[r] = make_uint2(0, 0);
    } else {
        if (l < cap) seg[l] = make_uint2(s.x, mid);
        if (r < cap) seg[r] = make_uint2(mid + 1, s.y);
    }
}

// side of record e relative to node's median on plane p: 0 left, 1 median, 2 right
__device__ __forceinline__ int side_of(const Rec& e, int p, uint2 m) {
    if (e.id == m.y) return 1;
    const float a = p == 0 ? e.c[0] : (p == 1 ? e.c[1] : e.c[2]), b = __uint_as_float(m.x);
    bool less = a < b || (a == b && e.id < m.y);
    return less ? 0 : 2;
}

// ---- one level's stable partition in one pass --------------------------------
// The three axis orders are one stream of 3n positions (axis-major). Each
// position carries a segmented-scan value: bit 31 = a segment starts here
// (the node's `from`, or a finished position), bit 30 = the node's median,
// bits 0-29 = goes left. A tile scans its positions in registers / LDS and
// takes its prefix from earlier tiles by decoupled look-back (epoch-tagged
// tile words, as in the u32 scan); within a node's segment the inclusive
// value then counts the left records (Lb) and the median (Mb) up to each
// position, which gives the destination directly:
//   left: from + Lb,   median: mid,   right: mid + 1 + (j - from - Lb - Mb).
// This replaces flags -> device scan -> scatter (three passes over 3n
// positions) by one.
constexpr int kPThreads = 512;                 // (kPItems * kPThreads / 64 wave rows <= 64: one wave scans them)
constexpr int kPItems = 8;                      // rows of kPThreads positions per tile
constexpr uint32_t kPTile = kPThreads * kPItems;  // 4096
static_assert(kPItems * (kPThreads / 64) <= 64, "the wave rows of a tile are scanned by one wave");
constexpr uint32_t kSegH = 1u << 31, kSegM = 1u << 30, kSegL = (1u << 30) - 1;
constexpr uint64_t kStAgg = 1, kStIncl = 2;

// x then y
__device__ __forceinline__ uint32_t seg_op(uint32_t x, uint32_t y) {
    if (y & kSegH) return y;
    return ((x & kSegL) + (y & kSegL)) | ((x | y) & kSegM) | (x & kSegH);
}

__device__ __forceinline__ uint64_t seg_state(uint32_t epoch, uint64_t status, uint32_t value) {
    return ((uint64_t)epoch << 34) | (status << 32) | value;
}

__global__ __launch_bounds__(kPThreads) void k_level_partition(
    const Rk* __restrict__ A0, const Rk* __restrict__ A1, const Rk* __restrict__ A2, Rk* __restrict__ B0,
    Rk* __restrict__ B1, Rk* __restrict__ B2, uint32_t n, const uint32_t* __restrict__ node_in,
    uint32_t* __restrict__ node_out, const uint2* __restrict__ seg, const int8_t* __restrict__ plane,
    const uint2* __restrict__ med, uint64_t* __restrict__ state, uint32_t* __restrict__ counter, uint32_t n_tiles) {
    __shared__ uint32_t s_tile, s_epoch;
    __shared__ uint32_t s_wagg[kPItems * (kPThreads / 64)];
    __shared__ uint32_t s_wpre[kPItems * (kPThreads / 64)];
    __shared__ uint32_t s_prefix;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) {
        // the epoch is advanced by k_level_nodes, the next launch: a tile
        // here may finish its look-back before earlier tiles have read it
        s_epoch = pm::lookback_epoch(counter);
        uint32_t id = atomicAdd(counter, 1u);
        if (id == n_tiles - 1) atomicExch(counter, 0u);  // every tile holds its id: reset for the next launch
        s_tile = id;
    }
    __syncthreads();
    const uint32_t tile = s_tile, epoch = s_epoch;
    const uint32_t total = 3u * n;  // n <= 2^30

    // position of item k: tile row k, wave, lane (rows of kPThreads consecutive
    // positions). All loads of the tile's items are issued before any item
    // is scanned: positions, then their nodes' segments, planes and medians.
    uint32_t val[kPItems], incl[kPItems], node[kPItems];
    uint2 sgk[kPItems];
    Rk rec[kPItems];
#pragma unroll
    for (int k = 0; k < kPItems; ++k) {
        const uint32_t g = tile * kPTile + (uint32_t)k * kPThreads + t;
        node[k] = kFinal;
        if (g < total) {
            const uint32_t a = g >= 2 * n ? 2u : (g >= n ? 1u : 0u), j = g - a * n;
            node[k] = node_in[j];
            rec[k] = (a == 0 ? A0 : a == 1 ? A1 : A2)[j];
        }
    }
    int8_t pl[kPItems];
    uint2 md[kPItems];
#pragma unroll
    for (int k = 0; k < kPItems; ++k) {
        if (node[k] != kFinal) {
            sgk[k] = seg[node[k]];
            pl[k] = plane[node[k]];
            md[k] = med[node[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < kPItems; ++k) {
        const uint32_t g = tile * kPTile + (uint32_t)k * kPThreads + t;
        uint32_t v = kSegH;
        if (node[k] != kFinal) {
            const uint32_t a = g >= 2 * n ? 2u : (g >= n ? 1u : 0u), j = g - a * n;
            // the split axis's own order: the position decides; another
            // order: the record's rank on the split axis against the median's
            int sd;
            if (pl[k] == (int8_t)a) {
                const uint32_t mid = sgk[k].x + (sgk[k].y - sgk[k].x) / 2;
                sd = j < mid ? 0 : (j == mid ? 1 : 2);
            } else {
                const uint32_t rr = rec[k].r[pl[k] == other_axis((int)a, 0) ? 0 : 1];
                sd = rr < md[k].x ? 0 : (rr == md[k].x ? 1 : 2);
            }
            v = (j == sgk[k].x ? kSegH : 0u) | (sd == 0 ? 1u : 0u) | (sd == 1 ? kSegM : 0u);
        }
        val[k] = v;
        // wave-level inclusive segmented scan over the row
        uint32_t x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t y = __shfl_up(x, d, 64);
            if (lane >= (uint32_t)d) x = seg_op(y, x);
        }
        incl[k] = x;
        if (lane == 63) s_wagg[k * (kPThreads / 64) + wave] = x;
    }
    __syncthreads();
    // exclusive prefixes of the 64 wave rows, in tile order (row-major)
    if (wave == 0) {
        constexpr uint32_t nw = kPItems * (kPThreads / 64);
        uint32_t x = lane < nw ? s_wagg[lane] : 0u;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t y = __shfl_up(x, d, 64);
            if (lane >= (uint32_t)d) x = seg_op(y, x);
        }
        // x = inclusive over rows 0..lane; exclusive = inclusive of lane - 1
        uint32_t ex = __shfl_up(x, 1, 64);
        if (lane < nw) s_wpre[lane] = lane == 0 ? 0u : ex;
        const uint32_t agg = __shfl(x, nw - 1, 64);
        // decoupled look-back for the tile prefix
        uint32_t prefix = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&state[0], seg_state(epoch, kStIncl, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&state[tile], seg_state(epoch, kStAgg, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            int64_t jt = (int64_t)tile - 1;
            bool have = false;  // prefix holds the combination of the tiles after the current window
            for (;;) {
                const int64_t idx = jt - (int64_t)lane;
                uint64_t w;
                uint32_t status;
                uint64_t stop_mask, inval_mask, upto;
                for (;;) {
                    if (idx >= 0) {
                        w = __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        status = ((uint32_t)(w >> 34) == epoch) ? (uint32_t)((w >> 32) & 3) : 0u;
                    } else {
                        w = 0;
                        status = (uint32_t)kStIncl;
                    }
                    // a tile whose value holds a segment start ends the walk too
                    stop_mask = __ballot(status == kStIncl || (status == kStAgg && ((uint32_t)w & kSegH)));
                    inval_mask = __ballot(status == 0);
                    const int first = stop_mask ? __builtin_ctzll(stop_mask) : 63;
                    upto = (first == 63) ? ~0ull : ((2ull << first) - 1);
                    if ((inval_mask & upto) == 0) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                // combine the window's values from the farthest used lane down to lane 0
                const int last = (stop_mask & upto) ? __builtin_ctzll(stop_mask) : 63;
                uint32_t win = 0;
                for (int l = last; l >= 0; --l) {
                    const uint32_t wl = (uint32_t)__shfl((uint32_t)w, l, 64);
                    win = (l == last) ? wl : seg_op(win, wl);
                }
                prefix = have ? seg_op(win, prefix) : win;
                have = true;
                if (stop_mask & upto) break;
                jt -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&state[tile], seg_state(epoch, kStIncl, seg_op(prefix, agg)), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    const uint32_t tprefix = s_prefix;
#pragma unroll
    for (int k = 0; k < kPItems; ++k) {
        const uint32_t g = tile * kPTile + (uint32_t)k * kPThreads + t;
        if (g >= total) continue;
        const uint32_t a = g >= 2 * n ? 2u : (g >= n ? 1u : 0u), j = g - a * n;
        const uint32_t s = node[k];
        uint32_t dst = j, nn = kFinal;
        if (s != kFinal) {
            const uint32_t before = seg_op(seg_op(tprefix, s_wpre[k * (kPThreads / 64) + wave]), incl[k]);
            const uint2 sg = sgk[k];
            const uint32_t from = sg.x, mid = sg.x + (sg.y - sg.x) / 2;
            const uint32_t own = val[k];
            const uint32_t lb = (before & kSegL) - (own & kSegL), mb = ((before & kSegM) ? 1u : 0u) -
                                                                       ((own & kSegM) ? 1u : 0u);
            if (own & 1u)
                dst = from + lb;
            else if (own & kSegM)
                dst = mid;
            else
                dst = mid + 1 + ((j - from) - lb - mb);
            if (!(sg.y - sg.x <= 1 || j == mid)) nn = j < mid ? 2 * s + 1 : 2 * s + 2;
        }
        if (dst < n) (a == 0 ? B0 : a == 1 ? B1 : B2)[dst] = rec[k];  // always true for a consistent tree
        if (a == 0) node_out[j] = nn;
    }
}

// ---- bottom levels in LDS ---------------------------------------------------
// Once every segment holds at most kLocal records, one workgroup finishes a
// whole subtree in LDS: the same median / stable-partition / child-plane
// steps as the level loop above, on the subtree's slice of the three axis
// orders, with workgroup scans instead of device scans. The subtree's records
// are loaded once (in axis-0 order); the three orders are permutations of
// 16-bit local indices, so a partition moves 2 bytes per position and a
// workgroup needs 46 KB of LDS (three workgroups per CU).
constexpr int kLocal = 1024;
constexpr int kLocalThreads = 256;

struct LocalLds {
    Rec rec[kLocal];          // the subtree's records in axis-0 order
    // per axis: local record index at each position, in buffers ordb[oi[a]];
    // the fourth buffer receives the next partition and the roles rotate
    uint16_t ordb[4][kLocal];
    uint16_t pref[kLocal];
    uint16_t node[kLocal];    // local node (at the current depth) of each position, 0xFFFF = final
    ushort2 seg[2][kLocal];   // node segments, current and next depth
    int8_t plane[2][kLocal];
    uint2 med[kLocal / 2];
    uint16_t mpos[kLocal / 2][3];
    uint8_t flag[kLocal];
    uint32_t wsum[kLocalThreads / 64];
};

// pos0[id] = position of photon id in the axis-0 order (for the subtrees'
// index maps).
// For the LDS levels: ids0[j] = the photon at position j of the axis-0
// order, and posr0[r] = the position in that order of the photon of axis-0
// rank r (the order-1 and order-2 records carry that rank).
__global__ __launch_bounds__(256) void k_pos0(Orders O, uint32_t* __restrict__ ids0, uint32_t* __restrict__ posr0) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= O.n) return;
    const uint32_t id = id_at(O, 0, j);
    ids0[j] = id;
    posr0[O.rank4[4 * (size_t)id]] = j;
}

// Exclusive scan of L.flag[0, m) (0/1) into L.pref.
__device__ __forceinline__ void local_scan(LocalLds& L, uint32_t m) {
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr uint32_t per = kLocal / kLocalThreads;  // 4 positions per thread
    uint32_t v[per], sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < per; ++i) {
        uint32_t j = t * per + i;
        v[i] = j < m ? L.flag[j] : 0u;
        sum += v[i];
    }
    uint32_t inc = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += y;
    }
    if (lane == 63) L.wsum[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += L.wsum[w];
    uint32_t run = before + inc - sum;
#pragma unroll
    for (uint32_t i = 0; i < per; ++i) {
        uint32_t j = t * per + i;
        if (j < m) L.pref[j] = (uint16_t)run;
        run += v[i];
    }
    __syncthreads();
}

// find_largest_plane over local segment [f, t) (PhotonMap.cuh:161-178)
__device__ __forceinline__ int8_t local_plane(const LocalLds& L, const int (&oi)[3], uint32_t f, uint32_t t) {
    const uint16_t* o0 = L.ordb[oi[0]];
    const uint16_t* o1 = L.ordb[oi[1]];
    const uint16_t* o2 = L.ordb[oi[2]];
    float d0 = fabsf(L.rec[o0[t - 1]].c[0] - L.rec[o0[f]].c[0]);
    float d1 = fabsf(L.rec[o1[t - 1]].c[1] - L.rec[o1[f]].c[1]);
    float d2 = fabsf(L.rec[o2[t - 1]].c[2] - L.rec[o2[f]].c[2]);
    int8_t p = 0;
    float mx = d0;
    if (d1 > mx) {
        mx = d1;
        p = 1;
    }
    if (d2 > mx) p = 2;
    return p;
}

__global__ __launch_bounds__(kLocalThreads) void k_local_subtrees(Orders O, const float4* __restrict__ prec,
                                                                 const uint32_t* __restrict__ ids0,
                                                                 const uint32_t* __restrict__ posr0, uint32_t first,
                                                                 uint32_t depth_left, const uint2* __restrict__ gseg,
                                                                 const int8_t* __restrict__ gplane,
                                                                 uint32_t* __restrict__ node_idx,
                                                                 int8_t* __restrict__ node_plane) {
    __shared__ LocalLds L;
    const uint32_t root = first + blockIdx.x, t = threadIdx.x;
    const uint2 rs = gseg[root];
    if (rs.y <= rs.x || rs.y - rs.x > (uint32_t)kLocal) return;  // empty subtree; larger ones never reach here
    const uint32_t m = rs.y - rs.x;
    // the segment holds the same photons in each order: the photons of the
    // axis-0 order with their coordinates, and axes 1 and 2 mapped to
    // axis-0 positions by their records' axis-0 ranks
    for (uint32_t j = t; j < m; j += kLocalThreads) {
        const uint32_t id = ids0[rs.x + j];
        const float4 c = prec[3 * (size_t)id];
        L.rec[j] = Rec{{c.x, c.y, c.z}, id};
        L.ordb[0][j] = (uint16_t)j;
        L.ordb[1][j] = (uint16_t)(posr0[O.A[1][rs.x + j].r[0]] - rs.x);
        L.ordb[2][j] = (uint16_t)(posr0[O.A[2][rs.x + j].r[0]] - rs.x);
        L.node[j] = 0;
    }
    if (t == 0) {
        L.seg[0][0] = make_ushort2(0, (uint16_t)m);
        L.plane[0][0] = gplane[root];
    }
    __syncthreads();
    int cur = 0;
    int oi[3] = {0, 1, 2}, spare = 3;  // (workgroup-uniform)
    for (uint32_t d = 0; d < depth_left; ++d) {
        const uint32_t nd = 1u << d;
        const uint32_t gfirst = (root + 1) * nd - 1;  // global index of local node 0 at this depth
        const bool last = d + 1 == depth_left;
        // medians (k_level_nodes) and the children's segments
        for (uint32_t i = t; i < nd; i += kLocalThreads) {
            const ushort2 s = L.seg[cur][i];
            ushort2 cl = make_ushort2(0, 0), cr = make_ushort2(0, 0);
            if (s.y > s.x) {
                const int8_t p = L.plane[cur][i];
                const uint32_t mid = s.x + (s.y - s.x) / 2;
                const Rec r = L.rec[L.ordb[p == 0 ? oi[0] : p == 1 ? oi[1] : oi[2]][mid]];
                node_idx[gfirst + i] = r.id;
                node_plane[gfirst + i] = p;
                if (!last) L.med[i] = make_uint2(__float_as_uint(r.c[p]), r.id);
                if (s.y - s.x > 1) {
                    cl = make_ushort2(s.x, (uint16_t)mid);
                    cr = make_ushort2((uint16_t)(mid + 1), s.y);
                }
            }
            if (!last) {
                L.seg[cur ^ 1][2 * i] = cl;
                L.seg[cur ^ 1][2 * i + 1] = cr;
            }
        }
        __syncthreads();
        if (last) break;
        // stable partition of each axis order around its node's median
        for (int a = 0; a < 3; ++a) {
            for (uint32_t j = t; j < m; j += kLocalThreads) {
                const uint32_t i = L.node[j];
                uint8_t f = 0;
                if (i != 0xFFFFu) {
                    const int sd = side_of(L.rec[L.ordb[oi[a]][j]], L.plane[cur][i], L.med[i]);
                    f = sd == 0;
                    if (sd == 1) L.mpos[i][a] = (uint16_t)j;
                }
                L.flag[j] = f;
            }
            __syncthreads();
            local_scan(L, m);
            for (uint32_t j = t; j < m; j += kLocalThreads) {
                const uint32_t i = L.node[j];
                uint32_t dst = j;
                if (i != 0xFFFFu) {
                    const ushort2 sg = L.seg[cur][i];
                    const uint32_t from = sg.x, mid = sg.x + (sg.y - sg.x) / 2;
                    const uint32_t before_l = (uint32_t)L.pref[j] - L.pref[from];
                    if (L.flag[j]) {
                        dst = from + before_l;
                    } else {
                        const uint32_t mp = L.mpos[i][a];
                        dst = mp == j ? mid : mid + 1 + ((j - from) - before_l - (mp < j ? 1u : 0u));
                    }
                }
                L.ordb[spare][dst] = L.ordb[oi[a]][j];
            }
            __syncthreads();
            const int was = oi[a];
            oi[a] = spare;
            spare = was;
        }
        // node of every position at the next depth
        for (uint32_t j = t; j < m; j += kLocalThreads) {
            const uint32_t i = L.node[j];
            if (i == 0xFFFFu) continue;
            const ushort2 sg = L.seg[cur][i];
            const uint32_t mid = sg.x + (sg.y - sg.x) / 2;
            L.node[j] = (sg.y - sg.x <= 1 || j == mid) ? (uint16_t)0xFFFFu : (uint16_t)(j < mid ? 2 * i : 2 * i + 1);
        }
        // child planes (as k_level_nodes)
        for (uint32_t c = t; c < 2 * nd; c += kLocalThreads) {
            const ushort2 s = L.seg[cur ^ 1][c];
            L.plane[cur ^ 1][c] = s.y > s.x ? local_plane(L, oi, s.x, s.y) : (int8_t)0;
        }
        __syncthreads();
        cur ^= 1;
    }
}

// The tree's photons in node order (nine planes), and each node's 16-byte
// descent record {x, y, z, photon index | plane << 30} (w = 0xFFFFFFFF for
// an empty slot): the kNN reads one record per visited node.
__global__ __launch_bounds__(256) void k_tree_photons(const float4* __restrict__ prec,
                                                     const uint32_t* __restrict__ node_idx,
                                                     const int8_t* __restrict__ node_plane, uint32_t cap,
                                                     float* __restrict__ out9, float4* __restrict__ nodes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    uint32_t id = node_idx[i];
    float v[9];
    if (id != kFinal) {
        const float4 a = prec[3 * (size_t)id], b = prec[3 * (size_t)id + 1], c = prec[3 * (size_t)id + 2];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w; v[8] = c.x;
    } else {
#pragma unroll
        for (int p = 0; p < 9; ++p) v[p] = p < 3 ? __builtin_inff() : 0.f;
    }
#pragma unroll
    for (int p = 0; p < 9; ++p) out9[(size_t)p * cap + i] = v[p];
    const uint32_t w = id != kFinal ? id | ((uint32_t)node_plane[i] << 30) : kFinal;
    nodes[i] = make_float4(v[0], v[1], v[2], __uint_as_float(w));
}

// ---- model AABB (the live effect of create_kd_tree, KdTree.cu:395-464) ----
// One workgroup per model: min/max over the model's vertices with wave
// reductions. Replaces per-triangle AABBs -> 32-triangle chunks -> pairwise
// segmented reduction (KdTree.cu:82-327).
__global__ __launch_bounds__(256) void k_model_aabb(const float* __restrict__ tris9, const uint32_t* __restrict__ tb,
                                                   float* __restrict__ out) {
    __shared__ float s[4][6];
    const uint32_t m = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t v0 = tb[m] * 3, v1 = tb[m + 1] * 3;  // vertex range
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t v = v0 + t; v < v1; v += blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float p = tris9[(size_t)v * 3 + a];
            mn[a] = fminf(mn[a], p);
            mx[a] = fmaxf(mx[a], p);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], d, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], d, 64));
        }
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            s[wave][a] = mn[a];
            s[wave][a + 3] = mx[a];
        }
    __syncthreads();
    if (t < 6) {
        float r = s[0][t];
        for (int w = 1; w < 4; ++w) r = t < 3 ? fminf(r, s[w][t]) : fmaxf(r, s[w][t]);
        out[(size_t)m * 6 + t] = r;
    }
}

}  // namespace

extern "C" pm_status pm_kdtree_build(pm_ctx* ctx, const float* d_photons9, uint32_t n, pm_kdtree** out) {
    if (!ctx || !out) return PM_ERR_INVALID;
    *out = nullptr;
    if (n == 0 || !d_photons9) return pm::fail(ctx, PM_ERR_INVALID, "pm_kdtree_build: no photons");
    if (n > (1u << 30)) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_kdtree_build: n too large");
    uint32_t L = 0;
    while (((1ull << L) - 1) < (uint64_t)n) ++L;
    const uint32_t cap = (uint32_t)((1ull << L) - 1);

    pm_kdtree* T = new pm_kdtree();
    T->n = n;
    T->cap = cap;
    T->levels = L;
    size_t bytes = pm::carve_size({(size_t)cap * 4, (size_t)cap, (size_t)cap * 36, (size_t)cap * 16});
    // (a block of the context's cache when one fits, as for grids)
    hipError_t e = pm::block_take(ctx, bytes, &T->mem, &T->mem_bytes);
    if (e != hipSuccess) {
        delete T;
        return pm::hip_fail(ctx, e, "hipMalloc(kdtree)");
    }
    T->owner = ctx;
    pm::Carve tc{static_cast<char*>(T->mem)};
    T->node_idx = tc.take<uint32_t>(cap);
    T->node_plane = tc.take<int8_t>(cap);
    T->photons = tc.take<float>((size_t)cap * 9);
    T->nodes = tc.take<float4>(cap);
    auto bail = [&](pm_status s) {
        pm_kdtree_destroy(T);
        return s;
    };

    // the three axis sorts run as one segmented sort of 3n keys
    size_t rs = pm::radix_sort_scratch_bytes(n, 3);
    const size_t rb = (size_t)n * sizeof(Rk), n3 = 3 * (size_t)n;
    size_t need = pm::carve_size({rb, rb, rb, rb, rb, rb, (size_t)n * 4, (size_t)n * 4, (size_t)cap * 8, (size_t)cap,
                                  (size_t)cap * 8, n3 * 4, n3 * 4, n3 * 4, (size_t)n * 16, rs, (size_t)n * 48});
    void* work = nullptr;
    pm_status st = pm::scratch_reserve(ctx, need, &work);
    if (st != PM_OK) return bail(st);
    pm::Carve cv{static_cast<char*>(work)};
    Rk* A[3];
    Rk* B[3];
    for (int a = 0; a < 3; ++a) A[a] = cv.take<Rk>(n);
    for (int a = 0; a < 3; ++a) B[a] = cv.take<Rk>(n);
    uint32_t* node_of = cv.take<uint32_t>(n);
    uint32_t* node_of2 = cv.take<uint32_t>(n);
    uint2* seg = cv.take<uint2>(cap);
    int8_t* plane = cv.take<int8_t>(cap);
    uint2* med = cv.take<uint2>(cap);
    uint32_t* keys = cv.take<uint32_t>(n3);   // sort keys
    uint32_t* skeys = cv.take<uint32_t>(n3);  // sorted keys, then ids0 and posr0 (the LDS levels)
    uint32_t* order = cv.take<uint32_t>(n3);  // each axis's photons in rank order
    uint32_t* rank4 = cv.take<uint32_t>(4 * (size_t)n);  // each photon's three ranks
    void* rw = cv.take<char>(rs);
    float4* prec = cv.take<float4>((size_t)n * 3);

    const uint32_t nb = pm::div_up(n, 256);
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_photon_records, dim3(nb), dim3(256), 0, s, d_photons9, n, prec);
    // planes 0-2 of the block are the three coordinates, back to back
    hipLaunchKernelGGL(k_axis_keys, dim3(pm::div_up((uint32_t)n3, 256)), dim3(256), 0, s, d_photons9, (uint32_t)n3,
                       keys);
    if ((st = pm::radix_sort_pairs_auto(ctx, keys, nullptr, skeys, order, n, 32, rw, nullptr, 3)) != PM_OK)
        return bail(st);
    hipLaunchKernelGGL(k_ranks, dim3(nb, 3), dim3(256), 0, s, order, n, rank4);
    hipLaunchKernelGGL(k_rank_recs, dim3(nb, 3), dim3(256), 0, s, order, reinterpret_cast<const uint4*>(rank4), n,
                       A[0], A[1], A[2]);
    auto orders = [&] {
        Orders O;
        for (int a = 0; a < 3; ++a) O.A[a] = A[a];
        O.order = order;
        O.rank4 = rank4;
        O.ph9 = d_photons9;
        O.n = n;
        return O;
    };
    if ((st = pm::fill_u32(ctx, T->node_idx, kFinal, cap)) != PM_OK) return bail(st);
    // node_plane = -1; the carve pads every block to 256 bytes, so whole
    // words past cap stay inside the allocation.
    if ((st = pm::fill_u32(ctx, reinterpret_cast<uint32_t*>(T->node_plane), 0xFFFFFFFFu, (cap + 3) / 4)) != PM_OK)
        return bail(st);
    if ((st = pm::fill_u32(ctx, node_of, 0u, n)) != PM_OK) return bail(st);

    // Levels [0, Lg) run on the whole arrays; from level Lg every segment
    // holds at most kLocal records (a node of size s has children of at most
    // s / 2) and each subtree is finished by one workgroup in LDS.
    uint32_t Lg = 0;
    for (uint32_t smax = n; smax > (uint32_t)kLocal; smax /= 2) ++Lg;
    // a segment of at most 1024 records spans at most 11 levels (node arrays of kLocal)
    if (L - Lg > 11) return bail(pm::fail(ctx, PM_ERR_INVALID, "pm_kdtree_build: internal level split"));
    // the root's plane (and, above the LDS levels, its median)
    hipLaunchKernelGGL(k_level_nodes, dim3(1), dim3(256), 0, s, orders(), 0u, 1u, cap, n, Lg > 0, seg, plane,
                       T->node_idx, T->node_plane, med, nullptr);
    for (uint32_t lvl = 0; lvl < L; ++lvl) {
        uint32_t first = (1u << lvl) - 1, count = 1u << lvl;
        if (lvl == Lg) {
            uint32_t *ids0 = skeys, *posr0 = skeys + n;
            hipLaunchKernelGGL(k_pos0, dim3(nb), dim3(256), 0, s, orders(), ids0, posr0);
            hipLaunchKernelGGL(k_local_subtrees, dim3(count), dim3(kLocalThreads), 0, s, orders(), prec, ids0, posr0,
                               first, L - lvl, seg, plane, T->node_idx, T->node_plane);
            break;
        }
        if (lvl + 1 == L) break;  // every segment of the last level has size <= 1
        const uint32_t tiles = pm::div_up(3ull * n, kPTile);
        uint64_t* lstate;
        uint32_t* lcounter;
        if ((st = pm::lookback_begin(ctx, tiles, &lstate, &lcounter)) != PM_OK) return bail(st);
        hipLaunchKernelGGL(k_level_partition, dim3(tiles), dim3(kPThreads), 0, s, A[0], A[1], A[2], B[0], B[1], B[2],
                           n, node_of, node_of2, seg, plane, med, lstate, lcounter, tiles);
        for (int a = 0; a < 3; ++a) std::swap(A[a], B[a]);
        std::swap(node_of, node_of2);
        // the children's planes, and their medians above the LDS levels
        uint32_t cf = (1u << (lvl + 1)) - 1, cc = 1u << (lvl + 1);
        hipLaunchKernelGGL(k_level_nodes, dim3(pm::div_up(cc, 256)), dim3(256), 0, s, orders(), cf, cc, cap, 0u,
                           lvl + 1 < Lg, seg, plane, T->node_idx, T->node_plane, med, lcounter);
    }
    hipLaunchKernelGGL(k_tree_photons, dim3(pm::div_up(cap, 256)), dim3(256), 0, s, prec, T->node_idx,
                       T->node_plane, cap, T->photons, T->nodes);
    if (hipGetLastError() != hipSuccess) return bail(pm::fail(ctx, PM_ERR_HIP, "pm_kdtree_build: launch failed"));
    *out = T;
    return PM_OK;
}

extern "C" pm_status pm_kdtree_describe(const pm_kdtree* t, pm_kdtree_desc* out) {
    if (!t || !out) return PM_ERR_INVALID;
    out->n_photons = t->n;
    out->n_slots = t->cap;
    out->levels = t->levels;
    out->d_node_idx = t->node_idx;
    out->d_node_plane = t->node_plane;
    out->d_photons = t->photons;
    return PM_OK;
}

extern "C" pm_status pm_kdtree_copy_to_host(pm_ctx* ctx, const pm_kdtree* t, uint32_t* h_node_idx,
                                            int8_t* h_node_plane) {
    if (!ctx || !t) return PM_ERR_INVALID;
    if (h_node_idx) PM_HIP(ctx, hipMemcpyAsync(h_node_idx, t->node_idx, (size_t)t->cap * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (h_node_plane) PM_HIP(ctx, hipMemcpyAsync(h_node_plane, t->node_plane, t->cap, hipMemcpyDeviceToHost, ctx->stream));
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PM_OK;
}

extern "C" pm_status pm_kdtree_destroy(pm_kdtree* t) {
    if (!t) return PM_OK;
    if (t->mem && !pm::block_give(t->owner, t->mem, t->mem_bytes)) (void)hipFree(t->mem);
    delete t;
    return PM_OK;
}

namespace pm {
pm_status model_aabb_launch(pm_ctx* ctx, const float* tris9, const uint32_t* tb, int n_models, float* out6) {
    if (n_models <= 0) return PM_OK;
    hipLaunchKernelGGL(k_model_aabb, dim3(n_models), dim3(256), 0, ctx->stream, tris9, tb, out6);
    PM_CHECK_LAUNCH(ctx, "k_model_aabb");
    return PM_OK;
}
}  // namespace pm

extern "C" pm_status pm_model_aabb(pm_ctx* ctx, const float* d_tris9, const uint32_t* d_tri_begin, int32_t n_models,
                                   float* d_aabb6) {
    if (!ctx) return PM_ERR_INVALID;
    if (n_models < 0 || (n_models > 0 && (!d_tris9 || !d_tri_begin || !d_aabb6)))
        return pm::fail(ctx, PM_ERR_INVALID, "pm_model_aabb: bad arguments");
    if (n_models == 0) return PM_OK;
    hipLaunchKernelGGL(k_model_aabb, dim3(n_models), dim3(256), 0, ctx->stream, d_tris9, d_tri_begin, d_aabb6);
    PM_CHECK_LAUNCH(ctx, "k_model_aabb");
    return PM_OK;
}
