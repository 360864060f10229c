// CPU oracle: photon kd-tree (PhotonMap) build and kd kNN.
// TEST INFRASTRUCTURE ONLY (see oracle_math.hpp).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <queue>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "oracle_math.hpp"

using namespace orc;

namespace {

// PhotonMap::find_largest_plane (PhotonMap.cuh:161-178): argmax of the
// extents with a strict '>' so the lowest axis wins ties.
static int largest_plane(const float big[3], const float small[3]) {
    float dims[3];
    for (int a = 0; a < 3; ++a) dims[a] = std::fabs(big[a] - small[a]);
    int plane = 0;
    float mx = dims[0];
    for (int a = 1; a < 3; ++a)
        if (dims[a] > mx) {
            mx = dims[a];
            plane = a;
        }
    return plane;
}

// PhotonMap::update_cube (PhotonMap.cuh:179-188).
static inline void update_cube(const float p[3], float big[3], float small[3]) {
    for (int a = 0; a < 3; ++a) {
        if (p[a] > big[a]) big[a] = p[a];
        if (p[a] < small[a]) small[a] = p[a];
    }
}

struct Frame {
    size_t from, to;
    int plane;
    uint32_t node;
};

}  // namespace

extern "C" int orc_kdtree_build(const float* pos3, uint32_t n, uint32_t cap, uint32_t* node_idx,
                                int8_t* node_plane) {
    for (uint32_t i = 0; i < cap; ++i) {
        node_idx[i] = 0xFFFFFFFFu;
        node_plane[i] = -1;
    }
    if (n == 0) return 0;
    const float* P[3] = {pos3, pos3 + n, pos3 + 2 * (size_t)n};
    auto pos = [&](uint32_t i, float out[3]) {
        out[0] = P[0][i];
        out[1] = P[1][i];
        out[2] = P[2][i];
    };
    std::vector<uint32_t> ptrs(n);
    float big[3], small[3], p[3];
    pos(0, big);
    pos(0, small);
    for (uint32_t i = 0; i < n; ++i) {  // PhotonMap.cuh:205-209
        ptrs[i] = i;
        pos(i, p);
        update_cube(p, big, small);
    }
    std::vector<Frame> stack;
    stack.push_back(Frame{0, n, largest_plane(big, small), 0});
    while (!stack.empty()) {  // PhotonMap.cuh:221-280
        Frame f = stack.back();
        stack.pop_back();
        if (f.to - f.from == 0) continue;
        if (f.node >= cap) return -2;
        if (f.to - f.from == 1) {
            node_idx[f.node] = ptrs[f.from];
            node_plane[f.node] = (int8_t)f.plane;
            continue;
        }
        // quick_sort by pos[plane] (Sort.cuh:8-48); ties ordered by photon
        // index so the median is unique.
        const float* key = P[f.plane];
        std::sort(ptrs.begin() + f.from, ptrs.begin() + f.to, [key](uint32_t a, uint32_t b) {
            return key[a] < key[b] || (key[a] == key[b] && a < b);
        });
        size_t mid = (f.to - f.from) / 2 + f.from;
        float lb[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX}, ls[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
        float rb[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX}, rs[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
        for (size_t j = f.from; j < mid; ++j) {
            pos(ptrs[j], p);
            update_cube(p, lb, ls);
        }
        for (size_t j = mid + 1; j < f.to; ++j) {
            pos(ptrs[j], p);
            update_cube(p, rb, rs);
        }
        node_idx[f.node] = ptrs[mid];
        node_plane[f.node] = (int8_t)f.plane;
        stack.push_back(Frame{f.from, mid, largest_plane(lb, ls), 2 * f.node + 1});
        stack.push_back(Frame{mid + 1, f.to, largest_plane(rb, rs), 2 * f.node + 2});
    }
    return 0;
}

namespace {
struct KC {
    uint32_t bits, idx;
    float d2;
};
static inline bool kc_less(const KC& a, const KC& b) {
    return a.bits < b.bits || (a.bits == b.bits && a.idx < b.idx);
}
struct KCMax {
    bool operator()(const KC& a, const KC& b) const { return kc_less(a, b); }
};
}  // namespace

extern "C" int orc_kdtree_knn(const float* pos3, uint32_t n, const uint32_t* node_idx, const int8_t* node_plane,
                              uint32_t cap, const float* queries3, uint32_t nq, float r, int32_t k, int32_t threads,
                              uint32_t* out_idx, float* out_dist, uint32_t* out_count) {
    // A standard depth-first kd kNN replaces the debug locate_q
    // (PhotonMap.cuh:289-371). The far side is pruned only when the plane
    // distance squared exceeds the current bound by a relative margin, so no
    // photon inside the radius is ever skipped; the result is therefore the
    // exact (d2, index) selection a brute-force scan returns.
    if (k <= 0 || r < 0.f) return -1;
    const float* P[3] = {pos3, pos3 + n, pos3 + 2 * (size_t)n};
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        float q[3] = {queries3[qi], queries3[nq + qi], queries3[2 * (size_t)nq + qi]};
        std::priority_queue<KC, std::vector<KC>, KCMax> h;
        double r2 = (double)r * (double)r;
        std::vector<uint32_t> st;
        st.push_back(0);
        while (!st.empty()) {
            uint32_t node = st.back();
            st.pop_back();
            if (node >= cap || node_idx[node] == 0xFFFFFFFFu) continue;
            uint32_t pi = node_idx[node];
            int plane = node_plane[node];
            float dx = P[0][pi] - q[0], dy = P[1][pi] - q[1], dz = P[2][pi] - q[2];
            float d2 = dx * dx + dy * dy + dz * dz;
            if (std::sqrt(d2) <= r) {
                KC c{fbits(d2), pi, d2};
                if ((int)h.size() < k)
                    h.push(c);
                else if (kc_less(c, h.top())) {
                    h.pop();
                    h.push(c);
                }
            }
            double bound = r2;
            if ((int)h.size() == k) bound = std::min(bound, (double)h.top().d2);
            double diff = (double)q[plane] - (double)P[plane][pi];
            uint32_t nearc = diff <= 0 ? 2 * node + 1 : 2 * node + 2;
            uint32_t farc = diff <= 0 ? 2 * node + 2 : 2 * node + 1;
            if (diff * diff <= bound * (1.0 + 1e-5) + 1e-30) st.push_back(farc);
            st.push_back(nearc);
        }
        int m = (int)h.size();
        std::vector<KC> v(m);
        for (int i = m - 1; i >= 0; --i) {
            v[i] = h.top();
            h.pop();
        }
        for (int i = 0; i < k; ++i) {
            out_idx[(size_t)qi * k + i] = i < m ? v[i].idx : 0xFFFFFFFFu;
            out_dist[(size_t)qi * k + i] = i < m ? std::sqrt(v[i].d2) : INFINITY;
        }
        out_count[qi] = (uint32_t)m;
    }
    return 0;
}
