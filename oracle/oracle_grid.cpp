// CPU oracle: prefix scan, hash grid, grid kNN, radiance estimate, model AABB.
// TEST INFRASTRUCTURE ONLY (see oracle_math.hpp). Each function cites the
// reference lines it restates.
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

extern "C" void orc_scan_u32(const uint32_t* in, uint32_t n, int inclusive, uint32_t* out, uint64_t* total) {
    // PrefixScanSum.cuh:12-148 computes, in u32, the exclusive (inclusive)
    // running sum; the KAT in DynamicArrayTest.cu:128-144 checks it against
    // a sequential sum.
    uint32_t s = 0;
    uint64_t s64 = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (inclusive) {
            s += in[i];
            out[i] = s;
        } else {
            out[i] = s;
            s += in[i];
        }
        s64 += in[i];
    }
    if (total) *total = s64;
}

namespace {

struct Dims {
    float mn[3];
    float cell;
    int g[3];
    int layout;
};

// separated_hash_position (PhotonGrid.cu:10-23): i = (int)((p - min) / cell),
// clamped to [0, g-1]. The float is range-checked before the conversion so
// out-of-range and NaN inputs have a defined result (the reference's cast is
// undefined there); inside the AABB this equals the reference.
static inline int cell_coord(float p, float mn, float cell, int g) {
    float f = (p - mn) / cell;
    if (!(f < (float)g)) return g - 1;
    if (f < 1.0f) return 0;
    int i = (int)f;
    return i < g - 1 ? i : g - 1;
}

static inline uint32_t spread3(uint32_t v) {
    v &= 0x3FF;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

// cell_idx (PhotonGrid.cu:25-31) for layout 0; Morton code for layout 1.
static inline uint32_t cell_index(const Dims& d, int x, int y, int z) {
    if (d.layout == 1) return spread3((uint32_t)x) | (spread3((uint32_t)y) << 1) | (spread3((uint32_t)z) << 2);
    return (uint32_t)x + (uint32_t)y * (uint32_t)d.g[0] + (uint32_t)z * (uint32_t)d.g[0] * (uint32_t)d.g[1];
}

static inline uint32_t n_cells_of(const Dims& d) {
    if (d.layout == 1) {
        int m = std::max(d.g[0], std::max(d.g[1], d.g[2]));
        uint32_t side = 1;
        while ((int)side < m) side <<= 1;
        return side * side * side;
    }
    return (uint32_t)d.g[0] * (uint32_t)d.g[1] * (uint32_t)d.g[2];
}

static Dims make_dims(const float aabb[6], float cell, int layout) {
    Dims d;
    for (int a = 0; a < 3; ++a) d.mn[a] = aabb[a];
    d.cell = cell;
    d.layout = layout;
    for (int a = 0; a < 3; ++a) {
        // AABB::length_by_axis (AABB.cuh:212-214) and PhotonGrid.cu:53-55.
        float size = std::fabs(aabb[a + 3] - aabb[a]);
        int g = (int)std::ceil(size / cell);
        d.g[a] = g < 1 ? 1 : g;
    }
    return d;
}

// compute_grid_and_cell_size (PhotonGrid.cu:41-56).
static float reference_cell(uint32_t n, const float aabb[6], int desired) {
    float sx = std::fabs(aabb[3] - aabb[0]);
    float sy = std::fabs(aabb[4] - aabb[1]);
    float sz = std::fabs(aabb[5] - aabb[2]);
    int N = ((int)n + desired - 1) / desired;
    float volume = sx * sy * sz;
    return std::cbrt(volume / (float)N);
}

static void count_cells(const float* ph, uint32_t n, const Dims& d, std::vector<uint32_t>& counts) {
    counts.assign(n_cells_of(d), 0u);
    const float* px = ph;
    const float* py = ph + n;
    const float* pz = ph + 2 * (size_t)n;
    for (uint32_t i = 0; i < n; ++i) {
        int x = cell_coord(px[i], d.mn[0], d.cell, d.g[0]);
        int y = cell_coord(py[i], d.mn[1], d.cell, d.g[1]);
        int z = cell_coord(pz[i], d.mn[2], d.cell, d.g[2]);
        counts[cell_index(d, x, y, z)]++;
    }
}

}  // namespace

extern "C" int orc_grid_plan(const float* photons9, uint32_t n, const float aabb[6], float cell_size,
                             int32_t desired, int32_t layout, float* out_cell, int32_t out_dims[3],
                             uint32_t* out_ncells) {
    if (!aabb || !out_cell || !out_dims || layout < 0 || layout > 1) return -1;
    float cell;
    if (cell_size > 0.f) {
        cell = cell_size;
    } else {
        if (desired <= 0 || n == 0 || !photons9) return -1;
        cell = reference_cell(n, aabb, desired);
        // analyze_density (PhotonGrid.cu:92-141): re-size when the fullest
        // cell holds more than 5 * desired photons.
        Dims d = make_dims(aabb, cell, layout);
        std::vector<uint32_t> counts;
        count_cells(photons9, n, d, counts);
        uint32_t mx = 0;
        for (uint32_t c : counts) mx = std::max(mx, c);
        if (mx > 5u * (uint32_t)desired) cell = reference_cell(n, aabb, desired / 5);
    }
    Dims d = make_dims(aabb, cell, layout);
    *out_cell = cell;
    for (int a = 0; a < 3; ++a) out_dims[a] = d.g[a];
    if (out_ncells) *out_ncells = n_cells_of(d);
    return 0;
}

extern "C" int orc_grid_fill(const float* photons9, uint32_t n, const float aabb[6], float cell,
                             const int32_t dims[3], int32_t layout, uint32_t* counts, uint32_t* starts,
                             float* sorted9, uint32_t* perm) {
    Dims d;
    for (int a = 0; a < 3; ++a) {
        d.mn[a] = aabb[a];
        d.g[a] = dims[a];
    }
    d.cell = cell;
    d.layout = layout;
    uint32_t nc = n_cells_of(d);
    std::vector<uint32_t> cnt;
    count_cells(photons9, n, d, cnt);
    std::vector<uint32_t> st(nc);
    orc_scan_u32(cnt.data(), nc, 0, st.data(), nullptr);
    std::vector<uint32_t> fill(nc, 0u);
    const float* px = photons9;
    const float* py = photons9 + n;
    const float* pz = photons9 + 2 * (size_t)n;
    for (uint32_t i = 0; i < n; ++i) {  // ascending i: stable in-cell order
        int x = cell_coord(px[i], d.mn[0], d.cell, d.g[0]);
        int y = cell_coord(py[i], d.mn[1], d.cell, d.g[1]);
        int z = cell_coord(pz[i], d.mn[2], d.cell, d.g[2]);
        uint32_t c = cell_index(d, x, y, z);
        uint32_t dst = st[c] + fill[c]++;
        if (perm) perm[dst] = i;
        if (sorted9)
            for (int p = 0; p < 9; ++p) sorted9[(size_t)p * n + dst] = photons9[(size_t)p * n + i];
    }
    if (counts) std::copy(cnt.begin(), cnt.end(), counts);
    if (starts) std::copy(st.begin(), st.end(), starts);
    return 0;
}

namespace {

struct Cand {
    uint32_t d2bits;
    uint32_t idx;
    float d2;
};
static inline bool cand_less(const Cand& a, const Cand& b) {
    return a.d2bits < b.d2bits || (a.d2bits == b.d2bits && a.idx < b.idx);
}
struct CandMax {
    bool operator()(const Cand& a, const Cand& b) const { return cand_less(a, b); }
};

// Bounded max-heap keeping the k smallest (d2, idx) keys. Replaces
// PhotonMaxHeap::push (PhotonMaxHeap.cuh:90-107), whose strict '<' made the
// kept set depend on visit order under ties.
struct TopK {
    int k;
    std::priority_queue<Cand, std::vector<Cand>, CandMax> h;
    void push(const Cand& c) {
        if ((int)h.size() < k) {
            h.push(c);
        } else if (cand_less(c, h.top())) {
            h.pop();
            h.push(c);
        }
    }
    void emit(uint32_t* idx, float* dist, uint32_t* count) {
        int m = (int)h.size();
        std::vector<Cand> v(m);
        for (int i = m - 1; i >= 0; --i) {
            v[i] = h.top();
            h.pop();
        }
        for (int i = 0; i < k; ++i) {
            if (i < m) {
                idx[i] = v[i].idx;
                dist[i] = std::sqrt(v[i].d2);
            } else {
                idx[i] = 0xFFFFFFFFu;
                dist[i] = INFINITY;
            }
        }
        *count = (uint32_t)m;
    }
};

}  // namespace

extern "C" int orc_grid_knn(const uint32_t* counts, const uint32_t* starts, const float* sorted9, uint32_t n,
                            const float aabb[6], float cell, const int32_t dims[3], int32_t layout,
                            const float* queries3, uint32_t nq, float r, int32_t k, int32_t threads,
                            uint32_t* out_idx, float* out_dist, uint32_t* out_count, uint64_t* sum_cells,
                            uint64_t* sum_cands, uint64_t* sum_found) {
    if (k <= 0 || r < 0.f) return -1;
    Dims d;
    for (int a = 0; a < 3; ++a) {
        d.mn[a] = aabb[a];
        d.g[a] = dims[a];
    }
    d.cell = cell;
    d.layout = layout;
    const float* px = sorted9;
    const float* py = sorted9 + n;
    const float* pz = sorted9 + 2 * (size_t)n;
    const float* qx = queries3;
    const float* qy = queries3 + nq;
    const float* qz = queries3 + 2 * (size_t)nq;
    uint64_t tc = 0, tp = 0, tf = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tc, tp, tf)
#endif
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        float q[3] = {qx[qi], qy[qi], qz[qi]};
        int c[3];
        for (int a = 0; a < 3; ++a) c[a] = cell_coord(q[a], d.mn[a], d.cell, d.g[a]);
        uint32_t own = cell_index(d, c[0], c[1], c[2]);
        // PhotonGrid.cu:266-267: rc = ceilf(r / cell); box clamped to the grid.
        int rc = (int)std::ceil(r / d.cell);
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::max(0, c[a] - rc);
            hi[a] = std::min(d.g[a] - 1, c[a] + rc);
        }
        TopK top;
        top.k = k;
        uint32_t found = 0;
        auto visit = [&](uint32_t cid) {
            uint32_t s = starts[cid], m = counts[cid];
            tp += m;
            for (uint32_t j = s; j < s + m; ++j) {
                // find_nearests_in_cell (PhotonGrid.cu:241-255): pos -= q;
                // distance = length(); push if distance <= radius.
                float dx = px[j] - q[0], dy = py[j] - q[1], dz = pz[j] - q[2];
                float d2 = dx * dx + dy * dy + dz * dz;
                if (std::sqrt(d2) <= r) {
                    ++found;
                    top.push(Cand{fbits(d2), j, d2});
                }
            }
        };
        // PhotonGrid.cu:269-281: query cell first, then z / y / x skipping it.
        visit(own);
        tc += 1;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    uint32_t cid = cell_index(d, x, y, z);
                    if (cid == own) continue;
                    tc += 1;
                    visit(cid);
                }
        tf += found;
        top.emit(out_idx + (size_t)qi * k, out_dist + (size_t)qi * k, out_count + qi);
    }
    if (sum_cells) *sum_cells = tc;
    if (sum_cands) *sum_cands = tp;
    if (sum_found) *sum_found = tf;
    return 0;
}

extern "C" int orc_knn_brute(const float* pos3, uint32_t n, const float* queries3, uint32_t nq, float r,
                             int32_t k, uint32_t* out_idx, float* out_dist, uint32_t* out_count) {
    if (k <= 0 || r < 0.f) return -1;
    const float* px = pos3;
    const float* py = pos3 + n;
    const float* pz = pos3 + 2 * (size_t)n;
    for (uint32_t qi = 0; qi < nq; ++qi) {
        float q[3] = {queries3[qi], queries3[nq + qi], queries3[2 * (size_t)nq + qi]};
        TopK top;
        top.k = k;
        for (uint32_t j = 0; j < n; ++j) {
            float dx = px[j] - q[0], dy = py[j] - q[1], dz = pz[j] - q[2];
            float d2 = dx * dx + dy * dy + dz * dz;
            if (std::sqrt(d2) <= r) top.push(Cand{fbits(d2), j, d2});
        }
        top.emit(out_idx + (size_t)qi * k, out_dist + (size_t)qi * k, out_count + qi);
    }
    return 0;
}

extern "C" int orc_radiance(const uint32_t* idx, const float* dist, const uint32_t* count, uint32_t nq,
                            int32_t k, const float* power3, const float* dir3, uint32_t n,
                            const float* normals3, float* out_rgb32, double* out_rgb64) {
    // Raytracer::get_color_from_photon_grid (Raytracer.cuh:110-138).
    const float alpha = 1.818f, beta = 1.953f;
    const float pi_f = 3.141592654f;  // CUDART_PI_F
    for (uint32_t q = 0; q < nq; ++q) {
        uint32_t m = count[q];
        V3 nrm = v3(normals3[q], normals3[nq + q], normals3[2 * (size_t)nq + q]);
        float res[3] = {0.f, 0.f, 0.f};
        double res64[3] = {0.0, 0.0, 0.0};
        if (m > 0) {
            const uint32_t* qi = idx + (size_t)q * k;
            const float* qd = dist + (size_t)q * k;
            float r = qd[m - 1];  // heap root = largest distance of the set
            float filter_r = r * r;
            double r64 = (double)r;
            double fr64 = r64 * r64;
            for (uint32_t i = 0; i < m; ++i) {
                uint32_t p = qi[i];
                V3 dirp = v3(dir3[p], dir3[n + p], dir3[2 * (size_t)n + p]);
                float cosNL = dot(nrm, neg(dirp));
                if (cosNL > 0) {
                    float distance = qd[i];
                    float filter = alpha * (1.f - (1.f - std::exp(-beta * distance * distance / (2 * filter_r))) /
                                                       (1.f - std::exp(-beta)));
                    for (int c = 0; c < 3; ++c) res[c] += power3[(size_t)c * n + p] * filter;
                    double d64 = (double)distance;
                    double f64 = (double)alpha * (1.0 - (1.0 - std::exp(-(double)beta * d64 * d64 / (2.0 * fr64))) /
                                                           (1.0 - std::exp(-(double)beta)));
                    for (int c = 0; c < 3; ++c) res64[c] += (double)power3[(size_t)c * n + p] * f64;
                }
            }
            float coef = 1.f / (pi_f * (r * r) * 1);
            for (int c = 0; c < 3; ++c) res[c] *= coef;
            double coef64 = 1.0 / ((double)pi_f * (r64 * r64));
            for (int c = 0; c < 3; ++c) res64[c] *= coef64;
        }
        for (int c = 0; c < 3; ++c) {
            if (out_rgb32) out_rgb32[(size_t)c * nq + q] = res[c];
            if (out_rgb64) out_rgb64[(size_t)c * nq + q] = res64[c];
        }
    }
    return 0;
}

extern "C" float orc_radius_threshold(float r) {
    if (!(r >= 0.f)) return -1.f;
    float t = r * r;
    while (std::sqrt(t) > r) t = std::nextafter(t, 0.f);
    for (;;) {
        float u = std::nextafter(t, INFINITY);
        if (std::sqrt(u) <= r)
            t = u;
        else
            break;
    }
    return t;
}

extern "C" int orc_model_aabb(const float* verts9, const uint32_t* tri_begin, int32_t n_models, float* out) {
    // AABB::fill / add (AABB.cuh:143-167): first vertex initialises, then
    // strict > max / strict < min per axis.
    for (int m = 0; m < n_models; ++m) {
        float* b = out + 6 * (size_t)m;
        uint32_t t0 = tri_begin[m], t1 = tri_begin[m + 1];
        if (t1 <= t0) {
            b[0] = b[1] = b[2] = FLT_MAX;
            b[3] = b[4] = b[5] = -FLT_MAX;
            continue;
        }
        for (int a = 0; a < 3; ++a) b[a] = b[a + 3] = verts9[9 * (size_t)t0 + a];
        for (uint32_t t = t0; t < t1; ++t)
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    float p = verts9[9 * (size_t)t + 3 * v + a];
                    if (p > b[a + 3]) b[a + 3] = p;
                    if (p < b[a]) b[a] = p;
                }
    }
    return 0;
}
