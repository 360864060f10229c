// k-nearest-photon gather and radiance estimate on the hash grid.
//
// Replaces PhotonGrid::find_nearests + PhotonMaxHeap (PhotonGrid.cu:222-284,
// PhotonMaxHeap.cuh:90-107) and Raytracer::get_color_from_photon_grid
// (Raytracer.cuh:110-138). One wave64 serves one query:
//   1. the candidate box (query cell +- ceil(r/cell), clamped) is split into
//      contiguous photon runs (one per (y,z) row of cells in the linear
//      layout, one per cell in the Morton layout), one run per lane;
//   2. the candidates stream 64 at a time (coalesced SoA loads, next step
//      prefetched); d2 = (dx*dx + dy*dy) + dz*dz exactly as the reference
//      host computes it, radius test d2 <= T with T the largest float whose
//      sqrtf is <= r (identical to the reference's sqrtf(d2) <= r);
//   3. in-radius keys (d2 bits << 32 | grid slot) are appended to an LDS
//      buffer with ballot + mbcnt; when it fills, a histogram over the top
//      float bits of d2 (two bins per octave) finds the smallest bin edge
//      keeping >= k keys, the buffer is compacted to it and the acceptance
//      bound tightens (a refinement splits a crowded bin 64 ways);
//   4. a register bitonic sort across the wave orders the <= 256 survivors;
//      the first k are the answer — the exact k smallest by (d2, slot);
//   5. optionally the radiance estimate runs on those k in the same wave.
// A two-pass variant (histogram pass, then collect pass) is kept selectable
// for tuning.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pm_select.hpp"

namespace {

using pmd::GridGeom;
using namespace pms;

constexpr int kWaves = 4;  // queries per 256-thread workgroup
constexpr int kMaxK = 2048;

struct KnnParams {
    GridGeom G;
    const uint32_t* starts;
    const uint32_t* sizes;
    const float* px;
    const float* py;
    const float* pz;
    const float* power;  // 3 planes of n
    const float* dir;    // 3 planes of n
    uint32_t n;
    const float* qx;
    const float* qy;
    const float* qz;
    const float* nrm;  // 3 planes of nq (radiance only)
    uint32_t nq;
    const uint32_t* order;  // optional processing order
    float r;
    float T;      // largest d2 with sqrtf(d2) <= r
    uint32_t bT;  // bits(T) >> 22
    int rc;
    int k;
    uint32_t* out_idx;
    float* out_dist;
    uint32_t* out_count;
    float* rgb;  // 3 planes of nq (radiance only)
    int remap;
    int h;  // inner box half-width in cells (0: walk the whole box at once)
    // (nullable) the query count on the device, at most nq: positions past
    // it are skipped; nq stays the planes' stride
    const uint32_t* nq_dev;
};

struct Box {
    int lo[3], hi[3];
};

// One FP32 plane as a raw buffer: loads take a 32-bit byte offset (one VALU
// address for the three position planes) and return 0 past the plane's end.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* p, uint32_t n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, (int)(n * 4u), 0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0));
}

// Lower bound of |v - p| along one axis over the photons hashed into cell c
// (cells 0 and g-1 also hold every photon clamped into them). The cell's
// bounds are widened by a relative 1e-6, far above the rounding of the hash
// (p - mn) / cell and of the bounds themselves.
__device__ __forceinline__ float axis_gap(float v, int c, int g, float mn, float cell) {
    const float lo = mn + (float)c * cell, hi = mn + (float)(c + 1) * cell;
    const float e = (fabsf(lo) + fabsf(hi) + fabsf(mn)) * 1e-6f;
    if (c > 0 && v < lo - e) return (lo - e) - v;
    if (c < g - 1 && v > hi + e) return v - (hi + e);
    return 0.f;
}

// Lower bound of |v - p| along one axis over the photons hashed outside the
// cells [lo, hi] (cells 0 and g-1 are unbounded outward, so a side at the
// grid's edge has no photons beyond it). The cell edges are widened by a
// relative 1e-6, far above the rounding of the hash and of the edges.
__device__ __forceinline__ float inner_gap(float v, int lo, int hi, int g, float mn, float cell) {
    float gap = __builtin_inff();
    if (lo > 0) {
        const float b = mn + (float)lo * cell;
        gap = fminf(gap, v - (b + (fabsf(b) + fabsf(mn)) * 1e-6f));
    }
    if (hi < g - 1) {
        const float b = mn + (float)(hi + 1) * cell;
        gap = fminf(gap, (b - (fabsf(b) + fabsf(mn)) * 1e-6f) - v);
    }
    return fmaxf(gap, 0.f);
}

// Photon run s of the box (linear layout: the cells of one (y, z) row;
// Morton: one cell) as [b, b + len) in grid-slot order, pruned to the cells
// that can hold a photon within the radius: a row whose (y, z) gap alone
// exceeds it is skipped, and the x extent is cut to the cells meeting the
// slab |x - qx| <= sqrt(T - gap_yz^2). Every bound is widened so that a
// photon passing the FP32 test d2 <= T is never cut: the squared gaps are
// compared with T (1 + 1e-3) (rounding of d2 and of the gaps is a few ulps),
// and the slab half-width is sqrt(T (1 + 1e-3) - gap_yz^2 + 1e-3 T) plus a
// relative 1e-6 before its cells are bounded (cell_bounds, conservative
// against the grid's own monotone cell_coord).
// The kept candidates are a subset of the reference's box holding all of
// its candidates within r, so results are unchanged.
// The cells [xl, xh] that the grid's hash (cell_coord: the correctly
// rounded quotient (p - mn) / cell, truncated and clamped to [0, g - 1];
// monotone in p) gives the positions in [lo, hi]: the quotients are taken
// with the hardware reciprocal, within 2^-21 relative of the divided ones,
// moved out by a relative and an absolute 2^-20, floored and clamped like
// the hash, so every such position's cell is inside (a bound can only come
// out one cell wider).
__device__ __forceinline__ void cell_bounds(float lo, float hi, float mn, float rcell, int g, int& xl, int& xh) {
    float tl = (lo - mn) * rcell, th = (hi - mn) * rcell;
    tl = tl - (fabsf(tl) * 0x1p-20f + 0x1p-20f);
    th = th + (fabsf(th) * 0x1p-20f + 0x1p-20f);
    const float top = (float)(g - 1);
    xl = (int)floorf(fminf(fmaxf(tl, 0.f), top));  // (NaN gives 0 and g - 1)
    xh = (int)floorf(fmaxf(fminf(th, top), 0.f));
}

// q = s / d and r = s % d for s < 2^20 (and d >= 1): (s + 0.5) / d is at
// least 0.5 / d from the integers around it, and the product with the
// hardware reciprocal is within (s + 0.5) 2^-22 / d of it.
__device__ __forceinline__ void small_divmod(int s, int d, float rd, int& q, int& r) {
    q = (int)(((float)s + 0.5f) * rd);
    r = s - q * d;
}

__device__ __forceinline__ bool row_span(const GridGeom& G, const uint32_t* __restrict__ starts,
                                         const uint32_t* __restrict__ sizes, const Box& B, int s, float qx, float qy,
                                         float qz, float T, uint32_t& b, uint32_t& len) {
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const float thr = T * 1.001f;
    b = 0;
    len = 0;
    if (G.layout == PM_LAYOUT_LINEAR) {
        int zi, yi;
        if (ny * nz < (1 << 20))
            small_divmod(s, ny, __builtin_amdgcn_rcpf((float)ny), zi, yi);
        else
            zi = s / ny, yi = s % ny;
        const int y = B.lo[1] + yi, z = B.lo[2] + zi;
        const float gy = axis_gap(qy, y, G.g[1], G.mn[1], G.cell), gz = axis_gap(qz, z, G.g[2], G.mn[2], G.cell);
        const float dyz2 = gy * gy + gz * gz;
        if (!(dyz2 <= thr)) return false;
        // (the hardware square root: its ulp is far inside the margins)
        const float dx = __builtin_amdgcn_sqrtf(thr - dyz2 + 1e-3f * T);
        const float e = (fabsf(qx) + dx) * 1e-6f;
        int xl, xh;
        cell_bounds(qx - dx - e, qx + dx + e, G.mn[0], __builtin_amdgcn_rcpf(G.cell), G.g[0], xl, xh);
        xl = max(B.lo[0], xl);
        xh = min(B.hi[0], xh);
        if (xl > xh) return false;
        const uint32_t c0 = pmd::cell_index(G, xl, y, z);
        const uint32_t c1 = c0 + (uint32_t)(xh - xl);
        b = starts[c0];
        len = starts[c1] + sizes[c1] - b;
    } else {
        const int x = B.lo[0] + s % nx, t = s / nx;
        const int y = B.lo[1] + t % ny, z = B.lo[2] + t / ny;
        const float gx = axis_gap(qx, x, G.g[0], G.mn[0], G.cell), gy = axis_gap(qy, y, G.g[1], G.mn[1], G.cell),
                    gz = axis_gap(qz, z, G.g[2], G.mn[2], G.cell);
        if (!(gx * gx + gy * gy + gz * gz <= thr)) return false;
        const uint32_t c = pmd::cell_index(G, x, y, z);
        b = starts[c];
        len = sizes[c];
    }
    return true;
}

// Visit every candidate of the box, 64 per step, calling
// f(valid, slot, d2) on all lanes (valid = lane holds a candidate; d2 is
// computed on every lane). Runs of photons (rows) are held one per lane; a
// scalar cursor walks them (all control is wave-uniform, in SGPRs), and the
// next step's positions are loaded, unconditionally, before the current step
// is processed: a lane past its run reads a neighbouring photon, or 0 past
// the end of the block, and is masked by `valid`.
template <class F>
__device__ __forceinline__ void walk(const KnnParams& P, uint32_t lane, const Box& B, float qx, float qy, float qz,
                                     F&& f) {
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const bool linear = P.G.layout == PM_LAYOUT_LINEAR;
    const int nseg = linear ? ny * nz : nx * ny * nz;
    const __amdgpu_buffer_rsrc_t rx = plane_rsrc(P.px, P.n), ry = plane_rsrc(P.py, P.n), rz = plane_rsrc(P.pz, P.n);
    for (int sbase = 0; sbase < nseg; sbase += 64) {
        const int s = sbase + (int)lane;
        uint32_t b = 0, len = 0;
        if (s < nseg) row_span(P.G, P.starts, P.sizes, B, s, qx, qy, qz, P.T, b, len);
        const int ns = min(64, nseg - sbase);
        // Non-empty rows as a lane mask; the cursor visits them in order.
        uint64_t rows = wballot(len > 0) & (ns == 64 ? ~0ull : ((1ull << ns) - 1));
        if (!rows) continue;
        int cs = __builtin_ctzll(rows);
        rows &= rows - 1;
        uint32_t rb = __builtin_amdgcn_readlane(b, cs), rl = __builtin_amdgcn_readlane(len, cs), cc = 0;
        bool v0 = lane < rl;
        uint32_t o0 = (rb + lane) * 4u;
        float x0 = bload(rx, o0), y0 = bload(ry, o0), z0 = bload(rz, o0);
        for (;;) {
            // next step: same row if it has more, else the next non-empty row
            bool has_next = true;
            uint32_t ncc = cc + 64, nrb = rb, nrl = rl;
            if (ncc >= rl) {
                if (!rows) {
                    has_next = false;
                } else {
                    int ncs = __builtin_ctzll(rows);
                    rows &= rows - 1;
                    nrb = __builtin_amdgcn_readlane(b, ncs);
                    nrl = __builtin_amdgcn_readlane(len, ncs);
                    ncc = 0;
                }
            }
            // (after the last step this reads past the row once, harmlessly)
            const bool v1 = ncc + lane < nrl;
            const uint32_t o1 = (nrb + ncc + lane) * 4u;
            const float x1 = bload(rx, o1), y1 = bload(ry, o1), z1 = bload(rz, o1);
            f(v0, o0 >> 2, dist2(x0, y0, z0, qx, qy, qz));
            if (!has_next) break;
            cc = __builtin_amdgcn_readfirstlane(ncc);
            rb = __builtin_amdgcn_readfirstlane(nrb);
            rl = __builtin_amdgcn_readfirstlane(nrl);
            v0 = v1;
            o0 = o1;
            x0 = x1;
            y0 = y1;
            z0 = z1;
        }
    }
}

// walk with 128-candidate steps over the runs laid end to end: a step is
// the next 128 positions of the concatenated runs (lane l holds positions
// p + l and p + 64 + l), so runs shorter than a step share one instead of
// leaving lanes idle, and the candidates arrive in the same order as run by
// run. Position t of run j (its stream start pre_j, first slot b_j) is slot
// b_j + t - pre_j: each lane carries the byte offset 4 (b_j - pre_j) of the
// run holding its position, updated only for the runs that start inside the
// step. The cursor (step start, next run, its start) is wave-uniform, in
// SGPRs. f(am, sa, d2a, bm, sb, d2b): am / bm are the lane masks of the
// lanes that hold a candidate (wave-uniform, SGPRs); d2 is computed on every
// lane. Rows and their cells are pruned against Trow (T, or a smaller bound
// when only keys below it can be kept).
template <class F>
__device__ __forceinline__ void walk2(const KnnParams& P, uint32_t lane, const Box& B, float qx, float qy, float qz,
                                      float Trow, F&& f) {
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const bool linear = P.G.layout == PM_LAYOUT_LINEAR;
    const int nseg = linear ? ny * nz : nx * ny * nz;
    const __amdgpu_buffer_rsrc_t rx = plane_rsrc(P.px, P.n), ry = plane_rsrc(P.py, P.n), rz = plane_rsrc(P.pz, P.n);
    const int ilane = (int)lane;
    const uint32_t lane4 = lane * 4u;
    for (int sbase = 0; sbase < nseg; sbase += 64) {
        const int s = sbase + ilane;
        uint32_t b = 0, len = 0;
        if (s < nseg) row_span(P.G, P.starts, P.sizes, B, s, qx, qy, qz, Trow, b, len);
        const uint32_t incl = wave_incl_scan(len, lane);
        const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
        if (total == 0) continue;
        const uint32_t pre = incl - len;
        const uint32_t d4 = (b - pre) * 4u;  // mod 2^32: (d4 + 4 t) is 4 x slot
        uint64_t rows = wballot(len > 0);
        // the first non-empty run starts the stream at 0
        uint32_t nstart = 0, cur = 0;
        struct Step {
            int left;  // stream positions left from the step's first
            uint32_t oa, ob;
            float xa, ya, za, xb, yb, zb;
        };
        // offsets and positions of the step at p; the runs starting before
        // p + 128 are entered (lanes at or past a run's start take its offset)
        auto load = [&](Step& w, uint32_t p) {
            uint32_t da = cur, db = cur;
            while (rows && nstart < p + 128u) {
                const int j = __builtin_ctzll(rows);
                rows &= rows - 1;
                const uint32_t dj = __builtin_amdgcn_readlane(d4, j);
                const int t0 = (int)(nstart - p);
                da = ilane >= t0 ? dj : da;
                db = ilane >= t0 - 64 ? dj : db;
                cur = dj;
                nstart = rows ? __builtin_amdgcn_readlane(pre, __builtin_ctzll(rows)) : 0xffffffffu;
            }
            w.left = (int)(total - p);
            w.oa = da + lane4 + p * 4u;
            w.ob = db + lane4 + p * 4u + 256u;
            w.xa = bload(rx, w.oa), w.ya = bload(ry, w.oa), w.za = bload(rz, w.oa);
            w.xb = bload(rx, w.ob), w.yb = bload(ry, w.ob), w.zb = bload(rz, w.ob);
        };
        // Two register sets: each half of the loop body loads the next step
        // into one set while the other is processed, so nothing is copied.
        Step s0, s1;
        uint32_t p = 0;
        load(s0, 0);
        auto half = [&](const Step& c, Step& nxt) -> bool {
            const uint32_t pn = p + 128u;
            const bool more = pn < total;
            if (more) load(nxt, pn);
            // (each mask is one lane compare's ballot)
            f(wballot(ilane < c.left), c.oa >> 2, dist2(c.xa, c.ya, c.za, qx, qy, qz), wballot(ilane + 64 < c.left),
              c.ob >> 2, dist2(c.xb, c.yb, c.zb, qx, qy, qz));
            p = pn;
            return more;
        };
        while (half(s0, s1) && half(s1, s0)) {
        }
    }
}

// Radiance estimate of Raytracer.cuh:110-138 over the m kept photons held in
// v (sorted); lanes own elements j*64 + lane.
template <int R>
__device__ __forceinline__ void radiance_epilogue(const KnnParams& P, const uint64_t (&v)[R], uint32_t m,
                                                  uint32_t lane, uint32_t q) {
    const float alpha = 1.818f, beta = 1.953f;
    const float pi_f = 3.141592654f;
    float res0 = 0.f, res1 = 0.f, res2 = 0.f;
    float rmax = 0.f;
    if (m > 0) {
        const int jm = (int)((m - 1) / 64);
        float dm = 0.f;
#pragma unroll
        for (int j = 0; j < R; ++j)
            if (j == jm) dm = sqrtf(__uint_as_float((uint32_t)(v[j] >> 32)));
        rmax = __shfl(dm, (int)((m - 1) & 63), 64);
        const float filter_r = __fmul_rn(rmax, rmax);
        const float denom = 1.f - expf(-beta);
        const float nx = P.nrm[q], ny = P.nrm[P.nq + q], nz = P.nrm[2 * (size_t)P.nq + q];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            uint32_t e = (uint32_t)j * 64 + lane;
            if (e < m) {
                uint32_t idx = (uint32_t)v[j];
                float dx = P.dir[idx], dy = P.dir[P.n + idx], dz = P.dir[2 * (size_t)P.n + idx];
                float cosNL = __fadd_rn(__fadd_rn(__fmul_rn(nx, -dx), __fmul_rn(ny, -dy)), __fmul_rn(nz, -dz));
                if (cosNL > 0.f) {
                    float d = sqrtf(__uint_as_float((uint32_t)(v[j] >> 32)));
                    float x = __fdiv_rn(__fmul_rn(__fmul_rn(-beta, d), d), __fmul_rn(2.f, filter_r));
                    float w = __fmul_rn(alpha, 1.f - __fdiv_rn(1.f - expf(x), denom));
                    res0 += P.power[idx] * w;
                    res1 += P.power[P.n + idx] * w;
                    res2 += P.power[2 * (size_t)P.n + idx] * w;
                }
            }
        }
    }
    res0 = wave_sum_f32(res0);
    res1 = wave_sum_f32(res1);
    res2 = wave_sum_f32(res2);
    if (lane == 0) {
        float coef = 1.f;
        if (m > 0) coef = __fdiv_rn(1.f, __fmul_rn(__fmul_rn(pi_f, __fmul_rn(rmax, rmax)), 1.f));
        P.rgb[q] = m > 0 ? res0 * coef : 0.f;
        P.rgb[P.nq + q] = m > 0 ? res1 * coef : 0.f;
        P.rgb[2 * (size_t)P.nq + q] = m > 0 ? res2 * coef : 0.f;
    }
}

template <int R, int CAP>
__device__ __forceinline__ void sort_and_emit(const KnnParams& P, WaveLds<CAP>& L, uint32_t lane, uint32_t q,
                                              uint32_t cnt, uint32_t m) {
    uint64_t v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        uint32_t e = (uint32_t)j * 64 + lane;
        v[j] = e < cnt ? L.buf[e] : ~0ull;
    }
    bitonic_sort<R>(v, lane);
    if (P.out_idx) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            uint32_t e = (uint32_t)j * 64 + lane;
            if (e < (uint32_t)P.k) {
                bool keep = e < m;
                P.out_idx[(size_t)q * P.k + e] = keep ? (uint32_t)v[j] : 0xFFFFFFFFu;
                P.out_dist[(size_t)q * P.k + e] =
                    keep ? sqrtf(__uint_as_float((uint32_t)(v[j] >> 32))) : __builtin_inff();
            }
        }
        // slots beyond the sorted registers (k > 64*R >= cnt >= m) are empty
        for (uint32_t e = 64u * R + lane; e < (uint32_t)P.k; e += 64) {
            P.out_idx[(size_t)q * P.k + e] = 0xFFFFFFFFu;
            P.out_dist[(size_t)q * P.k + e] = __builtin_inff();
        }
    }
    if (P.out_count && lane == 0) P.out_count[q] = m;
    if (P.rgb) radiance_epilogue<R>(P, v, m, lane, q);
}

// Radiance estimate (Raytracer.cuh:110-138) over the m kept keys left in
// the LDS buffer, unsorted: the filter radius is their largest distance and
// the sum runs in buffer order (lane-strided, then a wave reduction).
//
// The filter weight w = alpha (1 - (1 - e^{-beta d^2 / (2 r^2)}) / (1 - e^{-beta}))
// is evaluated as A + B e^{x} with x = d^2 (-beta / (2 r^2)) from the squared
// distance itself (no square root), the hardware exponential, and the
// filter radius and the per-query factors (-beta / (2 r^2), 1 / (pi r^2))
// from the hardware square root and reciprocal; against the reference's FP32
// sequence of square roots, divisions and expf this moves the estimate by a
// few 1e-7 relative, inside its 1e-5 bound. The cos > 0 test keeps the
// reference's rounding.
template <int CAP>
__device__ __forceinline__ void radiance_buffer(const KnnParams& P, const WaveLds<CAP>& L, uint32_t lane, uint32_t q,
                                                uint32_t m) {
    const float alpha = 1.818f, beta = 1.953f;
    const float pi_f = 3.141592654f;
    // 1 / (1 - e^{-beta}) and log2(e), rounded to FP32
    const float inv_denom = 1.16529458f, log2e = 1.44269504f;
    float dmax = 0.f;
    for (uint32_t e = lane; e < m; e += 64) dmax = fmaxf(dmax, __uint_as_float((uint32_t)(L.buf[e] >> 32)));
    dmax = wave_max_nonneg_f32(dmax);
    // hardware square root and reciprocals (a few 1e-7 relative on the
    // estimate, inside its 1e-5 bound)
    const float rmax = __builtin_amdgcn_sqrtf(dmax);
    float res0 = 0.f, res1 = 0.f, res2 = 0.f;
    if (m > 0) {
        const float filter_r = __fmul_rn(rmax, rmax);
        // x log2(e) = d^2 * s
        const float s = -beta * __builtin_amdgcn_rcpf(2.f * filter_r) * log2e;
        const float wa = alpha * (1.f - inv_denom), wb = alpha * inv_denom;
        const float nx = P.nrm[q], ny = P.nrm[P.nq + q], nz = P.nrm[2 * (size_t)P.nq + q];
        for (uint32_t e = lane; e < m; e += 64) {
            const uint64_t v = L.buf[e];
            const uint32_t idx = (uint32_t)v;
            // power is read with the direction (one memory round trip),
            // and used only where the photon faces the surface
            const float dx = P.dir[idx], dy = P.dir[P.n + idx], dz = P.dir[2 * (size_t)P.n + idx];
            const float p0 = P.power[idx], p1 = P.power[P.n + idx], p2 = P.power[2 * (size_t)P.n + idx];
            const float cosNL = __fadd_rn(__fadd_rn(__fmul_rn(nx, -dx), __fmul_rn(ny, -dy)), __fmul_rn(nz, -dz));
            if (cosNL > 0.f) {
                const float w = wa + wb * __builtin_amdgcn_exp2f(__uint_as_float((uint32_t)(v >> 32)) * s);
                res0 += p0 * w;
                res1 += p1 * w;
                res2 += p2 * w;
            }
        }
    }
    res0 = wave_sum_f32(res0);
    res1 = wave_sum_f32(res1);
    res2 = wave_sum_f32(res2);
    if (lane == 0) {
        float coef = 1.f;
        if (m > 0) coef = __builtin_amdgcn_rcpf(__fmul_rn(pi_f, __fmul_rn(rmax, rmax)));
        P.rgb[q] = m > 0 ? res0 * coef : 0.f;
        P.rgb[P.nq + q] = m > 0 ? res1 * coef : 0.f;
        P.rgb[2 * (size_t)P.nq + q] = m > 0 ? res2 * coef : 0.f;
    }
}

template <int CAP>
__device__ __forceinline__ void finish(const KnnParams& P, WaveLds<CAP>& L, uint32_t lane, uint32_t q, uint32_t cnt,
                                       uint32_t found, uint64_t khi) {
    const uint32_t k = (uint32_t)P.k;
    const uint32_t m = found < k ? found : k;
    if (!P.out_idx) {
        // no neighbour lists: narrow to exactly the k smallest and estimate
        // from the buffer without sorting (up to 256 keys: selected from
        // registers, the keys read once)
        if (cnt <= 256) {
            uint64_t v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t e = (uint32_t)j * 64 + lane;
                v[j] = e < cnt ? L.buf[e] : ~0ull;
            }
            cnt = reg_select<4, CAP>(L, lane, v, cnt, k, P.bT, khi);
        } else {
            cnt = buf_select<CAP>(L, lane, cnt, k, k, P.bT, khi);
        }
        wave_sync();
        if (P.out_count && lane == 0) P.out_count[q] = m;
        if (P.rgb) radiance_buffer<CAP>(P, L, lane, q, m);
        return;
    }
    // narrowed to the smallest sort that holds k keys (64, 128 or 256)
    cnt = buf_select<CAP>(L, lane, cnt, k, k <= 64 ? 64u : k <= 128 ? 128u : 256u, P.bT, khi);
    wave_sync();
    if (cnt <= 64)
        sort_and_emit<1, CAP>(P, L, lane, q, cnt, m);
    else if (cnt <= 128)
        sort_and_emit<2, CAP>(P, L, lane, q, cnt, m);
    else
        sort_and_emit<4, CAP>(P, L, lane, q, cnt, m);
}

// The candidate box of query q: its cell +- rc, clamped (PhotonGrid.cu:262-272).
__device__ __forceinline__ void box_of(const KnnParams& P, uint32_t q, Box& B, float& qx, float& qy, float& qz,
                                       int* cell = nullptr) {
    // the query is the wave's: its coordinates (and so the box, the row
    // count and the cursor of walk) are wave-uniform, held in SGPRs
    qx = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.qx[q])));
    qy = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.qy[q])));
    qz = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.qz[q])));
    const float qv[3] = {qx, qy, qz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int c = pmd::cell_coord(qv[a], P.G.mn[a], P.G.cell, P.G.g[a]);
        B.lo[a] = max(0, c - P.rc);
        B.hi[a] = min(P.G.g[a] - 1, c + P.rc);
        if (cell) cell[a] = c;
    }
}

// XCD-chunked order: XCD x takes chunks x, x + 8, ... of kXcdChunk
// consecutive workgroups (the grid is rounded up to a multiple of 8 chunks).
constexpr uint32_t kXcdChunk = 256;
__device__ __forceinline__ uint32_t xcd_chunked(uint32_t b) {
    const uint32_t x = b % 8, i = b / 8;
    return ((i / kXcdChunk) * 8 + x) * kXcdChunk + (i % kXcdChunk);
}

// Query of this wave in a one-wave-per-query launch (XCD-aware block order,
// optional processing order); false past the end (wave-uniform).
__device__ __forceinline__ bool query_of_wave(const KnnParams& P, uint32_t& q) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t blk = P.remap == 2 ? xcd_chunked(blockIdx.x)
                         : P.remap   ? pmd::xcd_remap(blockIdx.x, gridDim.x)
                                     : blockIdx.x;
    const uint32_t ql = blk * kWaves + wave;
    if (ql >= P.nq || (P.nq_dev && ql >= *P.nq_dev)) return false;
    q = __builtin_amdgcn_readfirstlane(P.order ? P.order[ql] : ql);
    return true;
}

// One step of the streaming selection over two candidates per lane (index
// sa / sb, squared distance d2a / d2b, lanes holding one: masks am / bm):
// appends the keys below the acceptance bound khi to the LDS buffer, and
// narrows the buffer first when the step would overflow it. khi never
// exceeds (bits(T) + 1) << 32, so a key below it is within the radius (d2
// >= 0: its bits order as its value; a NaN d2 is above every bound). Every
// predicate is a ballot of a plain compare combined in SGPRs; the writing
// lanes come back from their mask.
template <int CAP>
__device__ __forceinline__ void push2(WaveLds<CAP>& L, uint32_t lane, uint32_t k, uint32_t bT, uint64_t am,
                                      uint32_t sa, float d2a, uint64_t bm, uint32_t sb, float d2b, uint32_t& cnt,
                                      uint64_t& khi) {
    const uint64_t ka = make_key(d2a, sa), kb = make_key(d2b, sb);
    uint64_t bala = am & wballot(ka < khi), balb = bm & wballot(kb < khi);
    uint32_t na = (uint32_t)__popcll(bala), nb = (uint32_t)__popcll(balb);
    if (cnt + na + nb > (uint32_t)CAP) {
        // leaves at most CAP / 2 keys, room for this step's 128
        wave_sync();
        cnt = buf_select<CAP>(L, lane, cnt, k, CAP / 2, bT, khi);
        bala = am & wballot(ka < khi);
        balb = bm & wballot(kb < khi);
        na = (uint32_t)__popcll(bala);
        nb = (uint32_t)__popcll(balb);
    }
    if (__builtin_amdgcn_inverse_ballot_w64(bala)) L.buf[cnt + mbcnt(bala)] = ka;
    if (__builtin_amdgcn_inverse_ballot_w64(balb)) L.buf[cnt + na + mbcnt(balb)] = kb;
    cnt += na + nb;
}

// Per-query set-up of the streaming gather: the query's position (x, y,
// z: set by the caller, wave-uniform), its cell, and with the inner box the
// d2 bits below which a key is closer than every photon outside it (less a
// few ulps of rounding: every such photon has d2 >= gap^2). The three axes
// are worked side by side in lanes 0..2 (the cell hash's division and the
// inner box's two face gaps, each axis by the same formulas as one at a
// time) and read back; the gap is the exact minimum over the axes.
struct QSetup {
    uint32_t q;
    float x, y, z;
    int c[3];
    uint32_t gbits;
};
__device__ __forceinline__ void query_setup(const KnnParams& P, QSetup& S, uint32_t lane) {
    // (the query's coordinates and the per-axis grid values are taken as
    // wave-uniform values first, so the lane selects below stay selects and
    // never become an indexed load from memory)
    const float mn0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.G.mn[0])));
    const float mn1 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.G.mn[1])));
    const float mn2 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.G.mn[2])));
    const int g0 = __builtin_amdgcn_readfirstlane(P.G.g[0]), g1 = __builtin_amdgcn_readfirstlane(P.G.g[1]),
              g2 = __builtin_amdgcn_readfirstlane(P.G.g[2]);
    const float x = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(S.x)));
    const float y = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(S.y)));
    const float z = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(S.z)));
    const int a = lane < 3u ? (int)lane : 0;
    const float v = a == 0 ? x : a == 1 ? y : z;
    const float mn = a == 0 ? mn0 : a == 1 ? mn1 : mn2;
    const int g = a == 0 ? g0 : a == 1 ? g1 : g2;
    const int c = pmd::cell_coord(v, mn, P.G.cell, g);
#pragma unroll
    for (int i = 0; i < 3; ++i) S.c[i] = __builtin_amdgcn_readlane(c, i);
    S.gbits = 0u;
    if (P.h > 0) {
        const int lo = max(0, c - P.h), hi = min(g - 1, c + P.h);
        const uint32_t gb = __float_as_uint(inner_gap(v, lo, hi, g, mn, P.G.cell));
        const float gap = fminf(fminf(__uint_as_float(__builtin_amdgcn_readlane(gb, 0)),
                                      __uint_as_float(__builtin_amdgcn_readlane(gb, 1))),
                                __uint_as_float(__builtin_amdgcn_readlane(gb, 2)));
        S.gbits = min(__float_as_uint(gap * gap * 0.999f), __float_as_uint(P.T) + 1u);
    }
}

// Single pass: in-radius keys stream into an LDS buffer of CAP keys; when it
// would overflow, buf_select keeps only the smallest bin range holding >= k
// keys and tightens the acceptance bound. Exact: a dropped key has >= k kept
// keys below it.
//
// With P.h > 0 the walk first covers only the inner box (the query's cell
// +- h). If that finds at least k keys within the radius and the k-th
// smallest of them is closer than every photon outside the inner box can
// be, those k keys are the answer (m = k); otherwise the query walks the
// whole box again from an empty buffer. Both outcomes are exact.
// The finish's parameters, re-read where P came from (scalar loads) instead
// of kept in SGPRs across the walk: kFrom = kFromArgs when P is the launch's
// own kernel argument (at offset 0 of the argument segment), kFromRecord
// when P was loaded from the record src; kFromNone uses P.
constexpr int kFromNone = 0, kFromArgs = 1, kFromRecord = 2;
template <class Ptr>
__device__ __forceinline__ void finish_params(KnnParams& Pf, Ptr kp) {
    // (a pointer the compiler cannot see through, so the loads stay here;
    // every field the finish reads is taken from it)
    asm volatile("" : "+s"(kp));
    Pf.k = kp->k;
    Pf.n = kp->n;
    Pf.nq = kp->nq;
    Pf.bT = kp->bT;
    Pf.power = kp->power;
    Pf.dir = kp->dir;
    Pf.nrm = kp->nrm;
    Pf.rgb = kp->rgb;
    Pf.out_idx = kp->out_idx;
    Pf.out_dist = kp->out_dist;
    Pf.out_count = kp->out_count;
}

template <int CAP, int kFrom = kFromNone>
__device__ __forceinline__ void stream_core(const KnnParams& P, WaveLds<CAP>& L, uint32_t lane, const QSetup& S,
                                            const KnnParams* src = nullptr) {
    const uint32_t q = S.q;
    const float qx = S.x, qy = S.y, qz = S.z;
    const float T = P.T;
    const uint32_t k = (uint32_t)P.k;
    // the candidate box: the query's cell +- rc, clamped (PhotonGrid.cu:262-272)
    Box B;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        B.lo[a] = max(0, S.c[a] - P.rc);
        B.hi[a] = min(P.G.g[a] - 1, S.c[a] + P.rc);
    }
    // the inner box (the query's cell +- h) first when asked for
    Box W = B;
    const uint32_t gbits = S.gbits;
    bool inner = P.h > 0;
    if (inner) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            W.lo[a] = max(0, S.c[a] - P.h);
            W.hi[a] = min(P.G.g[a] - 1, S.c[a] + P.h);
        }
    }
    uint32_t cnt;
    uint64_t khi;
    for (;;) {
        // the acceptance bound starts at the radius (d2 <= T), and in the
        // inner box at gbits: only keys closer than every photon outside it
        // are buffered there
        cnt = 0;
        khi = inner ? (uint64_t)gbits << 32 : (uint64_t)(__float_as_uint(T) + 1u) << 32;
        // (the inner walk keeps only keys below gbits: its rows are pruned
        // against that bound)
        const float Trow = inner ? fminf(T, __uint_as_float(gbits)) : T;
        walk2(P, lane, W, qx, qy, qz, Trow,
              [&](uint64_t am, uint32_t sa, float d2a, uint64_t bm, uint32_t sb, float d2b) {
                  push2<CAP>(L, lane, k, P.bT, am, sa, d2a, bm, sb, d2b, cnt, khi);
              });
        wave_sync();
        // with at least k such keys, the k smallest keys overall are among
        // them (the buffer holds every in-radius key below its bound, or at
        // least k of them once narrowed)
        if (!inner || cnt >= k) break;
        // otherwise the whole box from an empty buffer
        W = B;
        inner = false;
        wave_sync();
    }
    // The count within the radius that the finish needs is min(found, k),
    // and that is min(cnt, k): without a narrowing the buffer holds every
    // key within the radius (cnt = found), and a narrowing keeps at least k
    // (it starts above CAP - 128 >= k keys), so cnt >= k after one.
    const uint32_t found = cnt;
    if constexpr (kFrom == kFromArgs) {
        KnnParams Pf = P;
        finish_params(Pf, (const __attribute__((address_space(4))) KnnParams*)__builtin_amdgcn_kernarg_segment_ptr());
        finish<CAP>(Pf, L, lane, q, cnt, found, khi);
    } else if constexpr (kFrom == kFromRecord) {
        KnnParams Pf = P;
        finish_params(Pf, src);
        finish<CAP>(Pf, L, lane, q, cnt, found, khi);
    } else {
        finish<CAP>(P, L, lane, q, cnt, found, khi);
    }
}

// One query on the wave (its set-up wave-uniform).
template <int CAP, int kFrom = kFromNone>
__device__ __forceinline__ void stream_one(const KnnParams& P, WaveLds<CAP>& L, uint32_t lane, uint32_t q,
                                           const KnnParams* src = nullptr) {
    // the query is the wave's: its coordinates (and so the box, the row
    // count and the cursor of walk) are wave-uniform, held in SGPRs
    QSetup S;
    S.q = q;
    S.x = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.qx[q])));
    S.y = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.qy[q])));
    S.z = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.qz[q])));
    query_setup(P, S, lane);
    stream_core<CAP, kFrom>(P, L, lane, S, src);
}

// 512-key buffer: 17 KB of LDS per workgroup, so 7 waves per SIMD fit when
// the registers do (68 VGPRs under this occupancy target, no spills).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_knn_stream512(KnnParams P) {
    __shared__ WaveLds<512> lds[kWaves];
    uint32_t q;
    if (!query_of_wave(P, q)) return;  // wave-uniform; no workgroup barriers below
    stream_one<512, kFromArgs>(P, lds[threadIdx.x >> 6], threadIdx.x & 63, q);
}

// 1024-key buffer: fewer refinements per query, LDS-limited to 4 waves per SIMD.
__global__ __launch_bounds__(256) void k_knn_stream1024(KnnParams P) {
    __shared__ WaveLds<1024> lds[kWaves];
    uint32_t q;
    if (!query_of_wave(P, q)) return;
    stream_one<1024, kFromArgs>(P, lds[threadIdx.x >> 6], threadIdx.x & 63, q);
}

// Large k (257..2048, e.g. the shipped renderer's k = 2000 / 400): one wave
// per workgroup with a CAP-key LDS buffer (CAP >= 2k). The stream is the same;
// the final selection narrows the buffer to exactly the k smallest keys,
// which a bitonic network sorts in LDS.
template <int CAP>
__device__ __forceinline__ void large_one(const KnnParams& P, WaveLds<CAP>& L, uint32_t lane, uint32_t q) {
    Box B;
    float qx, qy, qz;
    box_of(P, q, B, qx, qy, qz);
    const float T = P.T;
    const uint32_t k = (uint32_t)P.k;
    uint32_t cnt = 0, found = 0;
    uint64_t khi = ~0ull;
    walk(P, lane, B, qx, qy, qz, [&](bool valid, uint32_t slot, float d2) {
        // each ballot is taken of a plain compare (no bool round trip): an
        // invalid lane gets a threshold no d2 >= 0 (or NaN) passes
        const bool in = d2 <= (valid ? T : -1.0f);
        const uint64_t bin = wballot(in);
        found += (uint32_t)__popcll(bin);
        uint64_t key = make_key(d2, slot);
        const bool below = key < khi;
        bool acc = in && below;
        uint64_t bal = bin & wballot(below);
        uint32_t na = (uint32_t)__popcll(bal);
        if (cnt + na > (uint32_t)CAP) {
            wave_sync();
            cnt = buf_select<CAP>(L, lane, cnt, k, CAP / 2, P.bT, khi);
            acc = acc && key < khi;
            bal = wballot(acc);
            na = (uint32_t)__popcll(bal);
        }
        if (acc) L.buf[cnt + mbcnt(bal)] = key;
        cnt += na;
    });
    wave_sync();
    const uint32_t m = found < k ? found : k;
    cnt = buf_select<CAP>(L, lane, cnt, k, k, P.bT, khi);  // exactly min(k, found) keys remain
    wave_sync();
    if (!P.out_idx) {  // radiance / counts only: no sort needed
        if (P.out_count && lane == 0) P.out_count[q] = m;
        if (P.rgb) radiance_buffer<CAP>(P, L, lane, q, m);
        return;
    }
    uint32_t n2 = 64;
    while (n2 < cnt) n2 <<= 1;
    for (uint32_t e = cnt + lane; e < n2; e += 64) L.buf[e] = ~0ull;
    wave_sync();
    lds_bitonic_sort(L.buf, n2, lane);
    if (P.out_idx) {
        for (uint32_t e = lane; e < k; e += 64) {
            const bool keep = e < m;
            const uint64_t v = keep ? L.buf[e] : 0ull;
            P.out_idx[(size_t)q * P.k + e] = keep ? (uint32_t)v : 0xFFFFFFFFu;
            P.out_dist[(size_t)q * P.k + e] = keep ? sqrtf(__uint_as_float((uint32_t)(v >> 32))) : __builtin_inff();
        }
    }
    if (P.out_count && lane == 0) P.out_count[q] = m;
    if (P.rgb) {
        // Raytracer.cuh:110-138 over the m kept photons
        const float alpha = 1.818f, beta = 1.953f;
        const float pi_f = 3.141592654f;
        float res0 = 0.f, res1 = 0.f, res2 = 0.f, rmax = 0.f;
        if (m > 0) {
            rmax = sqrtf(__uint_as_float((uint32_t)(L.buf[m - 1] >> 32)));
            const float filter_r = __fmul_rn(rmax, rmax);
            const float denom = 1.f - expf(-beta);
            const float nx = P.nrm[q], ny = P.nrm[P.nq + q], nz = P.nrm[2 * (size_t)P.nq + q];
            for (uint32_t e = lane; e < m; e += 64) {
                const uint64_t v = L.buf[e];
                const uint32_t idx = (uint32_t)v;
                float dx = P.dir[idx], dy = P.dir[P.n + idx], dz = P.dir[2 * (size_t)P.n + idx];
                float cosNL = __fadd_rn(__fadd_rn(__fmul_rn(nx, -dx), __fmul_rn(ny, -dy)), __fmul_rn(nz, -dz));
                if (cosNL > 0.f) {
                    float d = sqrtf(__uint_as_float((uint32_t)(v >> 32)));
                    float x = __fdiv_rn(__fmul_rn(__fmul_rn(-beta, d), d), __fmul_rn(2.f, filter_r));
                    float w = __fmul_rn(alpha, 1.f - __fdiv_rn(1.f - expf(x), denom));
                    res0 += P.power[idx] * w;
                    res1 += P.power[P.n + idx] * w;
                    res2 += P.power[2 * (size_t)P.n + idx] * w;
                }
            }
        }
        res0 = wave_sum_f32(res0);
        res1 = wave_sum_f32(res1);
        res2 = wave_sum_f32(res2);
        if (lane == 0) {
            float coef = 1.f;
            if (m > 0) coef = __fdiv_rn(1.f, __fmul_rn(__fmul_rn(pi_f, __fmul_rn(rmax, rmax)), 1.f));
            P.rgb[q] = m > 0 ? res0 * coef : 0.f;
            P.rgb[P.nq + q] = m > 0 ? res1 * coef : 0.f;
            P.rgb[2 * (size_t)P.nq + q] = m > 0 ? res2 * coef : 0.f;
        }
    }
}

template <int CAP>
__global__ __launch_bounds__(64) void k_knn_large(KnnParams P) {
    __shared__ WaveLds<CAP> L;
    const uint32_t blk = P.remap ? pmd::xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    if (blk >= P.nq) return;
    large_one<CAP>(P, L, threadIdx.x, P.order ? P.order[blk] : blk);
}

// The queries a k_gather_wide launch handed over (list[0, *count)), one wave
// per workgroup over a fixed grid.
template <int CAP>
__global__ __launch_bounds__(64) void k_knn_large_list(KnnParams P, const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ count) {
    __shared__ WaveLds<CAP> L;
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        large_one<CAP>(P, L, threadIdx.x, list[i]);
        wave_sync();
    }
}

// Radiance for 256 < k <= 2048 without a k-key buffer (the kernel behind the
// reference frame's k = 2000 / 400 gathers). One wave per query, three walks
// of the candidate box:
//   1. count the photons within r and histogram their d2 into 256 bins (8
//      per octave below the radius threshold, ascending in d2), keeping the
//      largest d2;
//   2. only if more than k are within r: collect the keys (d2, slot) of the
//      bin holding the k-th smallest into a 512-key LDS buffer and sort it in
//      registers; the (k - below)-th key there is the k-th key K* overall;
//   3. sum the estimate over the photons within r with key <= K* (all of them
//      when at most k are within r), the filter radius being the largest
//      kept distance.
// The kept set is exactly the k smallest (d2, slot) keys, as in the buffered
// kernels; only the order of the radiance sum differs (walk order). A query
// whose crossing bin holds more than 512 keys is appended to `ovf_list` for
// k_knn_large_list.
constexpr uint32_t kWideCap = 512;
struct WideLds {
    union {
        uint32_t hist[256];
        uint64_t buf[kWideCap];
    };
};

__device__ __forceinline__ uint32_t wide_bin(float d2, uint32_t bT20) {
    // d2 <= T, so bits(d2) >> 20 <= bT20; d2 <= T / 2^32 (and 0) share bin 0
    const uint32_t h = bT20 - (__float_as_uint(d2) >> 20);
    return 255u - (h < 255u ? h : 255u);
}

__global__ __launch_bounds__(256) void k_gather_wide(KnnParams P, uint32_t* __restrict__ ovf_list,
                                                     uint32_t* __restrict__ ovf_count) {
    __shared__ WideLds lds[kWaves];
    uint32_t q;
    if (!query_of_wave(P, q)) return;  // wave-uniform; no workgroup barriers below
    const uint32_t lane = threadIdx.x & 63;
    WideLds& L = lds[threadIdx.x >> 6];
    Box B;
    float qx, qy, qz;
    box_of(P, q, B, qx, qy, qz);
    const float T = P.T;
    const uint32_t k = (uint32_t)P.k;
    const uint32_t bT20 = __float_as_uint(T) >> 20;

    for (uint32_t e = lane; e < 256; e += 64) L.hist[e] = 0;
    wave_sync();
    uint32_t found = 0;
    float dmax = 0.f;
    walk2(P, lane, B, qx, qy, qz, T, [&](uint64_t am, uint32_t, float d2a, uint64_t bm, uint32_t, float d2b) {
        const bool va = __builtin_amdgcn_inverse_ballot_w64(am), vb = __builtin_amdgcn_inverse_ballot_w64(bm);
        const bool ina = d2a <= (va ? T : -1.0f), inb = d2b <= (vb ? T : -1.0f);
        found += (uint32_t)(__popcll(wballot(ina)) + __popcll(wballot(inb)));
        if (ina) {
            atomicAdd(&L.hist[wide_bin(d2a, bT20)], 1u);
            dmax = fmaxf(dmax, d2a);
        }
        if (inb) {
            atomicAdd(&L.hist[wide_bin(d2b, bT20)], 1u);
            dmax = fmaxf(dmax, d2b);
        }
    });
    wave_sync();
    found = __builtin_amdgcn_readfirstlane(found);
    const uint32_t m = found < k ? found : k;
    uint64_t kstar = ~0ull;
    float rmax2;
    if (found > k) {
        // lane l holds bins 4l .. 4l+3; the crossing bin is the first whose
        // inclusive count from the smallest d2 reaches k
        const uint32_t h0 = L.hist[4 * lane], h1 = L.hist[4 * lane + 1], h2 = L.hist[4 * lane + 2],
                       h3 = L.hist[4 * lane + 3];
        const uint32_t incl = wave_incl_scan(h0 + h1 + h2 + h3, lane);
        const int ls = __builtin_ctzll(wballot(incl >= k));
        uint32_t below = __shfl(incl - (h0 + h1 + h2 + h3), ls, 64);
        const uint32_t c0 = __shfl(h0, ls, 64), c1 = __shfl(h1, ls, 64), c2 = __shfl(h2, ls, 64);
        uint32_t bstar = 4u * (uint32_t)ls, nb = c0;
        if (below + c0 < k) {
            below += c0;
            ++bstar;
            nb = c1;
            if (below + c1 < k) {
                below += c1;
                ++bstar;
                nb = c2;
                if (below + c2 < k) {
                    below += c2;
                    ++bstar;
                    nb = __shfl(h3, ls, 64);
                }
            }
        }
        bstar = __builtin_amdgcn_readfirstlane(bstar);
        below = __builtin_amdgcn_readfirstlane(below);
        nb = __builtin_amdgcn_readfirstlane(nb);
        wave_sync();  // the histogram is read before the buffer (same LDS) is written
        if (nb > kWideCap) {
            if (lane == 0) ovf_list[atomicAdd(ovf_count, 1u)] = q;
            return;
        }
        uint32_t cnt = 0;
        walk2(P, lane, B, qx, qy, qz, T, [&](uint64_t am, uint32_t sa, float d2a, uint64_t bm, uint32_t sb, float d2b) {
            const bool va = __builtin_amdgcn_inverse_ballot_w64(am), vb = __builtin_amdgcn_inverse_ballot_w64(bm);
            const bool ina = d2a <= (va ? T : -1.0f) && wide_bin(d2a, bT20) == bstar;
            const bool inb = d2b <= (vb ? T : -1.0f) && wide_bin(d2b, bT20) == bstar;
            const uint64_t ba = wballot(ina), bb = wballot(inb);
            const uint32_t na = (uint32_t)__popcll(ba);
            if (ina) L.buf[cnt + mbcnt(ba)] = make_key(d2a, sa);
            if (inb) L.buf[cnt + na + mbcnt(bb)] = make_key(d2b, sb);
            cnt += na + (uint32_t)__popcll(bb);
        });
        wave_sync();
        uint64_t v[kWideCap / 64];
#pragma unroll
        for (int j = 0; j < (int)(kWideCap / 64); ++j) {
            const uint32_t e = (uint32_t)j * 64 + lane;
            v[j] = e < cnt ? L.buf[e] : ~0ull;
        }
        bitonic_sort<kWideCap / 64>(v, lane);
        const uint32_t r = k - below - 1;  // rank of K* in the bin (< nb)
        const int jr = (int)(r >> 6);
        uint64_t pick = 0;
#pragma unroll
        for (int j = 0; j < (int)(kWideCap / 64); ++j)
            if (j == jr) pick = v[j];
        const uint32_t lo = __shfl((uint32_t)pick, (int)(r & 63), 64);
        const uint32_t hi = __shfl((uint32_t)(pick >> 32), (int)(r & 63), 64);
        kstar = ((uint64_t)hi << 32) | lo;
        rmax2 = __uint_as_float(hi);
    } else {
        dmax = wave_max_nonneg_f32(dmax);
        rmax2 = dmax;
    }
    if (P.out_count && lane == 0) P.out_count[q] = m;

    // Raytracer.cuh:110-138 over the kept photons. Their keys queue in the
    // wave's LDS buffer (free again); every 64 queued, one lane per photon
    // loads its direction and power, and a batch's loads are consumed only
    // when the next batch is taken (or after the walk), some walk steps
    // later, so the walk never waits on them.
    const float alpha = 1.818f, beta = 1.953f;
    const float pi_f = 3.141592654f;
    float res0 = 0.f, res1 = 0.f, res2 = 0.f;
    const float rmax = sqrtf(rmax2);
    if (m > 0) {
        const float filter_r = __fmul_rn(rmax, rmax);
        const float denom = 1.f - expf(-beta);
        const float nx = P.nrm[q], ny = P.nrm[P.nq + q], nz = P.nrm[2 * (size_t)P.nq + q];
        const __amdgpu_buffer_rsrc_t rdx = plane_rsrc(P.dir, P.n), rdy = plane_rsrc(P.dir + P.n, P.n),
                                     rdz = plane_rsrc(P.dir + 2 * (size_t)P.n, P.n);
        const __amdgpu_buffer_rsrc_t rpr = plane_rsrc(P.power, P.n), rpg = plane_rsrc(P.power + P.n, P.n),
                                     rpb = plane_rsrc(P.power + 2 * (size_t)P.n, P.n);
        bool bv = false;
        float bd2 = 0.f, bx = 0.f, by = 0.f, bz = 0.f, br = 0.f, bg = 0.f, bb = 0.f;
        auto consume = [&]() {
            if (bv) {
                const float cosNL = __fadd_rn(__fadd_rn(__fmul_rn(nx, -bx), __fmul_rn(ny, -by)), __fmul_rn(nz, -bz));
                if (cosNL > 0.f) {
                    const float d = sqrtf(bd2);
                    const float x = __fdiv_rn(__fmul_rn(__fmul_rn(-beta, d), d), __fmul_rn(2.f, filter_r));
                    const float w = __fmul_rn(alpha, 1.f - __fdiv_rn(1.f - expf(x), denom));
                    res0 += br * w;
                    res1 += bg * w;
                    res2 += bb * w;
                }
            }
        };
        auto take = [&](uint32_t avail) {  // lanes < avail take queue entries 0..avail-1
            bv = lane < avail;
            const uint64_t key = bv ? L.buf[lane] : 0ull;
            const uint32_t off = (uint32_t)key * 4u;
            bd2 = __uint_as_float((uint32_t)(key >> 32));
            if (bv) {
                bx = bload(rdx, off);
                by = bload(rdy, off);
                bz = bload(rdz, off);
                br = bload(rpr, off);
                bg = bload(rpg, off);
                bb = bload(rpb, off);
            }
        };
        uint32_t qn = 0;
        wave_sync();  // the buffer's previous contents are dead
        walk2(P, lane, B, qx, qy, qz, T, [&](uint64_t am, uint32_t sa, float d2a, uint64_t bm, uint32_t sb, float d2b) {
            const bool va = __builtin_amdgcn_inverse_ballot_w64(am), vb = __builtin_amdgcn_inverse_ballot_w64(bm);
            const bool sela = d2a <= (va ? T : -1.0f) && make_key(d2a, sa) <= kstar;
            const bool selb = d2b <= (vb ? T : -1.0f) && make_key(d2b, sb) <= kstar;
            const uint64_t ba = wballot(sela), bbal = wballot(selb);
            const uint32_t na = (uint32_t)__popcll(ba);
            if (sela) L.buf[qn + mbcnt(ba)] = make_key(d2a, sa);
            if (selb) L.buf[qn + na + mbcnt(bbal)] = make_key(d2b, sb);
            qn += na + (uint32_t)__popcll(bbal);  // < 64 + 128: the queue holds < 64 between steps
            while (qn >= 64) {
                wave_sync();
                consume();
                take(64);
                // entries 64 .. qn-1 (fewer than 128) move to the front
                const uint64_t r0 = lane + 64 < qn ? L.buf[64 + lane] : 0ull;
                const uint64_t r1 = lane + 128 < qn ? L.buf[128 + lane] : 0ull;
                wave_sync();
                if (lane + 64 < qn) L.buf[lane] = r0;
                if (lane + 128 < qn) L.buf[64 + lane] = r1;
                qn -= 64;
                wave_sync();
            }
        });
        wave_sync();
        consume();
        take(qn);
        consume();
    }
    res0 = wave_sum_f32(res0);
    res1 = wave_sum_f32(res1);
    res2 = wave_sum_f32(res2);
    if (lane == 0) {
        float coef = 1.f;
        if (m > 0) coef = __fdiv_rn(1.f, __fmul_rn(__fmul_rn(pi_f, __fmul_rn(rmax, rmax)), 1.f));
        P.rgb[q] = m > 0 ? res0 * coef : 0.f;
        P.rgb[P.nq + q] = m > 0 ? res1 * coef : 0.f;
        P.rgb[2 * (size_t)P.nq + q] = m > 0 ? res2 * coef : 0.f;
    }
}

// Batched forms: the queries of many jobs (grids, radii, k) in one launch.
// A position p maps to a flattened query index g (through the batch's
// processing order, if any), whose job is the last one with qbase <= g.
__device__ __forceinline__ uint32_t job_of(const uint32_t* __restrict__ qbase, uint32_t n_jobs, uint32_t g) {
    uint32_t lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (qbase[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// One wave per position. Workgroups go to the 8 XCDs round-robin; the
// linear workgroup id is remapped so that XCD x takes chunks x, x + 8, ...
// of kXcdChunk consecutive workgroups: each XCD's L2 still sees runs of
// neighbouring cells, while at any moment all XCDs work inside the same
// region of the (job, cell) order and share the photon blocks in the MALL
// (a contiguous eighth per XCD would put eight different jobs in flight).
// The 2-D grid keeps every dimension's work-item count below 2^32; the
// workgroups past the rounded-up count exit at once.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_knn_batch_grid(
    const KnnParams* __restrict__ jobs, const uint32_t* __restrict__ qbase, uint32_t n_jobs,
    const uint32_t* __restrict__ order, uint32_t total, uint32_t n_blocks) {
    __shared__ WaveLds<512> lds[kWaves];
    const uint32_t b = blockIdx.y * gridDim.x + blockIdx.x;
    if (b >= n_blocks) return;  // n_blocks is a multiple of 8 * kXcdChunk
    const uint32_t p = xcd_chunked(b) * kWaves + (threadIdx.x >> 6);
    if (p >= total) return;  // wave-uniform
    const uint32_t g = order ? order[p] : p;
    const uint32_t j = job_of(qbase, n_jobs, g);
    const KnnParams P = jobs[j];
    stream_one<512, kFromRecord>(P, lds[threadIdx.x >> 6], threadIdx.x & 63,
                                 __builtin_amdgcn_readfirstlane(g - qbase[j]), jobs + j);
}

// Persistent form: workgroups stay resident and their waves take tiles of
// kBatchTile positions from a device work counter until it passes the total.
constexpr uint32_t kBatchTile = 4;

template <int CAP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_knn_batch(const KnnParams* __restrict__ jobs, const uint32_t* __restrict__ qbase,
                                                   uint32_t n_jobs, const uint32_t* __restrict__ order, uint32_t total,
                                                   uint32_t* __restrict__ counter) {
    __shared__ WaveLds<CAP> lds[kWaves];
    const uint32_t lane = threadIdx.x & 63;
    WaveLds<CAP>& L = lds[threadIdx.x >> 6];
    for (;;) {
        uint32_t pos = 0;
        if (lane == 0) pos = atomicAdd(counter, kBatchTile);
        pos = __builtin_amdgcn_readfirstlane(pos);
        if (pos >= total) break;
        const uint32_t end = min(total, pos + kBatchTile);
        for (uint32_t p = pos; p < end; ++p) {
            const uint32_t g = order ? order[p] : p;
            const uint32_t j = job_of(qbase, n_jobs, g);
            const KnnParams P = jobs[j];
            stream_one<CAP>(P, L, lane, __builtin_amdgcn_readfirstlane(g - qbase[j]));
            wave_sync();
        }
    }
}

// Two passes: histogram of in-radius d2 over the float-bit bins, then the
// keys below the bin edge that keeps >= k are compacted into the buffer.
__global__ __launch_bounds__(256) void k_knn_two_pass(KnnParams P) {
    constexpr int CAP = 512;
    __shared__ WaveLds<CAP> lds[kWaves];
    const uint32_t lane = threadIdx.x & 63;
    uint32_t q;
    Box B;
    float qx, qy, qz;
    if (!query_of_wave(P, q)) return;
    box_of(P, q, B, qx, qy, qz);
    WaveLds<CAP>& L = lds[threadIdx.x >> 6];
    const float T = P.T;
    const uint32_t bT = P.bT, k = (uint32_t)P.k;
    L.hist[lane] = 0;
    wave_sync();
    walk(P, lane, B, qx, qy, qz, [&](bool valid, uint32_t, float d2) {
        if (valid && d2 <= T) {
            uint32_t h = bT - (__float_as_uint(d2) >> 22);
            atomicAdd(&L.hist[h < 63u ? h : 63u], 1u);
        }
    });
    wave_sync();
    uint32_t cum = wave_incl_scan(L.hist[63 - lane], lane);
    const uint32_t found = __shfl(cum, 63, 64);
    uint64_t khi = ~0ull;
    if (found > k) {
        int ls = __builtin_ctzll(wballot(cum >= k));
        khi = (uint64_t)(bT - (63u - (uint32_t)ls) + 1u) << 54;
    }
    // collect; if the bin edge keeps more than CAP keys, tighten on the fly
    uint32_t cnt = 0;
    walk(P, lane, B, qx, qy, qz, [&](bool valid, uint32_t slot, float d2) {
        uint64_t key = make_key(d2, slot);
        bool acc = valid && d2 <= T && key < khi;
        uint64_t bal = wballot(acc);
        uint32_t na = (uint32_t)__popcll(bal);
        if (cnt + na > (uint32_t)CAP) {
            wave_sync();
            cnt = buf_select<CAP>(L, lane, cnt, k, CAP / 2, bT, khi);
            acc = acc && key < khi;
            bal = wballot(acc);
            na = (uint32_t)__popcll(bal);
        }
        if (acc) L.buf[cnt + mbcnt(bal)] = key;
        cnt += na;
    });
    wave_sync();
    finish<CAP>(P, L, lane, q, cnt, found, khi);
}

// Radiance from a stored kNN result (pm_radiance): one wave per query.
__global__ __launch_bounds__(256) void k_radiance(KnnParams P) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t q = blockIdx.x * kWaves + wave;
    if (q >= P.nq) return;
    const uint32_t m = min(P.out_count[q], (uint32_t)P.k);
    const float alpha = 1.818f, beta = 1.953f;
    const float pi_f = 3.141592654f;
    float res0 = 0.f, res1 = 0.f, res2 = 0.f;
    float rmax = m > 0 ? P.out_dist[(size_t)q * P.k + (m - 1)] : 0.f;
    if (m > 0) {
        const float filter_r = __fmul_rn(rmax, rmax);
        const float denom = 1.f - expf(-beta);
        const float nx = P.nrm[q], ny = P.nrm[P.nq + q], nz = P.nrm[2 * (size_t)P.nq + q];
        for (uint32_t e = lane; e < m; e += 64) {
            uint32_t idx = P.out_idx[(size_t)q * P.k + e];
            float dx = P.dir[idx], dy = P.dir[P.n + idx], dz = P.dir[2 * (size_t)P.n + idx];
            float cosNL = __fadd_rn(__fadd_rn(__fmul_rn(nx, -dx), __fmul_rn(ny, -dy)), __fmul_rn(nz, -dz));
            if (cosNL > 0.f) {
                float d = P.out_dist[(size_t)q * P.k + e];
                float x = __fdiv_rn(__fmul_rn(__fmul_rn(-beta, d), d), __fmul_rn(2.f, filter_r));
                float w = __fmul_rn(alpha, 1.f - __fdiv_rn(1.f - expf(x), denom));
                res0 += P.power[idx] * w;
                res1 += P.power[P.n + idx] * w;
                res2 += P.power[2 * (size_t)P.n + idx] * w;
            }
        }
    }
    res0 = wave_sum_f32(res0);
    res1 = wave_sum_f32(res1);
    res2 = wave_sum_f32(res2);
    if (lane == 0) {
        float coef = m > 0 ? __fdiv_rn(1.f, __fmul_rn(__fmul_rn(pi_f, __fmul_rn(rmax, rmax)), 1.f)) : 0.f;
        P.rgb[q] = res0 * coef;
        P.rgb[P.nq + q] = res1 * coef;
        P.rgb[2 * (size_t)P.nq + q] = res2 * coef;
    }
}

// Batch processing order keys: flattened query g of job j gets
// cell_base[j] + its cell, so one sort orders jobs, then cells within a job.
__global__ __launch_bounds__(256) void k_batch_query_keys(const KnnParams* __restrict__ jobs,
                                                         const uint32_t* __restrict__ qbase,
                                                         const uint32_t* __restrict__ cell_base, uint32_t n_jobs,
                                                         uint32_t total, uint32_t* __restrict__ keys) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    uint32_t lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (qbase[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    const KnnParams& P = jobs[lo];
    const uint32_t q = g - qbase[lo];
    int x = pmd::cell_coord(P.qx[q], P.G.mn[0], P.G.cell, P.G.g[0]);
    int y = pmd::cell_coord(P.qy[q], P.G.mn[1], P.G.cell, P.G.g[1]);
    int z = pmd::cell_coord(P.qz[q], P.G.mn[2], P.G.cell, P.G.g[2]);
    keys[g] = cell_base[lo] + pmd::cell_index(P.G, x, y, z);
}

// Cell key of each query for the optional cell-ordered processing.
__global__ __launch_bounds__(256) void k_query_keys(GridGeom G, const float* __restrict__ qx,
                                                   const float* __restrict__ qy, const float* __restrict__ qz,
                                                   uint32_t nq, uint32_t* __restrict__ keys,
                                                   const uint32_t* __restrict__ nq_dev) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq || (nq_dev && i >= *nq_dev)) return;
    int x = pmd::cell_coord(qx[i], G.mn[0], G.cell, G.g[0]);
    int y = pmd::cell_coord(qy[i], G.mn[1], G.cell, G.g[1]);
    int z = pmd::cell_coord(qz[i], G.mn[2], G.cell, G.g[2]);
    keys[i] = pmd::cell_index(G, x, y, z);
}

// Sum over queries of C_q (cells in the box) and P_q (photons in them).
__global__ __launch_bounds__(256) void k_box_stats(GridGeom G, const uint32_t* __restrict__ starts,
                                                  const uint32_t* __restrict__ sizes, const float* __restrict__ qx,
                                                  const float* __restrict__ qy, const float* __restrict__ qz,
                                                  uint32_t nq, int rc, float T, int reference,
                                                  unsigned long long* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cells = 0, cands = 0;
    if (i < nq) {
        const float qv[3] = {qx[i], qy[i], qz[i]};
        Box B;
        for (int a = 0; a < 3; ++a) {
            int c = pmd::cell_coord(qv[a], G.mn[a], G.cell, G.g[a]);
            B.lo[a] = max(0, c - rc);
            B.hi[a] = min(G.g[a] - 1, c + rc);
        }
        // the runs the gather walks (row_span): cells and photons visited
        const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
        const int nseg = G.layout == PM_LAYOUT_LINEAR ? ny * nz : nx * ny * nz;
        // (C_q stays the box's cell count in the linear layout, an upper
        // bound of the start / size words its row spans read)
        if (reference) {
            // the reference's whole box (PhotonGrid.cu:262-272)
            cells = (unsigned long long)nx * ny * nz;
            for (int z = B.lo[2]; z <= B.hi[2]; ++z)
                for (int y = B.lo[1]; y <= B.hi[1]; ++y) {
                    if (G.layout == PM_LAYOUT_LINEAR) {
                        const uint32_t c0 = pmd::cell_index(G, B.lo[0], y, z), c1 = c0 + (uint32_t)(nx - 1);
                        cands += starts[c1] + sizes[c1] - starts[c0];
                    } else {
                        for (int x = B.lo[0]; x <= B.hi[0]; ++x) cands += sizes[pmd::cell_index(G, x, y, z)];
                    }
                }
        } else {
            for (int s = 0; s < nseg; ++s) {
                uint32_t b, len;
                const bool kept = row_span(G, starts, sizes, B, s, qv[0], qv[1], qv[2], T, b, len);
                cands += len;
                if (G.layout != PM_LAYOUT_LINEAR) cells += kept;
            }
            if (G.layout == PM_LAYOUT_LINEAR) cells = (unsigned long long)nx * ny * nz;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        cells += __shfl_xor(cells, d, 64);
        cands += __shfl_xor(cands, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], cells);
        atomicAdd(&out[1], cands);
    }
}

__global__ void k_zero2(unsigned long long* p) {
    p[0] = 0;
    p[1] = 0;
}

pm_status setup(pm_ctx* ctx, const pm_grid* grid, const float* d_q, uint32_t nq, float radius, int32_t k,
                uint32_t flags, KnnParams& P, const char* who) {
    if (!grid) return pm::fail(ctx, PM_ERR_INVALID, "%s: null grid", who);
    if (nq > 0 && !d_q) return pm::fail(ctx, PM_ERR_INVALID, "%s: null queries", who);
    if (!(radius >= 0.f) || !std::isfinite(radius)) return pm::fail(ctx, PM_ERR_INVALID, "%s: bad radius", who);
    if (k < 1 || k > kMaxK) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "%s: k=%d outside [1, %d]", who, k, kMaxK);
    const pm_grid_desc& D = grid->desc;
    // candidate positions are read through raw buffers with 32-bit byte offsets
    if (D.n_photons >= (1u << 30))
        return pm::fail(ctx, PM_ERR_UNSUPPORTED, "%s: grid of %u photons (the gather takes < 2^30)", who, D.n_photons);
    P = KnnParams{};
    for (int a = 0; a < 3; ++a) {
        P.G.mn[a] = D.aabb[a];
        P.G.g[a] = D.dims[a];
    }
    P.G.cell = D.cell_size;
    P.G.layout = D.layout;
    P.starts = D.d_cell_starts;
    P.sizes = D.d_cell_sizes;
    P.n = D.n_photons;
    P.px = D.d_photons;
    P.py = D.d_photons + P.n;
    P.pz = D.d_photons + 2 * (size_t)P.n;
    P.power = D.d_photons + 3 * (size_t)P.n;
    P.dir = D.d_photons + 6 * (size_t)P.n;
    P.qx = d_q;
    P.qy = d_q + nq;
    P.qz = d_q + 2 * (size_t)nq;
    P.nq = nq;
    P.r = radius;
    P.T = radius_threshold(radius);
    uint32_t tb;
    std::memcpy(&tb, &P.T, 4);
    P.bT = tb >> 22;
    // PhotonGrid.cu:266: radius_cells = (int)ceilf(radius / cell_size)
    float rcf = std::ceil(radius / D.cell_size);
    if (!(rcf < 4096.f)) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "%s: radius/cell ratio too large", who);
    P.rc = (int)rcf;
    P.k = k;
    P.remap = (flags & PM_KNN_NO_XCD_REMAP) ? 0 : (flags & PM_KNN_XCD_CONTIGUOUS) ? 1 : 2;
    P.h = (flags & PM_KNN_INNER_BOX) ? P.rc / 2 : 0;
    return PM_OK;
}

pm_status query_order(pm_ctx* ctx, const pm_grid* grid, const KnnParams& P, uint32_t* order_out) {
    size_t rs = pm::radix_sort_scratch_bytes(P.nq);
    void* work = nullptr;
    pm_status st = pm::scratch_reserve(ctx, pm::carve_size({(size_t)P.nq * 4, (size_t)P.nq * 4, (size_t)P.nq * 4, rs}),
                                       &work);
    if (st != PM_OK) return st;
    pm::Carve cv{static_cast<char*>(work)};
    uint32_t* keys = cv.take<uint32_t>(P.nq);
    uint32_t* skeys = cv.take<uint32_t>(P.nq);
    uint32_t* own = cv.take<uint32_t>(P.nq);
    uint32_t* order = order_out ? order_out : own;
    void* rw = cv.take<char>(rs);
    hipLaunchKernelGGL(k_query_keys, dim3(pm::div_up(P.nq, 256)), dim3(256), 0, ctx->stream, P.G, P.qx, P.qy, P.qz,
                       P.nq, keys, P.nq_dev);
    PM_CHECK_LAUNCH(ctx, "k_query_keys");
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)grid->desc.n_cells) ++bits;
    return pm::radix_sort_pairs_auto(ctx, keys, nullptr, skeys, order, P.nq, bits, rw, P.nq_dev);
}

pm_status launch(pm_ctx* ctx, const pm_grid* grid, KnnParams& P, uint32_t flags) {
    if (P.nq == 0) return PM_OK;
    // radiance (and counts) without neighbour lists for k > 256: the
    // three-walk kernel, with an overflow list behind it
    // (256 < k <= 512 radiance-only: one walk with the 1024-key buffer)
    const bool mid = P.k > 256 && P.k <= 512 && !P.out_idx && P.rgb;
    const bool wide = P.k > 512 && !P.out_idx && P.rgb;
    const bool sort = (flags & PM_KNN_SORT_QUERIES) != 0;
    uint32_t* ovf_list = nullptr;
    uint32_t* ovf_count = nullptr;
    if (sort || wide) {
        // the order lives in the scratch block carved by query_order; the
        // overflow list follows it
        const size_t rs = sort ? pm::radix_sort_scratch_bytes(P.nq) : 0;
        const size_t a = sort ? (size_t)P.nq * 4 : 0;
        void* work = nullptr;
        pm_status st = pm::scratch_reserve(
            ctx, pm::carve_size({a, a, a, rs, wide ? (size_t)P.nq * 4 : 0, wide ? (size_t)64 : 0}), &work);
        if (st != PM_OK) return st;
        pm::Carve cv{static_cast<char*>(work)};
        cv.take<uint32_t>(a / 4);
        cv.take<uint32_t>(a / 4);
        uint32_t* order = cv.take<uint32_t>(a / 4);
        cv.take<char>(rs);
        if (wide) {
            ovf_list = cv.take<uint32_t>(P.nq);
            ovf_count = cv.take<uint32_t>(16);
        }
        if (sort) {
            if ((st = query_order(ctx, grid, P, order)) != PM_OK) return st;
            P.order = order;
        }
    }
    if (wide) {
        pm_status st = pm::fill_u32(ctx, ovf_count, 0u, 1);
        if (st != PM_OK) return st;
        uint32_t nblk = pm::div_up(P.nq, kWaves);
        if (P.remap == 2) nblk = pm::div_up(nblk, 8 * kXcdChunk) * 8 * kXcdChunk;
        hipLaunchKernelGGL(k_gather_wide, dim3(nblk), dim3(256), 0, ctx->stream, P, ovf_list, ovf_count);
        PM_CHECK_LAUNCH(ctx, "k_gather_wide");
        const uint32_t lblocks = std::min<uint32_t>(P.nq, (uint32_t)std::max(ctx->cu_count, 1) * 8u);
        if (P.k > 1024)
            hipLaunchKernelGGL(k_knn_large_list<4096>, dim3(lblocks), dim3(64), 0, ctx->stream, P, ovf_list, ovf_count);
        else
            hipLaunchKernelGGL(k_knn_large_list<2048>, dim3(lblocks), dim3(64), 0, ctx->stream, P, ovf_list, ovf_count);
        PM_CHECK_LAUNCH(ctx, "k_knn_large_list");
        return PM_OK;
    }
    if (mid) {
        uint32_t nblk = pm::div_up(P.nq, kWaves);
        if (P.remap == 2) nblk = pm::div_up(nblk, 8 * kXcdChunk) * 8 * kXcdChunk;
        hipLaunchKernelGGL(k_knn_stream1024, dim3(nblk), dim3(256), 0, ctx->stream, P);
        PM_CHECK_LAUNCH(ctx, "k_knn_stream1024");
        return PM_OK;
    }
    if (P.k > 1024) {
        hipLaunchKernelGGL(k_knn_large<4096>, dim3(P.nq), dim3(64), 0, ctx->stream, P);
        PM_CHECK_LAUNCH(ctx, "k_knn_large");
        return PM_OK;
    }
    if (P.k > 256) {
        hipLaunchKernelGGL(k_knn_large<2048>, dim3(P.nq), dim3(64), 0, ctx->stream, P);
        PM_CHECK_LAUNCH(ctx, "k_knn_large");
        return PM_OK;
    }
    uint32_t nblk = pm::div_up(P.nq, kWaves);
    if (P.remap == 2) nblk = pm::div_up(nblk, 8 * kXcdChunk) * 8 * kXcdChunk;
    const dim3 grid_dim(nblk), block(256);
    if (flags & PM_KNN_TWO_PASS)
        hipLaunchKernelGGL(k_knn_two_pass, grid_dim, block, 0, ctx->stream, P);
    else if (flags & PM_KNN_LARGE_BUFFER)
        hipLaunchKernelGGL(k_knn_stream1024, grid_dim, block, 0, ctx->stream, P);
    else
        hipLaunchKernelGGL(k_knn_stream512, grid_dim, block, 0, ctx->stream, P);
    PM_CHECK_LAUNCH(ctx, "k_knn");
    return PM_OK;
}

}  // namespace

extern "C" pm_status pm_grid_knn(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, uint32_t nq, float radius,
                                 int32_t k, uint32_t flags, uint32_t* d_idx, float* d_dist, uint32_t* d_count) {
    if (!ctx) return PM_ERR_INVALID;
    flags = pm::tuned_flags(ctx, grid, radius, k, flags);
    KnnParams P;
    pm_status st = setup(ctx, grid, d_queries3, nq, radius, k, flags, P, "pm_grid_knn");
    if (st != PM_OK) return st;
    if (nq > 0 && (!d_idx || !d_dist || !d_count)) return pm::fail(ctx, PM_ERR_INVALID, "pm_grid_knn: null output");
    P.out_idx = d_idx;
    P.out_dist = d_dist;
    P.out_count = d_count;
    return launch(ctx, grid, P, flags);
}

extern "C" pm_status pm_grid_gather_radiance(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3,
                                             const float* d_normals3, uint32_t nq, float radius, int32_t k,
                                             uint32_t flags, float* d_rgb3, uint32_t* d_count) {
    if (!ctx) return PM_ERR_INVALID;
    flags = pm::tuned_flags(ctx, grid, radius, k, flags);
    KnnParams P;
    pm_status st = setup(ctx, grid, d_queries3, nq, radius, k, flags, P, "pm_grid_gather_radiance");
    if (st != PM_OK) return st;
    if (nq > 0 && (!d_normals3 || !d_rgb3))
        return pm::fail(ctx, PM_ERR_INVALID, "pm_grid_gather_radiance: null normals/output");
    P.nrm = d_normals3;
    P.rgb = d_rgb3;
    P.out_count = d_count;
    return launch(ctx, grid, P, flags);
}

namespace pm {
pm_status gather_radiance_counted(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, const float* d_normals3,
                                  uint32_t cap, const uint32_t* d_nq, float radius, int32_t k, uint32_t flags,
                                  float* d_rgb3, uint32_t* d_count) {
    KnnParams P;
    pm_status st = setup(ctx, grid, d_queries3, cap, radius, k, flags, P, "gather (device count)");
    if (st != PM_OK) return st;
    P.nq_dev = d_nq;
    P.nrm = d_normals3;
    P.rgb = d_rgb3;
    P.out_count = d_count;
    return launch(ctx, grid, P, flags);
}
}  // namespace pm

extern "C" pm_status pm_radiance(pm_ctx* ctx, const pm_grid* grid, const uint32_t* d_idx, const float* d_dist,
                                 const uint32_t* d_count, const float* d_normals3, uint32_t nq, int32_t k,
                                 float* d_rgb3) {
    if (!ctx) return PM_ERR_INVALID;
    if (!grid) return pm::fail(ctx, PM_ERR_INVALID, "pm_radiance: null grid");
    if (k < 1) return pm::fail(ctx, PM_ERR_INVALID, "pm_radiance: k < 1");
    if (nq == 0) return PM_OK;
    if (!d_idx || !d_dist || !d_count || !d_normals3 || !d_rgb3)
        return pm::fail(ctx, PM_ERR_INVALID, "pm_radiance: null buffer");
    const pm_grid_desc& D = grid->desc;
    KnnParams P{};
    P.n = D.n_photons;
    P.power = D.d_photons + 3 * (size_t)P.n;
    P.dir = D.d_photons + 6 * (size_t)P.n;
    P.nq = nq;
    P.k = k;
    P.out_idx = const_cast<uint32_t*>(d_idx);
    P.out_dist = const_cast<float*>(d_dist);
    P.out_count = const_cast<uint32_t*>(d_count);
    P.nrm = d_normals3;
    P.rgb = d_rgb3;
    hipLaunchKernelGGL(k_radiance, dim3(pm::div_up(nq, kWaves)), dim3(256), 0, ctx->stream, P);
    PM_CHECK_LAUNCH(ctx, "k_radiance");
    return PM_OK;
}

extern "C" pm_status pm_grid_box_stats(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, uint32_t nq,
                                       float radius, uint32_t flags, uint64_t* h_sum_cells, uint64_t* h_sum_cands) {
    if (!ctx) return PM_ERR_INVALID;
    KnnParams P;
    pm_status st = setup(ctx, grid, d_queries3, nq, radius, 1, 0, P, "pm_grid_box_stats");
    if (st != PM_OK) return st;
    void* work = nullptr;
    if ((st = pm::scratch_reserve(ctx, 256, &work)) != PM_OK) return st;
    unsigned long long* acc = static_cast<unsigned long long*>(work);
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1), 0, ctx->stream, acc);
    if (nq > 0)
        hipLaunchKernelGGL(k_box_stats, dim3(pm::div_up(nq, 256)), dim3(256), 0, ctx->stream, P.G, P.starts, P.sizes,
                           P.qx, P.qy, P.qz, nq, P.rc, P.T, (flags & PM_BOX_REFERENCE) ? 1 : 0, acc);
    PM_CHECK_LAUNCH(ctx, "k_box_stats");
    unsigned long long h[2] = {0, 0};
    PM_HIP(ctx, hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h_sum_cells) *h_sum_cells = h[0];
    if (h_sum_cands) *h_sum_cands = h[1];
    return PM_OK;
}

extern "C" pm_status pm_grid_query_order(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, uint32_t nq,
                                         uint32_t* d_order) {
    if (!ctx) return PM_ERR_INVALID;
    KnnParams P;
    pm_status st = setup(ctx, grid, d_queries3, nq, 0.f, 1, 0, P, "pm_grid_query_order");
    if (st != PM_OK) return st;
    if (nq == 0) return PM_OK;
    if (!d_order) return pm::fail(ctx, PM_ERR_INVALID, "pm_grid_query_order: null output");
    return query_order(ctx, grid, P, d_order);
}

extern "C" pm_status pm_gather_batch(pm_ctx* ctx, const pm_gather_job* jobs, uint32_t n_jobs, uint32_t flags) {
    if (!ctx) return PM_ERR_INVALID;
    if (n_jobs == 0) return PM_OK;
    if (!jobs) return pm::fail(ctx, PM_ERR_INVALID, "pm_gather_batch: null job table");
    // Jobs with k <= 256 share one launch of the batch kernel; larger k
    // (one wave per query over a 2048/4096-entry LDS buffer) run as their own
    // launches of the single-map kernel, first, on the same stream.
    std::vector<KnnParams> hp;
    std::vector<std::pair<const pm_grid*, KnnParams>> large;
    std::vector<uint32_t> qbase, cell_base;
    hp.reserve(n_jobs);
    uint64_t total = 0, cells = 0;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const pm_gather_job& J = jobs[j];
        // this job's flags: the call's, its own, and its tuning
        const uint32_t jf = pm::tuned_flags(ctx, J.grid, J.radius, J.k, flags | J.flags);
        KnnParams P;
        pm_status st = setup(ctx, J.grid, J.d_queries3, J.nq, J.radius, J.k, jf, P, "pm_gather_batch");
        if (st != PM_OK) return st;
        if (J.nq > 0) {
            if ((J.d_idx == nullptr) != (J.d_dist == nullptr))
                return pm::fail(ctx, PM_ERR_INVALID, "pm_gather_batch: job %u needs both d_idx and d_dist", j);
            if (J.d_rgb3 && !J.d_normals3)
                return pm::fail(ctx, PM_ERR_INVALID, "pm_gather_batch: job %u radiance without normals", j);
            if (!J.d_idx && !J.d_count && !J.d_rgb3)
                return pm::fail(ctx, PM_ERR_INVALID, "pm_gather_batch: job %u has no output", j);
        }
        P.out_idx = J.d_idx;
        P.out_dist = J.d_dist;
        P.out_count = J.d_count;
        P.nrm = J.d_rgb3 ? J.d_normals3 : nullptr;
        P.rgb = J.d_rgb3;
        P.order = nullptr;
        if (J.k > 256) {
            if (J.nq > 0) large.emplace_back(J.grid, P);
            continue;
        }
        qbase.push_back((uint32_t)total);
        cell_base.push_back((uint32_t)cells);
        hp.push_back(P);
        total += J.nq;
        cells += J.grid->desc.n_cells;
        if (total > 0xFFFFFFFFull) return pm::fail(ctx, PM_ERR_UNSUPPORTED, "pm_gather_batch: over 2^32 queries");
    }
    for (auto& L : large) {
        pm_status st = launch(ctx, L.first, L.second, flags & PM_KNN_SORT_QUERIES);
        if (st != PM_OK) return st;
    }
    if (total == 0) return PM_OK;
    n_jobs = (uint32_t)hp.size();
    const bool sorted = (flags & PM_KNN_SORT_QUERIES) && cells <= 0xFFFFFFFFull;
    const uint32_t nt = (uint32_t)total;
    const size_t rs = sorted ? pm::radix_sort_scratch_bytes(nt) : 0;
    const size_t bytes = pm::carve_size({(size_t)n_jobs * sizeof(KnnParams), (size_t)n_jobs * 4, (size_t)n_jobs * 4,
                                         64, sorted ? (size_t)nt * 4 : 0, sorted ? (size_t)nt * 4 : 0,
                                         sorted ? (size_t)nt * 4 : 0, rs});
    void* work = nullptr;
    pm_status st = pm::scratch_reserve(ctx, bytes, &work);
    if (st != PM_OK) return st;
    pm::Carve cv{static_cast<char*>(work)};
    KnnParams* d_jobs = cv.take<KnnParams>(n_jobs);
    uint32_t* d_qbase = cv.take<uint32_t>(n_jobs);
    uint32_t* d_cbase = cv.take<uint32_t>(n_jobs);
    uint32_t* d_counter = cv.take<uint32_t>(16);
    PM_HIP(ctx, hipMemcpyAsync(d_jobs, hp.data(), hp.size() * sizeof(KnnParams), hipMemcpyHostToDevice, ctx->stream));
    PM_HIP(ctx, hipMemcpyAsync(d_qbase, qbase.data(), qbase.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    PM_HIP(ctx, hipMemcpyAsync(d_cbase, cell_base.data(), cell_base.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    // the tables are copied from pageable host vectors: wait before they go
    PM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t* order = nullptr;
    if (sorted) {
        uint32_t* keys = cv.take<uint32_t>(nt);
        uint32_t* skeys = cv.take<uint32_t>(nt);
        uint32_t* ord = cv.take<uint32_t>(nt);
        void* rw = cv.take<char>(rs);
        hipLaunchKernelGGL(k_batch_query_keys, dim3(pm::div_up(nt, 256)), dim3(256), 0, ctx->stream, d_jobs, d_qbase,
                           d_cbase, n_jobs, nt, keys);
        PM_CHECK_LAUNCH(ctx, "k_batch_query_keys");
        int bits = 1;
        while (bits < 32 && (1ull << bits) < cells) ++bits;
        if ((st = pm::radix_sort_pairs_auto(ctx, keys, nullptr, skeys, ord, nt, bits, rw)) != PM_OK) return st;
        order = ord;
    }
    if (flags & PM_KNN_BATCH_QUEUE) {
        if ((st = pm::fill_u32(ctx, d_counter, 0u, 1)) != PM_OK) return st;
        const uint32_t blocks = std::min<uint32_t>(pm::div_up(pm::div_up(nt, kBatchTile), kWaves),
                                                   (uint32_t)std::max(ctx->cu_count, 1) * 8u);
        hipLaunchKernelGGL(k_knn_batch<512>, dim3(blocks), dim3(256), 0, ctx->stream, d_jobs, d_qbase, n_jobs, order,
                           nt, d_counter);
    } else {
        const uint32_t step = 8 * kXcdChunk;
        const uint32_t blocks = pm::div_up(pm::div_up(nt, kWaves), step) * step;
        const uint32_t gx = std::min<uint32_t>(blocks, 65536u);
        hipLaunchKernelGGL(k_knn_batch_grid, dim3(gx, pm::div_up(blocks, gx)), dim3(256), 0, ctx->stream, d_jobs,
                           d_qbase, n_jobs, order, nt, blocks);
    }
    PM_CHECK_LAUNCH(ctx, "k_knn_batch");
    return PM_OK;
}
