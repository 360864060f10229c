/* CPU oracle for the photon pipeline — C ABI used by the Python test harness.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of the reference's CPU semantics,
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker. The product library (libpm355.so) never links or calls it.
 *
 * Photon blocks use the same layout as the product: 9 planes of n floats,
 * in the field order of the reference record (Photon.cuh:7-10):
 *   pos.x pos.y pos.z | power.r power.g power.b | inc_dir.x inc_dir.y inc_dir.z
 * Query/normal blocks are 3 planes of nq floats.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* PrefixScanSum.cuh:12-148 (prescan<T, inclusive>): sequential scan. */
void orc_scan_u32(const uint32_t* in, uint32_t n, int inclusive, uint32_t* out, uint64_t* total);

/* PhotonGrid.cu:41-56 + :92-141 — cell size and grid dims. cell_size > 0 is
 * used as given; otherwise the reference rule cbrt(V / ceil(n/desired)) with
 * the density re-check (max count > 5*desired -> desired/5). layout 0 =
 * linear x + y*gx + z*gx*gy (PhotonGrid.cu:25-27), 1 = Morton (bit-interleaved
 * x, y, z). */
int orc_grid_plan(const float* photons9, uint32_t n, const float aabb[6], float cell_size,
                  int32_t desired, int32_t layout, float* out_cell, int32_t out_dims[3],
                  uint32_t* out_ncells);

/* Count -> exclusive scan -> stable scatter (PhotonGrid.cu:58-218). In-cell
 * order is ascending photon index. perm[j] = original index of sorted slot j. */
int orc_grid_fill(const float* photons9, uint32_t n, const float aabb[6], float cell, const int32_t dims[3],
                  int32_t layout, uint32_t* counts, uint32_t* starts, float* sorted9, uint32_t* perm);

/* PhotonGrid::find_nearests (PhotonGrid.cu:222-284) with PhotonMaxHeap
 * (PhotonMaxHeap.cuh:90-107) replaced by exact selection on the total order
 * (d2, grid index), d2 = (dx*dx + dy*dy) + dz*dz in FP32, radius test
 * sqrtf(d2) <= r as in the reference. Outputs are [nq][k], ascending; slots
 * past count hold idx = 0xFFFFFFFF, dist = +inf. Optional stats: C_q (cells
 * in the clamped box) and P_q (photons in those cells), summed. */
int orc_grid_knn(const uint32_t* counts, const uint32_t* starts, const float* sorted9, uint32_t n,
                 const float aabb[6], float cell, const int32_t dims[3], int32_t layout,
                 const float* queries3, uint32_t nq, float r, int32_t k, int32_t threads,
                 uint32_t* out_idx, float* out_dist, uint32_t* out_count,
                 uint64_t* sum_cells, uint64_t* sum_cands, uint64_t* sum_found);

/* Brute force over all photons of a block: same order and output convention. */
int orc_knn_brute(const float* pos3, uint32_t n, const float* queries3, uint32_t nq, float r, int32_t k,
                  uint32_t* out_idx, float* out_dist, uint32_t* out_count);

/* Raytracer::get_color_from_photon_grid (Raytracer.cuh:110-138) in FP32, and
 * its FP64 recomputation over the same selected set with the same FP32
 * cos > 0 decisions. power3/dir3 point at the power and inc_dir planes of a
 * (grid-ordered) photon block of n photons. */
int orc_radiance(const uint32_t* idx, const float* dist, const uint32_t* count, uint32_t nq, int32_t k,
                 const float* power3, const float* dir3, uint32_t n, const float* normals3,
                 float* out_rgb32, double* out_rgb64);

/* PhotonMap::fill_balanced (PhotonMap.cuh:189-283) with the sort order made
 * total: (pos[plane], photon index). Implicit heap of cap = 2^L - 1 slots,
 * L = ceil(log2(n+1)); node_idx = original photon index (0xFFFFFFFF empty),
 * node_plane = split axis (-1 empty). */
int orc_kdtree_build(const float* pos3, uint32_t n, uint32_t cap, uint32_t* node_idx, int8_t* node_plane);

/* k nearest by (d2, photon index) within r, searched on the kd-tree built
 * above; indices are original photon indices. */
int orc_kdtree_knn(const float* pos3, uint32_t n, const uint32_t* node_idx, const int8_t* node_plane,
                   uint32_t cap, const float* queries3, uint32_t nq, float r, int32_t k, int32_t threads,
                   uint32_t* out_idx, float* out_dist, uint32_t* out_count);

/* AABB of each model's vertices (KdTree.cu:82-89,243-327 reduce to this:
 * min/max over the model's vertex positions). verts9 holds one triangle per
 * 9 floats (v0.xyz v1.xyz v2.xyz); model m spans triangles
 * [tri_begin[m], tri_begin[m+1]). */
int orc_model_aabb(const float* verts9, const uint32_t* tri_begin, int32_t n_models, float* out_aabb6);

/* Largest float t with sqrtf(t) <= r (the exact d2 threshold of the
 * reference's distance test). */
float orc_radius_threshold(float r);

/* Photon emission and bounce tracing (PhotonTracing.cu:278-375): paths in
 * index order, each with its own Philox stream (seed, 0, path); the maps
 * take the first n_global / n_caustic photons in (path, event) order.
 * models_raw points at n_models pm_model records (include/pm355.h).
 * out_stats = {paths traced, global stored, caustic stored}. */
/* Nearest hit for each ray of a 6-plane ray block (origin, direction):
 * t = -1, model = -1 on a miss. */
int orc_scene_hit(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
                  int32_t n_models, const float* rays6, uint32_t nr, float* out_t, int32_t* out_model,
                  float* out_n3);

/* Eye-ray frame: Raytracer::render_cpu / render_trace (Raytracer.cuh:140-299)
 * over an oracle global grid and an optional caustic grid (c_n = 0: none).
 * camera_raw = pm_camera; iparams = {max_depth, k_global, k_caustic, flags},
 * fparams = {r_global, r_caustic}. out_rgb3: 3 planes of W*H HDR-mapped
 * floats; out_rgb8: W*H*3 bytes (uchar3). out_stats[0] = eye-ray hits. */
int orc_render(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
               int32_t n_models, float default_ior, const uint32_t* g_counts, const uint32_t* g_starts,
               const float* g_photons9, uint32_t g_n, const float* g_aabb, float g_cell, const int32_t* g_dims,
               int32_t g_layout, const uint32_t* c_counts, const uint32_t* c_starts, const float* c_photons9,
               uint32_t c_n, const float* c_aabb, float c_cell, const int32_t* c_dims, int32_t c_layout,
               const void* camera_raw, const int32_t* iparams, const float* fparams, int32_t threads,
               float* out_rgb3, uint8_t* out_rgb8, uint64_t* out_stats);

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Threads of the OpenMP parts that take no thread argument (the tracer);
 * n <= 0 restores the default. Returns the count now in effect. */
int orc_set_threads(int32_t n);

int orc_trace_photons(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
                      int32_t n_models, int32_t light_model, const float intensity[3], float default_ior,
                      uint64_t seed, uint32_t n_global, uint32_t n_caustic, int32_t max_bounces, uint64_t max_paths,
                      float* out_global9, float* out_caustic9, uint64_t* out_stats);

/* One bounce of one light path, as orc_trace_photons decides it: the
 * incoming ray (o, d), the hit point p, the normal n turned against d, the
 * event dest (0 diffuse, 1 specular, 2 absorb, 3 refract: the Path values of
 * PathOperator.cuh), stored (0 none, 1 global map, 2 caustic map), flags
 * (bit 0: the critical-angle test ruled refraction out; bit 1: Ray::refract
 * failed and the path ended), the outgoing ray (zero when the path ends
 * there) and the refraction pair (n1, n2) consulted (0 on opaque models). */
typedef struct {
    uint32_t path;
    int32_t bounce, model, dest, stored, flags;
    float o[3], d[3], p[3], n[3], next_o[3], next_d[3];
    float n1, n2;
} orc_trace_event;

/* Every bounce event of paths [0, n_paths) of the seed's numbering, in
 * (path, bounce) order, into out[cap]; *n_events = total (returns -4 when it
 * exceeds cap). n_global sets the photon power as in orc_trace_photons. */
int orc_trace_events(const float* tris9, const float* normals9, uint32_t n_tris, const void* models_raw,
                     int32_t n_models, int32_t light_model, const float intensity[3], float default_ior,
                     uint64_t seed, uint32_t n_global, int32_t max_bounces, uint64_t n_paths, orc_trace_event* out,
                     uint64_t cap, uint64_t* n_events);

#ifdef __cplusplus
}
#endif
#endif
