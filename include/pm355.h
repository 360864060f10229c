/* pm355 — MI355X-native photon pipeline: C ABI.
 *
 * This is the drop-in boundary for the reference's GPU photon path
 * (Alator-ux/Cuda-Photon-Mapping, directory CudaPhotonMapping/). The
 * reference has no FFI; its entry points are C++ free functions and classes
 * compiled into one executable. Each function below names the reference
 * interface it replaces. All functions return a pm_status (0 = OK, negative
 * = error class); pm_last_error() gives the message. Nothing here exits the
 * process (the reference calls exit(99) on any CUDA error, CudaUtils.cu:6-14).
 *
 * Memory: pointers named d_* are device (HBM) pointers, h_* are host
 * pointers. Inputs are borrowed and never freed by the callee (unlike
 * PhotonGrid::build, which frees its input, PhotonGrid.cu:216). Handles
 * (pm_grid, pm_kdtree, ...) own their device buffers and are released with
 * their *_destroy call. All work is enqueued on the context's HIP stream;
 * pm_ctx_sync() waits for it.
 *
 * Photon block layout ("photons9"): 9 planes of n FP32 values in the field
 * order of cpm::Photon (Photon.cuh:7-10):
 *   pos.x pos.y pos.z | power.r power.g power.b | inc_dir.x inc_dir.y inc_dir.z
 * i.e. plane p of photon i is at photons9[p * n + i]. Query, normal and RGB
 * blocks are 3 planes of nq values.
 */
#ifndef PM355_H
#define PM355_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pm_status;
enum {
    PM_OK = 0,
    PM_ERR_INVALID = -1,     /* bad argument */
    PM_ERR_HIP = -2,         /* HIP runtime error */
    PM_ERR_NOMEM = -3,       /* device allocation failed */
    PM_ERR_CAPACITY = -4,    /* an output or internal capacity was exceeded */
    PM_ERR_UNSUPPORTED = -5, /* parameter outside the supported range */
    PM_ERR_NO_DEVICE = -6    /* no usable gfx950 device */
};

enum { PM_LAYOUT_LINEAR = 0, PM_LAYOUT_MORTON = 1 };

/* pm_grid_knn / pm_grid_gather_radiance flags */
enum {
    PM_KNN_SORT_QUERIES = 1u << 0, /* process queries in cell order (outputs stay in caller order) */
    PM_KNN_NO_XCD_REMAP = 1u << 1, /* disable the XCD-aware block mapping (hardware order) */
    PM_KNN_TWO_PASS = 1u << 4,     /* tuning: histogram pass + collect pass instead of one streaming pass */
    PM_KNN_LARGE_BUFFER = 1u << 5,   /* tuning: 1024-key LDS buffer per query instead of 512 (fewer
                                        narrowing passes, lower occupancy) */
    PM_KNN_BATCH_QUEUE = 1u << 6,    /* tuning (pm_gather_batch): persistent workgroups on a device work queue */
    PM_KNN_XCD_CONTIGUOUS = 1u << 7, /* tuning: each XCD takes one contiguous eighth of the workgroups
                                        instead of interleaved chunks of 256 */
    PM_KNN_INNER_BOX = 1u << 8,      /* k <= 256: each query first walks only its cell +- ceil(r/cell)/2 and
                                        stops there when the k-th nearest found is closer than anything
                                        outside; otherwise it walks the whole box (pays off when cells are
                                        small against r and k is well below the photons within r) */
    PM_KNN_TUNED = 1u << 9           /* use the context's autotuned variant (pm_tune_gather) for this
                                        gather's k and photon density in place of the tuning bits given;
                                        the bits given stay when nothing was tuned */
};

typedef struct pm_ctx pm_ctx;
typedef struct pm_grid pm_grid;
typedef struct pm_kdtree pm_kdtree;
typedef struct pm_accel pm_accel;
typedef struct pm_frame pm_frame;
typedef struct pm_server pm_server;

/* ---- context (replaces the reference's global device state: GlobalParams
 * __constant__ mirrors GlobalParams.cu:1-58, the heap pointers of
 * set_photon_heap_parameters PhotonMaxHeap.cu:9-17, and checkCudaErrors
 * Defines.cuh:36). Calls are asynchronous on the context's stream.
 * Graph capture: pm_scan_u32, pm_grid_knn, pm_grid_gather_radiance,
 * pm_radiance and pm_grid_query_order make no host synchronisation and no
 * allocation once the context has served a call of the same size, so they
 * can be captured into a HIP graph and replayed (the look-back scans keep
 * their epoch in device memory); a captured graph borrows the context's
 * scratch, so a later call that grows it invalidates the graph. The builds,
 * the tracer, pm_gather_batch, frames and the server allocate or read
 * counts back to the host and are not capturable. ------------------------- */
const char* pm_version(void);
pm_status pm_ctx_create(int32_t device, void* hip_stream, pm_ctx** out);
pm_status pm_ctx_destroy(pm_ctx* ctx);
pm_status pm_ctx_set_stream(pm_ctx* ctx, void* hip_stream);
pm_status pm_ctx_sync(pm_ctx* ctx);
/* Destroyed grids and frames leave their device blocks in their context's
 * cache (up to a quarter of the device's memory) for later builds and frames
 * of similar size; pm_ctx_trim synchronises the stream and frees them. (No
 * counterpart in the reference, which frees per call.) */
pm_status pm_ctx_trim(pm_ctx* ctx);
const char* pm_last_error(const pm_ctx* ctx);

/* ---- prefix scan: prescan<uint, inclusive> (PrefixScanSum.cuh:12-148) and
 * the dead cooperative_inclusive_prescan (CGPrefixScanSum.cu:65-101). Single
 * pass, wave64 scan with decoupled look-back; no grid barrier. d_total
 * (nullable, device u64) receives the sum of all inputs. */
pm_status pm_scan_u32(pm_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint32_t n, int32_t inclusive,
                      uint64_t* d_total);

/* ---- hash grid: PhotonGrid(int, AABB) + PhotonGrid::build
 * (PhotonGrid.cuh:44-66, PhotonGrid.cu:41-218). cell_size > 0 is used as
 * given; cell_size <= 0 applies the reference rule cbrt(V / ceil(n/desired))
 * with its density re-check (PhotonGrid.cu:92-141). The build sorts the
 * photons' cell keys stably (radix sort of (key, index)), takes each cell's
 * size from the sorted runs and its start from their exclusive scan, and
 * gathers the photons into cell order: in-cell order is ascending input
 * index, so the grid is deterministic (the reference's atomicAdd slot is
 * not, PhotonGrid.cu:167). */
typedef struct {
    float aabb[6];       /* min xyz, max xyz */
    float cell_size;
    int32_t dims[3];
    int32_t layout;      /* PM_LAYOUT_* */
    uint32_t n_cells;    /* linear: gx*gy*gz; Morton: side^3 */
    uint32_t n_photons;
    const uint32_t* d_cell_starts; /* [n_cells] exclusive scan of sizes */
    const uint32_t* d_cell_sizes;  /* [n_cells] photons per cell */
    const float* d_photons;        /* photons9 in grid order */
    const uint32_t* d_perm;        /* [n] grid slot -> input index */
} pm_grid_desc;

pm_status pm_grid_build(pm_ctx* ctx, const float* d_photons9, uint32_t n, const float aabb[6], float cell_size,
                        int32_t desired_per_cell, int32_t layout, pm_grid** out);

/* Rebuild a grid in place from n photons (n <= the photon count it was built
 * with), keeping its cell size, dims and layout: the per-frame rebuild of a
 * resident grid (the reference builds once per photon pass,
 * Initializer.cu:157-160). Allocates nothing and never synchronises with the
 * host once the context has served a build of n photons, so it can be
 * captured into a HIP graph. PM_ERR_CAPACITY if n is too large. */
pm_status pm_grid_rebuild(pm_ctx* ctx, pm_grid* grid, const float* d_photons9, uint32_t n);
pm_status pm_grid_describe(const pm_grid* grid, pm_grid_desc* out);
/* PhotonGrid::copy_to_cpu (PhotonGrid.cuh:75-98). Any output may be NULL. */
pm_status pm_grid_copy_to_host(pm_ctx* ctx, const pm_grid* grid, uint32_t* h_starts, uint32_t* h_sizes,
                               float* h_photons9, uint32_t* h_perm);
/* The grid's device memory goes back to the context that built it, whose
 * later grid builds and frames reuse it (ordered on that context's stream
 * after work already enqueued there; up to a quarter of the device's memory
 * per context, the rest is freed; pm_ctx_trim frees the cache). A grid
 * read by work on another stream must have that work completed (or the
 * stream synchronised) before it is destroyed. */
pm_status pm_grid_destroy(pm_grid* grid);

/* ---- k-nearest-photon gather: PhotonGrid::find_nearests +
 * PhotonMaxHeap::push (PhotonGrid.cu:222-284, PhotonMaxHeap.cuh:90-107),
 * batched over nq queries. Same candidate box (query cell +- ceil(r/cell),
 * clamped) and radius test (sqrtf(d2) <= r) as the reference; the kept set is
 * the k smallest by (d2, grid index) with d2 = (dx*dx + dy*dy) + dz*dz in
 * FP32. Outputs are [nq][k] in ascending order: d_idx = grid slot (as the
 * heap's PhotonMaxHeapItem::idx), d_dist = sqrtf(d2); slots past d_count[q]
 * hold idx 0xFFFFFFFF and dist +inf. 1 <= k <= 2048; k > 256 (the shipped
 * renderer uses k = 2000 and 400) runs a one-wave-per-workgroup kernel that
 * sorts in LDS. The grid holds fewer than 2^30 photons (positions are read
 * with 32-bit byte offsets); larger grids are rejected with
 * PM_ERR_UNSUPPORTED. */
pm_status pm_grid_knn(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, uint32_t nq, float radius,
                      int32_t k, uint32_t flags, uint32_t* d_idx, float* d_dist, uint32_t* d_count);

/* Radiance estimate: Raytracer::get_color_from_photon_grid
 * (Raytracer.cuh:110-138) over the output of pm_grid_knn. d_rgb3 is 3 planes. */
pm_status pm_radiance(pm_ctx* ctx, const pm_grid* grid, const uint32_t* d_idx, const float* d_dist,
                      const uint32_t* d_count, const float* d_normals3, uint32_t nq, int32_t k, float* d_rgb3);

/* Fused gather + radiance: the kNN selection of pm_grid_knn with the estimate
 * of pm_radiance in its epilogue. d_count (nullable) receives the kept count. */
pm_status pm_grid_gather_radiance(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3,
                                  const float* d_normals3, uint32_t nq, float radius, int32_t k, uint32_t flags,
                                  float* d_rgb3, uint32_t* d_count);

/* pm_grid_gather_radiance without allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that must not call hipMalloc, hosts that
 * forbid it). Same arguments, kernels and results as pm_grid_gather_radiance;
 * the call's scratch (query order, radix-sort passes, large-k overflow list,
 * look-back scan state) lives in a fixed device array of about 26 MB per GPU,
 * held until the process ends, and nq is limited to 1,048,576
 * (PM_ERR_UNSUPPORTED above). Calls on one GPU share that array: calls that
 * overlap on one GPU race. */
pm_status pm_grid_gather_radiance_noalloc(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3,
                                          const float* d_normals3, uint32_t nq, float radius, int32_t k,
                                          uint32_t flags, float* d_rgb3, uint32_t* d_count);

/* Cell-coherent processing order of a query block: d_order[i] = index of the
 * i-th query when queries are sorted (stably) by grid cell. Gathering the
 * queries in this order lets neighbouring waves share cells in cache; the
 * renderer's eye rays have this coherence naturally, the reference gets it
 * from its pixel order (Raytracer.cuh:267). */
pm_status pm_grid_query_order(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, uint32_t nq,
                              uint32_t* d_order);

/* Batched gather: the queries of many jobs (each with its own grid, radius,
 * k and outputs) in one persistent launch over a shared work queue, the
 * batched form of find_nearests + get_color_from_photon_grid for many scenes
 * and frames at once (Drawer.cu:87-101 launches one frame at a time). Each
 * job's outputs are exactly those of pm_grid_knn (d_idx + d_dist, d_count)
 * and/or pm_grid_gather_radiance (d_rgb3 from d_normals3) on its own.
 * flags: PM_KNN_SORT_QUERIES orders all queries by (job, cell) with one
 * sort. Jobs with k > 256 (the reference frame's k = 2000 / 400) run as
 * their own launches of the large-k gather (one wave per query) ahead of the
 * shared launch. Calls on one context are ordered on its stream; the job
 * table is read before the call returns. */
typedef struct {
    const pm_grid* grid;
    const float* d_queries3;  /* [3][nq] */
    uint32_t nq;
    float radius;
    int32_t k;                /* 1..2048 */
    uint32_t* d_idx;          /* optional [nq][k] grid slots, with d_dist */
    float* d_dist;            /* optional [nq][k] */
    uint32_t* d_count;        /* optional [nq] */
    const float* d_normals3;  /* [3][nq], needed with d_rgb3 */
    float* d_rgb3;            /* optional [3][nq] radiance */
    uint32_t flags;           /* this job's PM_KNN_INNER_BOX / PM_KNN_TUNED, OR-ed with the call's flags */
} pm_gather_job;

pm_status pm_gather_batch(pm_ctx* ctx, const pm_gather_job* h_jobs, uint32_t n_jobs, uint32_t flags);

/* ---- autotuner: the launch configuration of the fused gather, chosen by
 * measurement (the reference sizes its launches by fixed rules,
 * CudaUtils.cu:42-68, Defines.cuh:5-72). pm_tune_gather times, on the
 * caller's photons and queries, a grid build (rebuilt in place) at each
 * candidate cell size (cell / r in {1/4, 1/3, 0.4, 1/2, 2/3, 1}) and, on the
 * cell-ordered queries, the fused gather + radiance of each variant
 * (512-key stream with interleaved or contiguous XCD chunks, 1024-key
 * stream, two-pass, inner box), with HIP events on the context's stream.
 * The fastest build + gather is returned next to the default (cell = r,
 * no tuning bits) and remembered in the context under (k, photon density
 * class: round(2 log2(n r^3 / V)), V the AABB volume). pm_tuning_lookup
 * returns the remembered choice for a (k, density) (the nearest class with
 * the same k, else the default with found = 0). Gathers flagged
 * PM_KNN_TUNED take the remembered variant; pm_server jobs whose
 * params.gather_flags hold PM_KNN_TUNED also build their grids at the
 * remembered cell size. Results do not depend on the choice: every variant
 * and cell size returns the same neighbours and radiance. */
typedef struct {
    float cell_over_r;   /* grid cell size / radius */
    uint32_t flags;      /* PM_KNN_* tuning bits of the gather */
    float ms;            /* grid rebuild + gather of the choice */
    float ms_default;    /* the same at cell = r with no tuning bits */
    int32_t found;       /* pm_tuning_lookup: 1 if a tuning was remembered for this k */
} pm_gather_tuning;

pm_status pm_tune_gather(pm_ctx* ctx, const float* d_photons9, uint32_t n, const float aabb[6],
                         const float* d_queries3, const float* d_normals3, uint32_t nq, float radius, int32_t k,
                         pm_gather_tuning* out);
pm_status pm_tuning_lookup(const pm_ctx* ctx, uint32_t n, const float aabb[6], float radius, int32_t k,
                           pm_gather_tuning* out);
/* The last pm_tune_gather's measurements on this context: h_ms[6 * 5]
 * (cell sizes in the order above x variants 0, XCD_CONTIGUOUS,
 * LARGE_BUFFER, TWO_PASS, INNER_BOX) of grid rebuild + gather in ms, -1 for
 * a configuration not measured. */
pm_status pm_tuning_table(const pm_ctx* ctx, float* h_ms, uint32_t cap);

/* Per-query candidate statistics, summed: cells C_q and photons P_q.
 * flags = PM_BOX_REFERENCE: the reference's candidate box (cell +- ceil(r/cell),
 * clamped, PhotonGrid.cu:262-272), the C_q / P_q of the algorithmic-byte
 * count. flags = 0: what the gather visits, that box pruned to the rows
 * (linear layout) or cells (Morton) that can hold a photon within r (C_q of
 * the linear layout stays the box's cell count). */
enum { PM_BOX_REFERENCE = 1u };
pm_status pm_grid_box_stats(pm_ctx* ctx, const pm_grid* grid, const float* d_queries3, uint32_t nq,
                            float radius, uint32_t flags, uint64_t* h_sum_cells, uint64_t* h_sum_cands);

/* ---- photon kd-tree: cpm::PhotonMap(Type, Photon*, size_t) ->
 * PhotonMap::fill_balanced (PhotonMap.cuh:189-283, 373-375). Same split
 * rule: plane = argmax extent (lowest axis on ties), median at
 * mid = from + (to-from)/2 of the segment sorted on that plane, children in
 * heap slots 2i+1 / 2i+2. The sort order is made total with the photon
 * index as tie-break. n_slots = 2^L - 1 with L = ceil(log2(n+1)), so no node
 * is dropped (the reference's Tree drops slots >= n, Tree.cuh:10-13). */
typedef struct {
    uint32_t n_photons;
    uint32_t n_slots;
    uint32_t levels;
    const uint32_t* d_node_idx; /* [n_slots] input photon index, 0xFFFFFFFF = empty */
    const int8_t* d_node_plane; /* [n_slots] split axis 0..2, -1 = empty */
    const float* d_photons;     /* photons9 in slot order, [9][n_slots] */
} pm_kdtree_desc;

pm_status pm_kdtree_build(pm_ctx* ctx, const float* d_photons9, uint32_t n, pm_kdtree** out);
pm_status pm_kdtree_describe(const pm_kdtree* tree, pm_kdtree_desc* out);
pm_status pm_kdtree_copy_to_host(pm_ctx* ctx, const pm_kdtree* tree, uint32_t* h_node_idx, int8_t* h_node_plane);
/* The tree's device memory goes back to the cache of the context that built
 * it, as pm_grid_destroy's does (work reading the tree on another stream
 * must have completed). */
pm_status pm_kdtree_destroy(pm_kdtree* tree);

/* k nearest photons on the kd-tree, the query side of cpm::PhotonMap
 * (PhotonMap::locate_q, PhotonMap.cuh:289-371, whose debug queue and 10000
 * sentinel are replaced by an exact search). Same contract as pm_grid_knn:
 * photons with length(p - q) <= radius, the k smallest by (FP32 d2, photon
 * index), sorted ascending; d_idx[q*k + j] is the INPUT photon index
 * (0xFFFFFFFF past the count), d_dist the FP32 distance (+inf past the
 * count), d_count[q] = min(k, photons within radius). k in [1, 256]. */
pm_status pm_kdtree_knn(pm_ctx* ctx, const pm_kdtree* tree, const float* d_queries3, uint32_t nq, float radius,
                        int32_t k, uint32_t* d_idx, float* d_dist, uint32_t* d_count);

/* ---- scene (Scene.cuh:7-25, Model.cuh:24-30, Material.cuh:4-31,
 * Light.cuh:7-31). A model is a contiguous range of triangles with one
 * material; the light source is the model whose triangles emit (the
 * reference picks the OBJ group named "light", Initializer.cu:62-101). */
typedef struct {
    uint32_t tri_begin, tri_end; /* triangles [tri_begin, tri_end) */
    float kd[3];                 /* Material::diffuse */
    float ks[3];                 /* Material::specular */
    float ke[3];                 /* Material::emission */
    float refr_index;            /* Material::refr_index */
    float opaque;                /* Material::opaque (1 = opaque) */
    float shininess;             /* Material::shininess */
    int32_t id;                  /* model id (refraction medium tracking) */
    int32_t pad;
    float aabb[6];               /* filled by the library where needed */
} pm_model;

typedef struct {
    const float* d_tris9;    /* [n_tris][9] vertices v0 v1 v2 */
    const float* d_normals9; /* [n_tris][9] vertex normals n0 n1 n2 */
    uint32_t n_tris;
    const pm_model* h_models; /* host array [n_models] */
    int32_t n_models;
    int32_t light_model;
    float light_intensity[3]; /* LightSource::intensity */
    float default_refr_index; /* GlobalParams::default_refractive_index */
    const pm_accel* accel;    /* optional, from pm_accel_build on this geometry; NULL = test every triangle */
} pm_scene;

/* ---- scene acceleration: per-model triangle BVHs, the structure the
 * reference's KdTree was designed to be (KdTree.cuh:42-157, KdTree.cu:329-375,
 * build loop never enabled) for the brute-force Model::intersection loop
 * (Model.cuh:381-429). Built on the device (Morton order, radix sort, binary
 * radix tree, bottom-up boxes) for every model with at least 16 triangles.
 * Traversal returns exactly the brute-force nearest hit (smallest t, lowest
 * triangle index on ties; boxes are conservatively inflated), so photon maps
 * and frames are bit-identical with and without it. The scene's triangles and
 * model ranges must not change while the structure is in use. */
pm_status pm_accel_build(pm_ctx* ctx, const pm_scene* scene, pm_accel** out);
pm_status pm_accel_destroy(pm_accel* accel);

typedef struct {
    uint64_t paths;     /* light paths traced */
    uint32_t n_global;  /* photons stored in the global map */
    uint32_t n_caustic; /* photons stored in the caustic map */
} pm_trace_stats;

/* ---- photon emission and bounce tracing: trace_photons(gpu_scene,
 * cpu_scene, diffuse_num, specular_num) (PhotonTracing.cu:379-422) with emit
 * / trace / destiny (:245-375). Paths are numbered and path p draws from its
 * own Philox4x32-10 stream keyed by (seed, p); the maps take the first
 * n_global / n_caustic photons in (path, event) order, so the result does not
 * depend on launch shape. Tracing stops when both maps are full or after
 * max_paths paths (the reference spins forever when a map cannot fill,
 * PhotonTracing.cu:344). Output blocks are photons9 with plane stride
 * n_global / n_caustic. Photon power = intensity / n_global for both maps
 * (PhotonTracing.cu:369). 1 <= max_bounces <= 16 (reference: 10); at most
 * 65536 models. stats->paths counts the paths traced, which may exceed the
 * last path that stored a kept photon (whole batches are traced). */
pm_status pm_trace_photons(pm_ctx* ctx, const pm_scene* scene, uint64_t seed, uint32_t n_global,
                           uint32_t n_caustic, int32_t max_bounces, uint64_t max_paths, float* d_global9,
                           float* d_caustic9, pm_trace_stats* stats);

/* ---- eye-ray frame: Raytracer::render_gpu / render_cpu / render_trace
 * (Raytracer.cuh:140-299), Camera (Camera.cuh:43-81), RaytracePlanner
 * (RaytracePlanner.cuh:9-73). Camera fields are those of the reference's
 * Camera after construction (forward/right/up normalised, viewport sizes). */
typedef struct {
    float position[3], forward[3], right[3], up[3];
    float viewport_width, viewport_height, focal_length;
    int32_t width, height, pad;
} pm_camera;

enum {
    /* Raytracer.cuh:203-204 reads the global map's power/direction with the
     * caustic map's indices; this flag reproduces that (requires
     * n_caustic <= n_global). Default: the caustic map's own arrays. */
    PM_RENDER_REFERENCE_CAUSTIC = 1u << 0
};

typedef struct {
    int32_t max_depth;     /* GlobalParams::max_depth (kernel.cu:23: 4), <= 6 */
    int32_t k_global;      /* GlobalParams::global_photon_num (kernel.cu:25) */
    float r_global;        /* Raytracer.cuh:187 (0.1) */
    int32_t k_caustic;     /* GlobalParams::caustic_photon_num (kernel.cu:26) */
    float r_caustic;       /* Raytracer.cuh:200 (0.2) */
    uint32_t flags;        /* PM_RENDER_* */
    uint32_t gather_flags; /* PM_KNN_* tuning flags for the gathers */
} pm_render_params;

typedef struct {
    uint64_t hits;   /* eye-ray hits shaded (gather queries per map) */
    uint32_t pixels;
} pm_frame_stats;

/* Renders one frame: d_rgb3 (nullable) = 3 planes of W*H HDR-mapped floats,
 * d_rgb8 (nullable) = W*H uchar3 as the reference's canvas (row y, column x
 * at y*W + x). caustic_grid may be NULL (no caustic term). */
pm_status pm_render_frame(pm_ctx* ctx, const pm_scene* scene, const pm_grid* global_grid,
                          const pm_grid* caustic_grid, const pm_camera* camera, const pm_render_params* params,
                          float* d_rgb3, uint8_t* d_rgb8, pm_frame_stats* stats);

/* ---- phased frame: pm_render_frame split so that the gathers of many
 * frames can run as one pm_gather_batch (the reference renders one frame per
 * launch, Drawer.cu:87-101):
 *   pm_frame_rays        ray trees of every pixel + compacted query block;
 *                        *n_queries = hits (synchronizes to read it back);
 *   pm_frame_gather_jobs the frame's global (and caustic, if given) gather as
 *                        batch jobs writing into the frame (not available
 *                        with PM_RENDER_REFERENCE_CAUSTIC);
 *   (run the jobs, e.g. with pm_gather_batch, on the same stream)
 *   pm_frame_finish      per-pixel fold, HDR and uchar3, as pm_render_frame.
 * A pm_frame owns its device buffers and may be reused for the next frame;
 * pm_frame_destroy returns them to the cache of the context that last grew
 * them (as pm_grid_destroy). */
pm_status pm_frame_create(pm_ctx* ctx, pm_frame** out);
pm_status pm_frame_destroy(pm_frame* frame);
pm_status pm_frame_rays(pm_ctx* ctx, pm_frame* frame, const pm_scene* scene, const pm_camera* camera,
                        const pm_render_params* params, uint32_t* n_queries);
pm_status pm_frame_gather_jobs(pm_frame* frame, const pm_grid* global_grid, const pm_grid* caustic_grid,
                               pm_gather_job* jobs /* [2] */, uint32_t* n_jobs);
pm_status pm_frame_finish(pm_ctx* ctx, pm_frame* frame, float* d_rgb3, uint8_t* d_rgb8, pm_frame_stats* stats);
/* The frame's gather queries after pm_frame_rays (hit points and normals,
 * [3][nq] each) copied into caller device buffers of at most cap queries,
 * on the context's stream; *nq = the frame's query count (copied:
 * min(nq, cap)). For tuning a renderer's gathers on its own queries. */
pm_status pm_frame_copy_queries(pm_ctx* ctx, const pm_frame* frame, float* d_queries3, float* d_normals3,
                                uint32_t cap, uint32_t* nq);

/* ---- per-frame pipeline for graph capture: the whole frame of the reference
 * (trace_photons, PhotonGrid::build of both maps, Raytracer::render_gpu:
 * Initializer.cu:142-177, Drawer.cu:18-57) as one sequence of launches on the
 * context's stream with every count kept on the device, so it can be
 * captured into a HIP graph and replayed each frame. pm_pipeline_create
 * renders one planning frame eagerly (plan_seed) to size the photon maps,
 * the grids (built at the given cell sizes), the trace's path schedule (the
 * planning frame's paths plus a quarter, at most max_paths) and the query
 * block (the planning frame's hit count: ray trees depend only on the scene
 * and the camera), and grows the context's scratch; it
 * fails with PM_ERR_CAPACITY if max_paths cannot fill the maps.
 * pm_pipeline_frame enqueues one frame for `seed` (uchar3 and/or HDR floats
 * into caller buffers, either nullable): no allocation, no host
 * synchronisation. Its maps, grids and image equal the eager calls' for the
 * same seed whenever the scheduled paths fill both maps, which
 * pm_pipeline_status reports (it synchronises the stream). One pipeline per
 * context at a time: frames share the context's scratch. The
 * reference-caustic mode is not supported. */
typedef struct {
    const pm_scene* scene;
    float aabb[6];                     /* grid bounds */
    float cell_global, cell_caustic;   /* grid cell sizes (> 0) */
    uint32_t n_global, n_caustic;
    int32_t max_bounces;
    uint64_t max_paths;
    uint64_t plan_seed;
    pm_camera camera;
    pm_render_params params;
} pm_pipeline_desc;

typedef struct {
    uint64_t scheduled_paths;   /* paths each frame traces */
    uint32_t n_global, n_caustic;  /* photons the last frame stored (capped at the map sizes) */
    uint32_t complete;          /* 1 when both maps filled */
    uint64_t hits;              /* the last frame's eye-ray hits (queries per map) */
} pm_pipeline_info;

typedef struct pm_pipeline pm_pipeline;
pm_status pm_pipeline_create(pm_ctx* ctx, const pm_pipeline_desc* desc, pm_pipeline** out);
pm_status pm_pipeline_frame(pm_pipeline* pipeline, uint64_t seed, float* d_rgb3, uint8_t* d_rgb8);
pm_status pm_pipeline_status(pm_pipeline* pipeline, pm_pipeline_info* out);
pm_status pm_pipeline_grids(const pm_pipeline* pipeline, const pm_grid** global_grid, const pm_grid** caustic_grid);
pm_status pm_pipeline_destroy(pm_pipeline* pipeline);

/* ---- serving: continuous batching of render jobs (the reference renders one
 * scene, one frame at a time, Drawer.cu:18-101). A job is a resident scene
 * (with its pm_accel, if any), photon budgets, grid cell size, camera,
 * render parameters and a number of frames; frame f re-traces its photons
 * with seed + f * 0x9E3779B97F4A7C15 and writes uchar3 to
 * d_rgb8 + f * W * H * 3. Each pm_server_step admits queued jobs whose
 * device footprint fits the budget, then renders the next frame of every
 * active job with all of their gathers in one pm_gather_batch. A job's frames
 * are identical alone and in any mix. No reference-caustic mode. */
typedef struct {
    const pm_scene* scene;
    float aabb[6];            /* grid bounds */
    float cell_size;          /* > 0, or <= 0 for the reference sizing rule */
    uint32_t n_global, n_caustic;
    int32_t max_bounces;
    uint64_t max_paths;
    uint64_t seed;
    pm_camera camera;
    pm_render_params params;
    uint32_t frames;
    uint8_t* d_rgb8;          /* [frames][H*W*3], caller-owned */
} pm_job_desc;

typedef struct {
    uint32_t queued, active, finished;
    uint64_t frames;          /* frames completed */
    uint64_t queries;         /* gather queries run (both maps) */
    uint64_t bytes_in_use;    /* admitted jobs' footprint */
} pm_server_info;

/* budget_bytes = 0: 80% of the device memory free at creation (which
 * depends on what the process holds then; pass a fixed budget for an
 * admission that does not). Admission counts each job's worst-case
 * footprint: its photon maps and grid copies (76 B per photon) plus a frame
 * of 2^min(max_depth, 6) - 1 ray-tree slots of 100 B per pixel and 16 B per
 * pixel, plus 64 MiB; a frame allocates far less (its hits only). */
pm_status pm_server_create(pm_ctx* ctx, size_t budget_bytes, pm_server** out);
pm_status pm_server_destroy(pm_server* server);
pm_status pm_server_submit(pm_server* server, const pm_job_desc* job, uint64_t* job_id);
pm_status pm_server_step(pm_server* server, uint32_t* frames_done);
/* Worker contexts for the per-job stages (trace, grids, eye rays): with
 * n_workers > 0 each step runs those stages of its jobs on n_workers host
 * threads, each driving a context on a stream of its own (created by the
 * library at the default priority), and the shared gather launch on the
 * server's context waits on all of them. 0 (the default) runs them on the
 * server's context. Frames are identical either way. */
pm_status pm_server_set_workers(pm_server* server, int32_t n_workers);
pm_status pm_server_stats(const pm_server* server, pm_server_info* out);
/* per-frame latency in seconds (eligible -> completed), oldest first */
pm_status pm_server_latencies(const pm_server* server, double* out, uint64_t cap, uint64_t* n);

/* ---- photon-map files: the binary format of load() (Initializer.cu:102-141):
 * a uint64 photon count, then per photon 9 FP32 values (pos, power,
 * inc_dir), 36 bytes, little-endian. The reference only reads this format
 * (into an AoS device array); here files load into and save from photons9
 * blocks, for persisting resident maps across runs. Files stream through a
 * pinned host buffer in chunks and are transposed on the device. The calls
 * synchronise the context's stream.
 *   pm_photon_file_count: the count stored in the file's header;
 *   pm_photon_file_read:  the file into d_photons9 (plane stride = the
 *                         file's count); PM_ERR_CAPACITY when the count
 *                         exceeds capacity, PM_ERR_INVALID when the file is
 *                         missing or truncated (then *n_read = 0);
 *   pm_photon_file_write: n photons of d_photons9 (plane stride n). */
pm_status pm_photon_file_count(pm_ctx* ctx, const char* path, uint64_t* n);
pm_status pm_photon_file_read(pm_ctx* ctx, const char* path, float* d_photons9, uint32_t capacity,
                              uint32_t* n_read);
pm_status pm_photon_file_write(pm_ctx* ctx, const char* path, const float* d_photons9, uint32_t n);

/* ---- model AABB: create_kd_tree(const Model*) (KdTree.cu:395-464), whose
 * only live effect is the model's bounding box (min/max over its vertices).
 * d_tris9: one triangle per 9 floats (v0.xyz v1.xyz v2.xyz); model m spans
 * triangles [d_tri_begin[m], d_tri_begin[m+1]). d_aabb6: [n_models][6]. */
pm_status pm_model_aabb(pm_ctx* ctx, const float* d_tris9, const uint32_t* d_tri_begin, int32_t n_models,
                        float* d_aabb6);

#ifdef __cplusplus
}
#endif
#endif
