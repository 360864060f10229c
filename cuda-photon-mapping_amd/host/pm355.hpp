// pm355 C++ façade: the reference's entry-point names over the C ABI.
//
// Header-only. A C++ caller of the reference (Initializer.cu, Drawer.cu,
// Raytracer.cuh) swaps its includes for this header and links libpm355.so:
//   PhotonGrid(int, AABB) / build / find_nearests / copy_to_cpu
//        (PhotonGrid.cuh:44-98)              -> pm355::PhotonGrid
//   create_kd_tree(const Model*)  (KdTree.cu:461) -> pm355::create_kd_tree
//   cpm::PhotonMap(Type, Photon*, size_t) (PhotonMap.cuh:373)
//                                             -> pm355::PhotonMap
//   trace_photons(gpu, cpu, diffuse_num, specular_num) (PhotonTracing.cu:379)
//                                             -> pm355::trace_photons
//   Raytracer::render_gpu + RaytracePlanner (Raytracer.cuh:245,
//        RaytracePlanner.cu:18)              -> pm355::render_frame
//   load (Initializer.cu:102-141)            -> pm355::load_photons /
//                                               save_photons
//   loadOBJ / loadMaterial (ObjLoader.cpp:4-306), Camera (Camera.cuh:43-66)
//        and the GL canvas (Window.cuh)      -> scene_io.hpp: load_obj,
//                                               make_camera, write_ppm;
//                                               pm355::DeviceScene
// Differences from the reference are those of the C ABI: batched queries,
// borrowed inputs, status codes raised as pm355::Error (never exit()), and
// deterministic results.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pm355.h"
#include "scene_io.hpp"

namespace pm355 {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

class Context {
public:
    explicit Context(int device = 0, hipStream_t stream = nullptr) {
        int rc = pm_ctx_create(device, stream, &h_);
        if (rc != PM_OK) throw Error(rc, "pm_ctx_create failed");
    }
    ~Context() { pm_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    void check(int rc) const {
        if (rc != PM_OK) throw Error(rc, pm_last_error(h_));
    }
    void sync() const { check(pm_ctx_sync(h_)); }
    pm_ctx* get() const { return h_; }

private:
    pm_ctx* h_ = nullptr;
};

// A device buffer owned by the façade (the C ABI itself never allocates
// caller-visible buffers except inside handles).
template <class T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) : n_(n) {
        if (n && hipMalloc(&p_, n * sizeof(T)) != hipSuccess) throw Error(PM_ERR_NOMEM, "hipMalloc");
    }
    ~DeviceBuffer() {
        if (p_) (void)hipFree(p_);
    }
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    T* data() const { return p_; }
    size_t size() const { return n_; }
    std::vector<T> to_host() const {
        std::vector<T> h(n_);
        if (n_ && hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
            throw Error(PM_ERR_HIP, "hipMemcpy");
        return h;
    }
    void from_host(const std::vector<T>& h) {
        if (h.size() != n_) throw Error(PM_ERR_INVALID, "size mismatch");
        if (n_ && hipMemcpy(p_, h.data(), n_ * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw Error(PM_ERR_HIP, "hipMemcpy");
    }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// AABB with the reference's corner layout (AABB.cuh:11-12): min xyz, max xyz.
struct AABB {
    float corners[6];
};

// Host copy of a grid (PhotonGrid::copy_to_cpu, PhotonGrid.cuh:75-98).
struct HostGrid {
    pm_grid_desc desc;
    std::vector<uint32_t> cell_starts, cell_sizes, perm;
    std::vector<float> photons9;
};

// PhotonGrid (PhotonGrid.cuh:12-99).
class PhotonGrid {
public:
    PhotonGrid(Context& ctx, int num_photons, AABB scene_aabb) : ctx_(&ctx), n_(num_photons), aabb_(scene_aabb) {}
    ~PhotonGrid() { pm_grid_destroy(g_); }
    PhotonGrid(const PhotonGrid&) = delete;
    PhotonGrid& operator=(const PhotonGrid&) = delete;

    // PhotonGrid::build(photons, desired, max) (PhotonGrid.cuh:51-66). The
    // photon block is borrowed (the reference frees it). cell_size > 0
    // replaces the reference's density-based sizing.
    void build(const float* d_photons9, int desired_photons_per_cell = 1000, uint32_t /*max_photons_per_cell*/ = 5000,
               float cell_size = 0.f, int layout = PM_LAYOUT_LINEAR) {
        pm_grid_destroy(g_);
        g_ = nullptr;
        ctx_->check(pm_grid_build(ctx_->get(), d_photons9, (uint32_t)n_, aabb_.corners, cell_size,
                                  desired_photons_per_cell, layout, &g_));
    }

    // PhotonGrid::find_nearests (PhotonGrid.cu:257-284) for a batch of
    // points; results [nq][k] ascending, d_count[q] = kept photons
    // (find_nearests returned count == k).
    void find_nearests(const float* d_points3, uint32_t nq, float radius, int num_to_find, uint32_t* d_idx,
                       float* d_dist, uint32_t* d_count, uint32_t flags = 0) const {
        ctx_->check(pm_grid_knn(ctx_->get(), g_, d_points3, nq, radius, num_to_find, flags, d_idx, d_dist, d_count));
    }

    // find_nearests + Raytracer::get_color_from_photon_grid, fused.
    void gather_radiance(const float* d_points3, const float* d_normals3, uint32_t nq, float radius, int k,
                         float* d_rgb3, uint32_t* d_count = nullptr, uint32_t flags = 0) const {
        ctx_->check(pm_grid_gather_radiance(ctx_->get(), g_, d_points3, d_normals3, nq, radius, k, flags, d_rgb3,
                                            d_count));
    }

    HostGrid copy_to_cpu() const {
        HostGrid h;
        ctx_->check(pm_grid_describe(g_, &h.desc));
        h.cell_starts.resize(h.desc.n_cells);
        h.cell_sizes.resize(h.desc.n_cells);
        h.perm.resize(h.desc.n_photons);
        h.photons9.resize((size_t)h.desc.n_photons * 9);
        ctx_->check(pm_grid_copy_to_host(ctx_->get(), g_, h.cell_starts.data(), h.cell_sizes.data(),
                                         h.photons9.data(), h.perm.data()));
        return h;
    }

    pm_grid_desc describe() const {
        pm_grid_desc d;
        ctx_->check(pm_grid_describe(g_, &d));
        return d;
    }
    // PhotonGrid.cuh:71-73 accessors (device SoA planes in grid order)
    const float* get_photon_positions() const { return describe().d_photons; }
    const float* get_photon_powers() const { return describe().d_photons + 3 * (size_t)n_; }
    const float* get_photon_directions() const { return describe().d_photons + 6 * (size_t)n_; }
    const pm_grid* handle() const { return g_; }

private:
    Context* ctx_;
    int n_;
    AABB aabb_;
    pm_grid* g_ = nullptr;
};

// create_kd_tree (KdTree.cu:461): per-model AABBs of a triangle soup.
inline std::vector<AABB> create_kd_tree(Context& ctx, const float* d_tris9, const uint32_t* d_tri_begin,
                                        int n_models) {
    DeviceBuffer<float> out((size_t)n_models * 6);
    ctx.check(pm_model_aabb(ctx.get(), d_tris9, d_tri_begin, n_models, out.data()));
    ctx.sync();
    std::vector<float> h = out.to_host();
    std::vector<AABB> r(n_models);
    for (int m = 0; m < n_models; ++m)
        for (int c = 0; c < 6; ++c) r[m].corners[c] = h[(size_t)m * 6 + c];
    return r;
}

// cpm::PhotonMap (PhotonMap.cuh:19-379): balanced photon kd-tree.
class PhotonMap {
public:
    enum Type { def = 0, caustic };
    PhotonMap(Context& ctx, Type type, const float* d_photons9, size_t size) : ctx_(&ctx), type_(type) {
        ctx.check(pm_kdtree_build(ctx.get(), d_photons9, (uint32_t)size, &t_));
        ctx.check(pm_kdtree_describe(t_, &desc_));
    }
    ~PhotonMap() { pm_kdtree_destroy(t_); }
    PhotonMap(const PhotonMap&) = delete;
    PhotonMap& operator=(const PhotonMap&) = delete;
    // heap-slot view of the tree (Tree.cuh:7-118): node i -> photon index,
    // plane; children 2i+1, 2i+2
    void nodes(std::vector<uint32_t>& idx, std::vector<int8_t>& plane) const {
        idx.resize(desc_.n_slots);
        plane.resize(desc_.n_slots);
        ctx_->check(pm_kdtree_copy_to_host(ctx_->get(), t_, idx.data(), plane.data()));
    }
    // k nearest photons within radius, the query side of PhotonMap::locate_q
    // (PhotonMap.cuh:289-371); d_idx holds input photon indices
    void find_nearests(const float* d_queries3, uint32_t nq, float radius, int k, uint32_t* d_idx, float* d_dist,
                       uint32_t* d_count) const {
        ctx_->check(pm_kdtree_knn(ctx_->get(), t_, d_queries3, nq, radius, k, d_idx, d_dist, d_count));
    }
    const pm_kdtree_desc& desc() const { return desc_; }
    Type type() const { return type_; }

private:
    Context* ctx_;
    Type type_;
    pm_kdtree* t_ = nullptr;
    pm_kdtree_desc desc_{};
};

// Per-model triangle BVHs of a scene (the per-model KdTree of
// KdTree.cuh:42-157 as designed). Attach with scene.accel = accel.get() while
// the scene's triangles stay unchanged; results are identical without it.
class SceneAccel {
public:
    SceneAccel(Context& ctx, const pm_scene& scene) { ctx.check(pm_accel_build(ctx.get(), &scene, &a_)); }
    ~SceneAccel() { pm_accel_destroy(a_); }
    SceneAccel(const SceneAccel&) = delete;
    SceneAccel& operator=(const SceneAccel&) = delete;
    const pm_accel* get() const { return a_; }

private:
    pm_accel* a_ = nullptr;
};

// A HostScene (scene_io.hpp: load_obj) in device memory as a pm_scene
// (initialize_scene_and_photon_maps, Initializer.cu:179-201): triangles and
// normals copied once, model records kept on the host, light intensity 1
// and default IOR 1 (Light.cuh:8, kernel.cu:24); optional per-model BVHs.
class DeviceScene {
public:
    DeviceScene(Context& ctx, const HostScene& s, bool bvh = true)
        : tris_(s.tris9.size()), normals_(s.normals9.size()), models_(s.models) {
        tris_.from_host(s.tris9);
        normals_.from_host(s.normals9);
        rec_ = pm_scene{};
        rec_.d_tris9 = tris_.data();
        rec_.d_normals9 = normals_.data();
        rec_.n_tris = s.n_tris();
        rec_.h_models = models_.data();
        rec_.n_models = (int32_t)models_.size();
        rec_.light_model = s.light_model;
        for (int a = 0; a < 3; ++a) rec_.light_intensity[a] = 1.f;
        rec_.default_refr_index = 1.f;
        if (bvh) {
            accel_ = std::make_unique<SceneAccel>(ctx, rec_);
            rec_.accel = accel_->get();
        }
    }
    const pm_scene& get() const { return rec_; }

private:
    DeviceBuffer<float> tris_, normals_;
    std::vector<pm_model> models_;
    pm_scene rec_{};
    std::unique_ptr<SceneAccel> accel_;
};

// trace_photons (PhotonTracing.cu:379-422): returns the global ("diffuse")
// and caustic ("specular") photon blocks.
struct PhotonArrays {
    DeviceBuffer<float> global9, caustic9;
    pm_trace_stats stats;
};
inline PhotonArrays trace_photons(Context& ctx, const pm_scene& scene, uint32_t diffuse_num, uint32_t specular_num,
                                  uint64_t seed = 1234, int max_bounces = 10, uint64_t max_paths = 1ull << 40) {
    PhotonArrays a{DeviceBuffer<float>((size_t)diffuse_num * 9), DeviceBuffer<float>((size_t)specular_num * 9), {}};
    ctx.check(pm_trace_photons(ctx.get(), &scene, seed, diffuse_num, specular_num, max_bounces, max_paths,
                               a.global9.data(), a.caustic9.data(), &a.stats));
    return a;
}

// load (Initializer.cu:102-141): a photon-map file (u64 count, 36-byte
// records pos/power/inc_dir) into a device photons9 block; save_photons is
// its writer, for persisting resident maps.
inline DeviceBuffer<float> load_photons(Context& ctx, const std::string& path, uint32_t& count) {
    uint64_t n = 0;
    ctx.check(pm_photon_file_count(ctx.get(), path.c_str(), &n));
    if (n > 0xFFFFFFFFull) throw Error(PM_ERR_UNSUPPORTED, "photon file too large: " + path);
    DeviceBuffer<float> block((size_t)n * 9);
    ctx.check(pm_photon_file_read(ctx.get(), path.c_str(), block.data(), (uint32_t)n, &count));
    return block;
}
inline void save_photons(Context& ctx, const std::string& path, const float* d_photons9, uint32_t count) {
    ctx.check(pm_photon_file_write(ctx.get(), path.c_str(), d_photons9, count));
}

// Raytracer::render_gpu with the RaytracePlanner (Raytracer.cuh:245-279).
inline pm_frame_stats render_frame(Context& ctx, const pm_scene& scene, const PhotonGrid& global,
                                   const PhotonGrid* caustic, const pm_camera& camera, const pm_render_params& params,
                                   float* d_rgb3, uint8_t* d_rgb8) {
    pm_frame_stats st{};
    ctx.check(pm_render_frame(ctx.get(), &scene, global.handle(), caustic ? caustic->handle() : nullptr, &camera,
                              &params, d_rgb3, d_rgb8, &st));
    return st;
}

}  // namespace pm355
