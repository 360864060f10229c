// Drives the C++ façade the way the reference's Initializer / tests do:
// build a PhotonGrid from a photon block, query find_nearests, build a
// PhotonMap and query it, and write every result to <outdir> as raw little-endian arrays
// for the parity test to compare with the oracle.
//   facade_demo <outdir> <n_photons> <n_queries> <seed>
// and, in `render` mode, the reference application's frame on an OBJ scene.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "pm355.hpp"

static void write_file(const std::string& path, const void* p, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

// The reference application's flow on an OBJ scene (kernel.cu:8-40,
// Initializer.cu:179-201, Drawer.cu:18-101) through the façade: load_obj,
// DeviceScene with BVHs, trace_photons, both PhotonGrids with the reference
// sizing, render_frame with the reference's k / radii / depth and camera;
// writes the photon maps, the uchar3 canvas and frame.ppm.
//   facade_demo render <obj> <outdir> <W> <H> <n_global> <n_caustic> <seed>
static int render_main(int argc, char** argv) {
    if (argc != 9) {
        std::fprintf(stderr, "usage: %s render obj outdir W H n_global n_caustic seed\n", argv[0]);
        return 2;
    }
    const std::string out = argv[3];
    const int W = std::atoi(argv[4]), H = std::atoi(argv[5]);
    const uint32_t ng = (uint32_t)std::strtoul(argv[6], nullptr, 10), nc = (uint32_t)std::strtoul(argv[7], nullptr, 10);
    const uint64_t seed = std::strtoull(argv[8], nullptr, 10);
    try {
        const pm355::HostScene hs = pm355::load_obj(argv[2]);
        pm355::Context ctx(0);
        pm355::DeviceScene ds(ctx, hs, true);
        pm355::PhotonArrays maps = pm355::trace_photons(ctx, ds.get(), ng, nc, seed, 10);
        pm355::AABB box;
        hs.aabb(box.corners);
        pm355::PhotonGrid gg(ctx, (int)ng, box), gc(ctx, (int)nc, box);
        gg.build(maps.global9.data());
        gc.build(maps.caustic9.data());
        const pm_camera cam = pm355::reference_camera(W, H);
        pm_render_params p{};
        p.max_depth = 4;
        p.k_global = 2000;
        p.r_global = 0.1f;
        p.k_caustic = 400;
        p.r_caustic = 0.2f;
        pm355::DeviceBuffer<uint8_t> rgb8((size_t)W * H * 3);
        const pm_frame_stats st = pm355::render_frame(ctx, ds.get(), gg, &gc, cam, p, nullptr, rgb8.data());
        ctx.sync();
        const std::vector<uint8_t> h8 = rgb8.to_host();
        const std::vector<float> g9 = maps.global9.to_host(), c9 = maps.caustic9.to_host();
        // persist both maps in the reference's photon-map file format and
        // load the global one back (Initializer.cu:102-141)
        pm355::save_photons(ctx, out + "/global_map.bin", maps.global9.data(), maps.stats.n_global);
        pm355::save_photons(ctx, out + "/caustic_map.bin", maps.caustic9.data(), maps.stats.n_caustic);
        uint32_t n_back = 0;
        const pm355::DeviceBuffer<float> back = pm355::load_photons(ctx, out + "/global_map.bin", n_back);
        if (n_back != maps.stats.n_global || back.to_host() != g9) {
            std::fprintf(stderr, "photon file round trip differs\n");
            return 1;
        }
        write_file(out + "/global9.f32", g9.data(), g9.size() * 4);
        write_file(out + "/caustic9.f32", c9.data(), c9.size() * 4);
        write_file(out + "/rgb8.u8", h8.data(), h8.size());
        pm355::write_ppm(out + "/frame.ppm", h8.data(), W, H);
        std::printf("facade_demo render ok: models=%zu tris=%u global=%u caustic=%u hits=%llu\n", hs.models.size(),
                    hs.n_tris(), maps.stats.n_global, maps.stats.n_caustic, (unsigned long long)st.hits);
    } catch (const pm355::Error& e) {
        std::fprintf(stderr, "pm355 error %d: %s\n", e.code, e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::string(argv[1]) == "render") return render_main(argc, argv);
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s outdir n_photons n_queries seed\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const uint32_t n = (uint32_t)std::strtoul(argv[2], nullptr, 10);
    const uint32_t nq = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    uint64_t s = std::strtoull(argv[4], nullptr, 10) * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {  // xorshift64*, uniform in [0, 1)
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return (float)((s * 0x2545F4914F6CDD1Dull) >> 40) * (1.0f / 16777216.0f);
    };
    std::vector<float> ph((size_t)n * 9), q((size_t)nq * 3);
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) ph[(size_t)a * n + i] = rnd();
        ph[(size_t)(1 + (i % 3)) * n + i] = 0.f;  // photons on three planes
        for (int a = 3; a < 6; ++a) ph[(size_t)a * n + i] = 1e-6f * rnd();
        ph[(size_t)6 * n + i] = 0.f;
        ph[(size_t)7 * n + i] = -1.f;
        ph[(size_t)8 * n + i] = 0.f;
    }
    for (uint32_t i = 0; i < nq; ++i)
        for (int a = 0; a < 3; ++a) q[(size_t)a * nq + i] = rnd();
    try {
        pm355::Context ctx(0);
        pm355::DeviceBuffer<float> d_ph(ph.size()), d_q(q.size());
        d_ph.from_host(ph);
        d_q.from_host(q);
        pm355::AABB box{{0.f, 0.f, 0.f, 1.f, 1.f, 1.f}};
        pm355::PhotonGrid grid(ctx, (int)n, box);
        grid.build(d_ph.data(), 1000, 5000, 0.03f);
        const int k = 32;
        const float r = 0.05f;
        pm355::DeviceBuffer<uint32_t> idx((size_t)nq * k), cnt(nq);
        pm355::DeviceBuffer<float> dist((size_t)nq * k);
        grid.find_nearests(d_q.data(), nq, r, k, idx.data(), dist.data(), cnt.data());
        pm355::PhotonMap map(ctx, pm355::PhotonMap::def, d_ph.data(), n);
        std::vector<uint32_t> node_idx;
        std::vector<int8_t> node_plane;
        map.nodes(node_idx, node_plane);
        pm355::DeviceBuffer<uint32_t> kidx((size_t)nq * k), kcnt(nq);
        pm355::DeviceBuffer<float> kdist((size_t)nq * k);
        map.find_nearests(d_q.data(), nq, r, k, kidx.data(), kdist.data(), kcnt.data());
        ctx.sync();
        pm355::HostGrid hg = grid.copy_to_cpu();
        write_file(out + "/photons.f32", ph.data(), ph.size() * 4);
        write_file(out + "/queries.f32", q.data(), q.size() * 4);
        auto hi = idx.to_host();
        auto hd = dist.to_host();
        auto hc = cnt.to_host();
        write_file(out + "/knn_idx.u32", hi.data(), hi.size() * 4);
        write_file(out + "/knn_dist.f32", hd.data(), hd.size() * 4);
        write_file(out + "/knn_count.u32", hc.data(), hc.size() * 4);
        write_file(out + "/kd_idx.u32", node_idx.data(), node_idx.size() * 4);
        write_file(out + "/kd_plane.i8", node_plane.data(), node_plane.size());
        auto ki = kidx.to_host();
        auto kd = kdist.to_host();
        auto kc = kcnt.to_host();
        write_file(out + "/kdknn_idx.u32", ki.data(), ki.size() * 4);
        write_file(out + "/kdknn_dist.f32", kd.data(), kd.size() * 4);
        write_file(out + "/kdknn_count.u32", kc.data(), kc.size() * 4);
        write_file(out + "/grid_starts.u32", hg.cell_starts.data(), hg.cell_starts.size() * 4);
        write_file(out + "/grid_sizes.u32", hg.cell_sizes.data(), hg.cell_sizes.size() * 4);
        std::printf("facade_demo ok: n=%u nq=%u cells=%u slots=%u\n", n, nq, hg.desc.n_cells, map.desc().n_slots);
    } catch (const pm355::Error& e) {
        std::fprintf(stderr, "pm355 error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
