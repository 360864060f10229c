// Test helper: loads an OBJ with the C++ loader (host/scene_io.hpp) and
// writes its arrays, and the reference camera for W x H, as raw files for
// tests/test_objload.py to compare with the Python loader.
//   objload_dump <obj> <outdir> <W> <H>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../cuda-photon-mapping_amd/host/scene_io.hpp"

static void put(const std::string& path, const void* p, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) std::exit(3);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 5) return 2;
    const std::string out = argv[2];
    try {
        pm355::HostScene s = pm355::load_obj(argv[1]);
        put(out + "/tris9.f32", s.tris9.data(), s.tris9.size() * 4);
        put(out + "/normals9.f32", s.normals9.data(), s.normals9.size() * 4);
        put(out + "/models.bin", s.models.data(), s.models.size() * sizeof(pm_model));
        std::string names;
        for (const std::string& n : s.names) names += n + "\n";
        put(out + "/names.txt", names.data(), names.size());
        put(out + "/light.i32", &s.light_model, 4);
        const pm_camera c = pm355::reference_camera(std::atoi(argv[3]), std::atoi(argv[4]));
        put(out + "/camera.bin", &c, sizeof c);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
