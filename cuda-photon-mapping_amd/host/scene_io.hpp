// Host-side scene input and image output for C++ callers of pm355: the
// reference's OBJ/MTL loader (ObjLoader.cpp:4-306), its camera
// (Camera.cuh:43-66) and a headless replacement for the GL canvas (PPM).
//
// Plain C++17 over the C ABI's structs (no HIP): HostScene holds the
// triangle soup, vertex normals and one pm_model per OBJ group, ready for
// pm355::DeviceScene (pm355.hpp) or any caller filling a pm_scene.
//
// Loader rules (those of the reference loop, with its failure cases given a
// defined meaning):
//  * a model is the run of consecutive `f` lines; the first non-`f` line
//    after it closes it (ObjLoader.cpp:169-186), as does the end of the file;
//  * it is named by the last `g` / `o` (default "n1"); a taken name gets one
//    '1' appended and the running name keeps the suffix (:174-177);
//  * `usemtl` resets the material to Material() (diffuse 1, specular 1,
//    emission 0, shininess 16, IOR 1, opaque 1; Material.cuh:22-30) and
//    applies the first `newmtl` block of that name (Kd, Ks, Ke, Ns, Ni, `d`,
//    `Tr` as opaque = 1 - Tr); a model with no `usemtl` keeps the previous
//    model's material (:179-181);
//  * indices are 1-based, negative ones count back from the vertices read
//    (:96-112); a line of whitespace repeats the previous keyword with no
//    operands (the last values parsed are reused); lines end at '\n' only;
//  * where the reference misreads (`f v v v` / `f v/vt` faces, faces
//    without normals, mixed triangles and quads) the OBJ meaning is used:
//    positions, the face's geometric normal, each face by its own arity;
//    quads split as (v0 v1 v3) (v1 v2 v3) (Model.cuh:318-340); faces of 5+
//    vertices, index 0 and out-of-range indices throw; a bare `f` line
//    makes an empty model that takes a name but is dropped.
// The photon source is the model named exactly "light" (Initializer.cu:62-84).
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pm355.h"

namespace pm355 {

struct HostScene {
    std::vector<float> tris9;     // [n_tris][9] vertices v0 v1 v2
    std::vector<float> normals9;  // [n_tris][9] vertex normals
    std::vector<pm_model> models;
    std::vector<std::string> names;
    int light_model = -1;
    uint32_t n_tris() const { return (uint32_t)(tris9.size() / 9); }
    // union of the vertices (Initializer.cu:197-200): min xyz, max xyz
    void aabb(float out[6]) const {
        for (int a = 0; a < 3; ++a) {
            out[a] = INFINITY;
            out[3 + a] = -INFINITY;
        }
        for (size_t i = 0; i < tris9.size(); ++i) {
            const int a = (int)(i % 3);
            out[a] = std::fmin(out[a], tris9[i]);
            out[3 + a] = std::fmax(out[3 + a], tris9[i]);
        }
    }
};

namespace objio {

struct Material {
    float diffuse[3] = {1.f, 1.f, 1.f}, specular[3] = {1.f, 1.f, 1.f}, emission[3] = {0.f, 0.f, 0.f};
    float shininess = 16.f, refr_index = 1.f, opaque = 1.f;
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f' || c == '\n'; }

inline std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && is_space(s[i])) ++i;
        size_t j = i;
        while (j < s.size() && !is_space(s[j])) ++j;
        if (j > i) t.push_back(s.substr(i, j - i));
        i = j;
    }
    return t;
}

// (keyword, operands) of one line; see the header note on blank lines.
inline void split_line(const std::string& line, std::string& prefix, std::vector<std::string>& ops) {
    ops.clear();
    if (line.empty()) {
        prefix.clear();
        return;
    }
    std::vector<std::string> t = split_ws(line);
    if (t.empty()) return;  // whitespace only: the keyword stays
    prefix = t[0];
    ops.assign(t.begin() + 1, t.end());
}

inline float to_f32(const std::string& s) {
    char* end = nullptr;
    const float v = std::strtof(s.c_str(), &end);
    if (end == s.c_str()) throw std::runtime_error("objload: bad number '" + s + "'");
    return v;
}

// three floats, or the stale value when the line has no operands
inline void read_vec3(const std::vector<std::string>& ops, float v[3]) {
    if (ops.empty()) return;
    if (ops.size() < 3) throw std::runtime_error("objload: malformed vector operands");
    for (int a = 0; a < 3; ++a) v[a] = to_f32(ops[a]);
}

inline std::vector<std::string> read_lines(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("objload: cannot open " + path);
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) lines.push_back(line);
    return lines;
}

// Material `name` of an MTL file over Material() (ObjLoader.cpp:4-83).
inline Material load_material(const std::string& mtl_path, const std::string& name) {
    Material m;
    std::string prefix;
    std::vector<std::string> ops;
    float tv[3] = {0.f, 0.f, 0.f}, tf = 0.f;
    bool processing = false;
    for (const std::string& line : read_lines(mtl_path)) {
        split_line(line, prefix, ops);
        if (prefix == "#") continue;
        if (prefix == "newmtl") {
            if (processing) break;
            if ((ops.empty() ? std::string() : ops[0]) != name) continue;
            processing = true;
        } else if (!processing) {
            continue;
        } else if (prefix == "Ka" || prefix == "Kd" || prefix == "Ks" || prefix == "Ke") {
            read_vec3(ops, tv);
            float* dst = prefix == "Kd" ? m.diffuse : prefix == "Ks" ? m.specular : prefix == "Ke" ? m.emission : nullptr;
            if (dst)
                for (int a = 0; a < 3; ++a) dst[a] = tv[a];
        } else if (prefix == "d" || prefix == "Tr" || prefix == "Ns" || prefix == "Ni") {
            if (!ops.empty()) tf = to_f32(ops[0]);
            if (prefix == "d")
                m.opaque = tf;
            else if (prefix == "Tr")
                m.opaque = 1.f - tf;
            else if (prefix == "Ns")
                m.shininess = tf;
            else
                m.refr_index = tf;
        }
    }
    return m;
}

inline int index_of(const std::string& tok, size_t n, const char* what) {
    char* end = nullptr;
    const long i = std::strtol(tok.c_str(), &end, 10);
    if (end == tok.c_str() || *end) throw std::runtime_error(std::string("objload: bad ") + what + " index '" + tok + "'");
    if (i == 0) throw std::runtime_error(std::string("objload: ") + what + " index 0 (OBJ indices are 1-based)");
    const long j = i > 0 ? i - 1 : (long)n + i;
    if (j < 0 || j >= (long)n) throw std::runtime_error(std::string("objload: ") + what + " index out of range");
    return (int)j;
}

}  // namespace objio

// loadOBJ (ObjLoader.cpp:123-306): the OBJ at `path` and the MTL it names.
inline HostScene load_obj(const std::string& path) {
    using namespace objio;
    const size_t slash = path.find_last_of('/');
    const std::string folder = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
    struct Face {
        int p[3];
        int n[3];  // -1 = none
    };
    struct Open {
        std::string name;
        std::vector<Face> faces;
        Material mat;
    };
    std::vector<std::array<float, 3>> pos, nrm;
    std::vector<Open> models;
    std::set<std::string> seen;
    std::map<std::string, Material> cache;
    std::string prefix, mtl, cur_name = "n1";
    std::vector<std::string> ops;
    std::vector<Face> faces;
    Material cur_mat;
    bool finished = false;
    float tv[3] = {0.f, 0.f, 0.f};

    auto close_model = [&]() {
        if (seen.count(cur_name)) cur_name += '1';
        seen.insert(cur_name);
        if (!faces.empty()) models.push_back({cur_name, faces, cur_mat});
        faces.clear();
    };
    for (const std::string& line : read_lines(path)) {
        split_line(line, prefix, ops);
        if (finished && prefix != "f") {
            close_model();
            finished = false;
        }
        if (prefix == "mtllib") {
            if (!ops.empty()) mtl = ops[0];
        } else if (prefix == "g" || prefix == "o") {
            if (!ops.empty()) cur_name = ops[0];
        } else if (prefix == "usemtl") {
            const std::string name = ops.empty() ? std::string() : ops[0];
            if (mtl.empty()) throw std::runtime_error("objload: usemtl before mtllib");
            auto it = cache.find(name);
            if (it == cache.end()) it = cache.emplace(name, load_material(folder + "/" + mtl, name)).first;
            cur_mat = it->second;
        } else if (prefix == "v" || prefix == "vn") {
            read_vec3(ops, tv);
            (prefix == "v" ? pos : nrm).push_back({tv[0], tv[1], tv[2]});
        } else if (prefix == "f") {
            finished = true;
            if (ops.empty()) continue;
            if (ops.size() > 4) throw std::runtime_error("objload: faces of more than 4 vertices are not supported");
            if (ops.size() < 3) throw std::runtime_error("objload: face with fewer than 3 vertices");
            int pi[4], ni[4];
            for (size_t v = 0; v < ops.size(); ++v) {
                const std::string& t = ops[v];
                const size_t s1 = t.find('/');
                pi[v] = index_of(t.substr(0, s1), pos.size(), "position");
                ni[v] = -1;
                if (s1 != std::string::npos) {
                    const size_t s2 = t.find('/', s1 + 1);
                    if (s2 != std::string::npos && s2 + 1 < t.size())
                        ni[v] = index_of(t.substr(s2 + 1, t.find('/', s2 + 1) - s2 - 1), nrm.size(), "normal");
                }
            }
            if (ops.size() == 3) {
                faces.push_back({{pi[0], pi[1], pi[2]}, {ni[0], ni[1], ni[2]}});
            } else {  // Model.cuh:318-340
                faces.push_back({{pi[0], pi[1], pi[3]}, {ni[0], ni[1], ni[3]}});
                faces.push_back({{pi[1], pi[2], pi[3]}, {ni[1], ni[2], ni[3]}});
            }
        }
    }
    if (finished) close_model();
    if (models.empty()) throw std::runtime_error("objload: " + path + ": no faces");

    HostScene S;
    uint32_t t0 = 0;
    for (size_t mi = 0; mi < models.size(); ++mi) {
        const Open& M = models[mi];
        for (const Face& f : M.faces) {
            float v[3][3];
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) v[c][a] = pos[f.p[c]][a];
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) S.tris9.push_back(v[c][a]);
            float g[3] = {0.f, 0.f, 0.f};
            if (f.n[0] < 0 || f.n[1] < 0 || f.n[2] < 0) {
                // geometric normal: float edges, double cross and length
                double e1[3], e2[3];
                for (int a = 0; a < 3; ++a) {
                    e1[a] = (double)(v[1][a] - v[0][a]);
                    e2[a] = (double)(v[2][a] - v[0][a]);
                }
                const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                     e1[0] * e2[1] - e1[1] * e2[0]};
                const double ln = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                for (int a = 0; a < 3; ++a) g[a] = (float)(ln > 0 ? c[a] / ln : c[a]);
            }
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) S.normals9.push_back(f.n[c] >= 0 ? nrm[f.n[c]][a] : g[a]);
        }
        pm_model m{};
        m.tri_begin = t0;
        m.tri_end = t0 + (uint32_t)M.faces.size();
        for (int a = 0; a < 3; ++a) {
            m.kd[a] = M.mat.diffuse[a];
            m.ks[a] = M.mat.specular[a];
            m.ke[a] = M.mat.emission[a];
        }
        m.refr_index = M.mat.refr_index;
        m.opaque = M.mat.opaque;
        m.shininess = M.mat.shininess;
        m.id = (int32_t)mi;
        S.models.push_back(m);
        S.names.push_back(M.name);
        if (M.name == "light") S.light_model = (int)mi;
        t0 = m.tri_end;
    }
    if (S.light_model < 0) throw std::runtime_error("objload: " + path + ": no model named 'light'");
    return S;
}

// Camera(pos, dir, up, fov, W, H) (Camera.cuh:43-66) in float, the
// reference's host order; tan is taken in double and rounded once.
inline pm_camera make_camera(const float pos[3], const float dir[3], const float up[3], float vfov_deg, int width,
                             int height) {
    auto norm = [](float v[3]) {
        if (v[0] == 0.f && v[1] == 0.f && v[2] == 0.f) return;
        const float ln = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        for (int a = 0; a < 3; ++a) v[a] = v[a] / ln;
    };
    auto cross = [](const float a[3], const float b[3], float o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = -(a[0] * b[2] - a[2] * b[0]);
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    pm_camera c{};
    const float aspect = (float)width / (float)height;
    const float theta = (vfov_deg * 3.14159265358979f) / 180.f;
    const float h = (float)std::tan((double)(theta / 2.f));
    const float vh = 2.f * h;
    float fwd[3] = {dir[0], dir[1], dir[2]}, right[3], upv[3];
    norm(fwd);
    cross(fwd, up, right);
    norm(right);
    cross(right, fwd, upv);
    for (int a = 0; a < 3; ++a) {
        c.position[a] = pos[a];
        c.forward[a] = fwd[a];
        c.right[a] = right[a];
        c.up[a] = upv[a];
    }
    c.viewport_width = aspect * vh;
    c.viewport_height = vh;
    c.focal_length = 1.f;
    c.width = width;
    c.height = height;
    return c;
}

// The reference's camera for CornellBox-Sphere (kernel.cu:29-30).
inline pm_camera reference_camera(int width = 1024, int height = 768) {
    const float pos[3] = {-0.00999999046f, 0.795000017f, 2.35000001f}, dir[3] = {0.f, 0.f, -1.f},
                up[3] = {0.f, 1.f, 0.f};
    return make_camera(pos, dir, up, 60.f, width, height);
}

// Binary PPM (P6) of a W*H uchar3 canvas whose row 0 is the bottom of the
// picture (Camera.cuh:67-81 maps pixel_y = 0 to the lower viewport edge).
inline void write_ppm(const std::string& path, const uint8_t* rgb8, int width, int height) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("write_ppm: cannot open " + path);
    std::fprintf(f, "P6\n%d %d\n255\n", width, height);
    for (int y = height - 1; y >= 0; --y)
        if (std::fwrite(rgb8 + (size_t)y * width * 3, 1, (size_t)width * 3, f) != (size_t)width * 3) {
            std::fclose(f);
            throw std::runtime_error("write_ppm: short write");
        }
    std::fclose(f);
}

}  // namespace pm355
