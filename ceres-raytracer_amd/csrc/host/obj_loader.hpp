// obj_loader.hpp -- Wavefront OBJ reader with the reference loader's semantics.
//
// Follows lib/model_loaders/obj.hpp:57-125:
//   * `v x y z` parsed with strtof (FP32 values, then widened);
//   * `f a b c ...` fan-triangulated (a,b,c), (a,c,d), ...; indices may be
//     negative (relative) and may carry /vt/vn parts, which are ignored;
//   * a vertex normal is the normalised sum of the unnormalised geometric
//     normals (p0-p1)x(p2-p0) of the triangles using that vertex (these point
//     inwards for counter-clockwise outward meshes, like the reference).
// Differences (stricter, documented): lines of any length are accepted (the
// reference stops at the first line longer than 1023 characters) and an
// out-of-range face index raises instead of being undefined behaviour.
#pragma once
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace crt_host {

struct Mesh {
    std::vector<double> v;        // 3 per vertex (FP32 values widened to FP64)
    std::vector<uint32_t> f;      // 3 per triangle
    std::vector<double> vn;       // 3 per vertex
    size_t n_vertices() const { return v.size() / 3; }
    size_t n_tris() const { return f.size() / 3; }
};

inline const char* skip_ws(const char* p) {
    while (*p && std::isspace((unsigned char)*p)) ++p;
    return p;
}

// One face index: "i", "i/t", "i//n", "i/t/n"; returns false at end of line.
inline bool read_face_index(const char*& p, long& out) {
    p = skip_ws(p);
    if (!std::isdigit((unsigned char)*p) && *p != '-') return false;
    char* end = nullptr;
    out = std::strtol(p, &end, 10);
    p = skip_ws(end);
    if (*p == '/') {
        ++p;
        if (*p != '/') { std::strtol(p, &end, 10); p = end; }
        p = skip_ws(p);
        if (*p == '/') { ++p; std::strtol(p, &end, 10); p = end; }
    }
    return true;
}

inline Mesh load_obj(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open OBJ file: " + path);
    Mesh m;
    std::string line;
    while (std::getline(in, line)) {
        const char* p = skip_ws(line.c_str());
        if (*p == '\0' || *p == '#') continue;
        if (p[0] == 'v' && std::isspace((unsigned char)p[1])) {
            char* end = nullptr;
            const float x = std::strtof(p + 1, &end);
            const float y = std::strtof(end, &end);
            const float z = std::strtof(end, &end);
            m.v.push_back(x);
            m.v.push_back(y);
            m.v.push_back(z);
        } else if (p[0] == 'f' && std::isspace((unsigned char)p[1])) {
            p += 2;
            long idx;
            uint32_t first = 0, prev = 0;
            int k = 0;
            while (read_face_index(p, idx)) {
                const long nv = (long)m.n_vertices();
                const long j = idx < 0 ? nv + idx : idx - 1;
                if (j < 0 || j >= nv) throw std::runtime_error("OBJ face index out of range in " + path);
                const uint32_t u = (uint32_t)j;
                if (k == 0) first = u;
                else if (k >= 2) { m.f.push_back(first); m.f.push_back(prev); m.f.push_back(u); }
                prev = u;
                ++k;
            }
        }
    }
    // area-weighted (unnormalised cross product) vertex normals, inward
    m.vn.assign(m.v.size(), 0.0);
    for (size_t t = 0; t < m.n_tris(); ++t) {
        const uint32_t a = m.f[3 * t], b = m.f[3 * t + 1], c = m.f[3 * t + 2];
        const double e1[3] = {m.v[3 * a] - m.v[3 * b], m.v[3 * a + 1] - m.v[3 * b + 1], m.v[3 * a + 2] - m.v[3 * b + 2]};
        const double e2[3] = {m.v[3 * c] - m.v[3 * a], m.v[3 * c + 1] - m.v[3 * a + 1], m.v[3 * c + 2] - m.v[3 * a + 2]};
        const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        for (uint32_t id : {a, b, c})
            for (int k = 0; k < 3; ++k) m.vn[3 * id + k] += n[k];
    }
    for (size_t i = 0; i < m.n_vertices(); ++i) {
        const double l = std::sqrt(m.vn[3 * i] * m.vn[3 * i] + m.vn[3 * i + 1] * m.vn[3 * i + 1] +
                                   m.vn[3 * i + 2] * m.vn[3 * i + 2]);
        if (l > 0)
            for (int k = 0; k < 3; ++k) m.vn[3 * i + k] /= l;
    }
    return m;
}

}  // namespace crt_host
