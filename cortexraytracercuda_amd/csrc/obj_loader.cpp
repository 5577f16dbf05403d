// obj_loader.cpp - Wavefront OBJ reader producing an indexed crt_mesh-ready triangle list.
//
// The reference reads OBJ files with the vendored tinyobjloader (core/src/utilities/mesh_loader.cuh:94-146)
// and de-indexes every face into three vertices (core/src/gpu.cu:1776-1878). Here each distinct
// (position, texcoord, normal) corner becomes one vertex of an indexed mesh, polygons are fanned
// into triangles, and negative (relative) indices are resolved. Texture coordinates are read as
// pairs (the reference's stride-3 read of the texcoord array, mesh_loader.cuh:370-371, is not
// reproduced) and normals are kept only when every corner has one.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "cortex_rt.h"
#include "crt_runtime.h"

namespace crt {
namespace {

struct Corner { int v, t, n; };
struct CornerHash {
    size_t operator()(const Corner& c) const {
        uint64_t h = (uint64_t)(uint32_t)c.v * 0x9e3779b97f4a7c15ULL;
        h ^= (uint64_t)(uint32_t)c.t * 0xc2b2ae3d27d4eb4fULL + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)c.n * 0x165667b19e3779f9ULL + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
struct CornerEq {
    bool operator()(const Corner& a, const Corner& b) const { return a.v == b.v && a.t == b.t && a.n == b.n; }
};

// Parses "v", "v/t", "v//n", "v/t/n"; returns false on a malformed token.
bool parse_corner(const char*& p, int nv, int nt, int nn, Corner& c) {
    auto num = [](const char*& q, int count, int& out) -> bool {
        char* end;
        long x = std::strtol(q, &end, 10);
        if (end == q) return false;
        q = end;
        if (x < 0) x = count + x; else x = x - 1;   // 1-based, negative = relative to the end
        out = (int)x;
        return true;
    };
    c.v = c.t = c.n = -1;
    if (!num(p, nv, c.v)) return false;
    if (*p == '/') {
        ++p;
        if (*p != '/') { if (!num(p, nt, c.t)) return false; }
        if (*p == '/') { ++p; if (!num(p, nn, c.n)) return false; }
    }
    return true;
}

}  // namespace
}  // namespace crt

using namespace crt;

extern "C" {

int crt_obj_load(const char* path, crt_obj_mesh* out) {
    if (!path || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof(*out));
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(CRT_ERR_IO, std::string("cannot open ") + path);
    std::vector<float> P, T, N;
    std::vector<Corner> corners;     // 3 per triangle
    std::vector<Corner> poly;
    char line[4096];
    int lineno = 0;
    bool bad = false;
    std::string why;
    while (std::fgets(line, sizeof(line), f)) {
        ++lineno;
        const char* p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
            float x = 0, y = 0, z = 0;
            if (std::sscanf(p + 2, "%f %f %f", &x, &y, &z) != 3) { bad = true; why = "bad vertex"; break; }
            P.push_back(x); P.push_back(y); P.push_back(z);
        } else if (p[0] == 'v' && p[1] == 't' && (p[2] == ' ' || p[2] == '\t')) {
            float u = 0, v = 0;
            int k = std::sscanf(p + 3, "%f %f", &u, &v);
            if (k < 1) { bad = true; why = "bad texcoord"; break; }
            T.push_back(u); T.push_back(k > 1 ? v : 0.0f);
        } else if (p[0] == 'v' && p[1] == 'n' && (p[2] == ' ' || p[2] == '\t')) {
            float x = 0, y = 0, z = 0;
            if (std::sscanf(p + 3, "%f %f %f", &x, &y, &z) != 3) { bad = true; why = "bad normal"; break; }
            N.push_back(x); N.push_back(y); N.push_back(z);
        } else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            poly.clear();
            for (;;) {
                while (*p == ' ' || *p == '\t') ++p;
                if (*p == 0 || *p == '\n' || *p == '\r' || *p == '#') break;
                Corner c;
                if (!parse_corner(p, (int)(P.size() / 3), (int)(T.size() / 2), (int)(N.size() / 3), c)) {
                    bad = true; why = "bad face"; break;
                }
                if (c.v < 0 || c.v >= (int)(P.size() / 3) || c.t >= (int)(T.size() / 2) || c.n >= (int)(N.size() / 3) ||
                    (c.t < -1) || (c.n < -1)) {
                    bad = true; why = "face index out of range"; break;
                }
                poly.push_back(c);
                while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
            }
            if (bad) break;
            if (poly.size() < 3) { bad = true; why = "face with fewer than 3 corners"; break; }
            for (size_t k = 1; k + 1 < poly.size(); ++k) {
                corners.push_back(poly[0]); corners.push_back(poly[k]); corners.push_back(poly[k + 1]);
            }
        }
    }
    std::fclose(f);
    if (bad) return set_error(CRT_ERR_IO, std::string(path) + ":" + std::to_string(lineno) + ": " + why);
    bool all_t = !corners.empty(), all_n = !corners.empty();
    for (const Corner& c : corners) { all_t = all_t && c.t >= 0; all_n = all_n && c.n >= 0; }
    std::unordered_map<Corner, uint32_t, CornerHash, CornerEq> map;
    std::vector<float> vp, vt, vn;
    std::vector<uint32_t> idx;
    idx.reserve(corners.size());
    for (Corner c : corners) {
        if (!all_t) c.t = -1;
        if (!all_n) c.n = -1;
        auto it = map.find(c);
        if (it != map.end()) { idx.push_back(it->second); continue; }
        uint32_t id = (uint32_t)(vp.size() / 3);
        map.emplace(c, id);
        vp.insert(vp.end(), &P[3 * (size_t)c.v], &P[3 * (size_t)c.v] + 3);
        if (all_t) vt.insert(vt.end(), &T[2 * (size_t)c.t], &T[2 * (size_t)c.t] + 2);
        if (all_n) vn.insert(vn.end(), &N[3 * (size_t)c.n], &N[3 * (size_t)c.n] + 3);
        idx.push_back(id);
    }
    auto dup = [](const std::vector<float>& v) -> float* {
        if (v.empty()) return nullptr;
        float* p = (float*)std::malloc(v.size() * sizeof(float));
        std::memcpy(p, v.data(), v.size() * sizeof(float));
        return p;
    };
    out->positions = dup(vp);
    out->num_vertices = (uint32_t)(vp.size() / 3);
    out->normals = dup(vn);
    out->uvs = dup(vt);
    out->num_triangles = (uint32_t)(idx.size() / 3);
    if (!idx.empty()) {
        out->indices = (uint32_t*)std::malloc(idx.size() * sizeof(uint32_t));
        std::memcpy(out->indices, idx.data(), idx.size() * sizeof(uint32_t));
    }
    return CRT_OK;
}

int crt_obj_free(crt_obj_mesh* m) {
    if (!m) return CRT_OK;
    std::free(m->positions); std::free(m->indices); std::free(m->normals); std::free(m->uvs);
    std::memset(m, 0, sizeof(*m));
    return CRT_OK;
}

}  // extern "C"
