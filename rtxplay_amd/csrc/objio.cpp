// objio.cpp -- Wavefront OBJ meshes in and out (host code).
//
// Object(wavefront) of optx/object.cxx:23-93 reads a file with tinyobjloader
// and keeps one indexed triangle mesh per shape.  This reader parses the
// subset that loader keeps for geometry, written here from the OBJ format
// description, and fixes the reference's loader defects:
//   * the index bound check compares against the vertex count (the
//     reference compares against 3x that count, with > instead of >=);
//   * each shape's vertices are the ones its faces use, in first-use order,
//     and its indices are renumbered into that array (the reference keeps the
//     file's indices and only happens to work for files whose vertices are
//     in first-use order, such as the sphere writer's);
//   * the first-use map is per shape (the reference shares it).
// Lines: "v x y z [w]" (w ignored), "f a b c ..." with a, b, c as "i",
// "i/t", "i//n" or "i/t/n" (1-based; negative = relative to the end), "o"
// and "g" start a new shape; "vt", "vn", "s", "usemtl", "mtllib", comments
// and blank lines are skipped.  A face of n > 3 corners is split into the
// fan (0, k, k+1); a face of fewer than 3 corners is an error.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_util.h"
#include "rtx.h"

struct rtx_object {
    std::vector<std::vector<rtx_float3>> vces;
    std::vector<std::vector<rtx_uint3>> ices;
    std::vector<std::string> names;
};

namespace {

struct Shape {
    std::string name;
    std::vector<rtx_float3> v;
    std::vector<rtx_uint3> i;
    std::unordered_map<uint32_t, uint32_t> first_use;   // file vertex -> shape vertex
};

bool parse_index(const char* tok, long nverts, uint32_t& out) {
    char* end = nullptr;
    errno = 0;
    const long k = std::strtol(tok, &end, 10);
    if (end == tok || errno) return false;
    long idx = k > 0 ? k - 1 : (k < 0 ? nverts + k : -1);
    if (idx < 0 || idx >= nverts) return false;
    out = (uint32_t)idx;
    return true;
}

}  // namespace

extern "C" {

rtx_status rtx_object_load(const char* wavefront, rtx_object** out) {
    if (!wavefront || !out) return rtx::set_error(RTX_ERROR_INVALID_VALUE, "rtx_object_load: NULL argument");
    std::ifstream in(wavefront);
    if (!in) return rtx::set_error(RTX_ERROR_INVALID_VALUE, std::string(wavefront) + ": cannot open");
    std::vector<rtx_float3> verts;
    std::vector<Shape> shapes(1);
    std::string line;
    size_t lineno = 0;
    auto err = [&](const std::string& what) {
        return rtx::set_error(RTX_ERROR_INVALID_VALUE,
                              std::string(wavefront) + ":" + std::to_string(lineno) + ": " + what);
    };
    while (std::getline(in, line)) {
        lineno++;
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        if (key == "v") {
            float x, y, z;
            if (!(ls >> x >> y >> z)) return err("vertex needs three coordinates");
            verts.push_back({x, y, z});
        } else if (key == "o" || key == "g") {
            std::string name;
            std::getline(ls, name);
            const size_t b = name.find_first_not_of(" \t");
            name = b == std::string::npos ? std::string() : name.substr(b);
            if (!shapes.back().i.empty()) shapes.emplace_back();
            shapes.back().name = name;
        } else if (key == "f") {
            std::vector<uint32_t> corner;
            std::string tok;
            while (ls >> tok) {
                const std::string vtok = tok.substr(0, tok.find('/'));
                uint32_t idx;
                if (!parse_index(vtok.c_str(), (long)verts.size(), idx))
                    return err("vertex index out of bounds or malformed: " + tok);
                corner.push_back(idx);
            }
            if (corner.size() < 3) return err("face with fewer than 3 corners");
            Shape& sh = shapes.back();
            auto local = [&](uint32_t g) {
                auto it = sh.first_use.find(g);
                if (it != sh.first_use.end()) return it->second;
                const uint32_t k = (uint32_t)sh.v.size();
                sh.first_use.emplace(g, k);
                sh.v.push_back(verts[g]);
                return k;
            };
            for (size_t k = 1; k + 1 < corner.size(); k++) {
                const uint32_t a = local(corner[0]), b = local(corner[k]), c = local(corner[k + 1]);
                sh.i.push_back({a, b, c});
            }
        }
        // vt, vn, s, usemtl, mtllib, l, p and the rest carry no triangle geometry
    }
    rtx_object* o = new rtx_object();
    for (auto& sh : shapes) {
        if (sh.i.empty()) continue;
        o->vces.push_back(std::move(sh.v));
        o->ices.push_back(std::move(sh.i));
        o->names.push_back(sh.name);
    }
    *out = o;
    return RTX_SUCCESS;
}

rtx_status rtx_object_size(const rtx_object* o, uint32_t* n) {
    if (!o || !n) return rtx::set_error(RTX_ERROR_INVALID_VALUE, "rtx_object_size: NULL argument");
    *n = (uint32_t)o->vces.size();
    return RTX_SUCCESS;
}

rtx_status rtx_object_mesh(const rtx_object* o, uint32_t m, const rtx_float3** vces, uint32_t* nv,
                           const rtx_uint3** ices, uint32_t* ni) {
    if (!o || !vces || !nv || !ices || !ni) return rtx::set_error(RTX_ERROR_INVALID_VALUE, "rtx_object_mesh: NULL argument");
    if (m >= o->vces.size()) return rtx::set_error(RTX_ERROR_INVALID_VALUE, "rtx_object_mesh: no such submesh");
    *vces = o->vces[m].data();
    *nv = (uint32_t)o->vces[m].size();
    *ices = o->ices[m].data();
    *ni = (uint32_t)o->ices[m].size();
    return RTX_SUCCESS;
}

rtx_status rtx_object_free(rtx_object* o) {
    delete o;
    return RTX_SUCCESS;
}

rtx_status rtx_object_write(const char* path, const char* name, const rtx_float3* const* vces, const uint32_t* nv,
                            const rtx_uint3* const* ices, const uint32_t* ni, uint32_t nmesh) {
    if (!path || (nmesh && (!vces || !nv || !ices || !ni)))
        return rtx::set_error(RTX_ERROR_INVALID_VALUE, "rtx_object_write: NULL argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return rtx::set_error(RTX_ERROR_INVALID_VALUE, std::string(path) + ": cannot write");
    std::fprintf(f, "o %s\n", name ? name : "object");
    uint64_t base = 0, vsum = 0;
    for (uint32_t m = 0; m < nmesh; m++) {
        for (uint32_t v = 0; v < nv[m]; v++)
            std::fprintf(f, "v %.9g %.9g %.9g\n", vces[m][v].x, vces[m][v].y, vces[m][v].z);
        vsum += nv[m];
    }
    std::fprintf(f, "# %llu vertices\n", (unsigned long long)vsum);
    for (uint32_t m = 0; m < nmesh; m++) {
        for (uint32_t t = 0; t < ni[m]; t++) {
            const rtx_uint3 i = ices[m][t];
            if (i.x >= nv[m] || i.y >= nv[m] || i.z >= nv[m]) {
                std::fclose(f);
                return rtx::set_error(RTX_ERROR_INVALID_VALUE, "rtx_object_write: index out of bounds");
            }
            std::fprintf(f, "f %llu %llu %llu\n", (unsigned long long)(base + i.x + 1),
                         (unsigned long long)(base + i.y + 1), (unsigned long long)(base + i.z + 1));
        }
        std::fprintf(f, "# %u triangles\n", ni[m]);
        base += nv[m];
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? RTX_SUCCESS : rtx::set_error(RTX_ERROR_INVALID_VALUE, std::string(path) + ": write failed");
}

}  // extern "C"
