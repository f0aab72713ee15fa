// obj.hip — Wavefront OBJ meshes -> the vertex / index / geometry arrays a BLAS build takes (host code).
//
// The reference's next step after the ray-traced triangle is "load an .obj file -> build the acceleration
// structure" (vulkan-raytracing-basic/README.md:225-226).  Subset read: `v x y z [w]` (w ignored), `f` with
// `v`, `v/vt`, `v//vn` or `v/vt/vn` references (1-based, or negative = relative to the vertices read so far),
// polygons of n >= 3 vertices fan-triangulated (v0, v_k, v_k+1); `o` and `g` start a new geometry (empty ones
// are dropped); comments, `vt`, `vn`, `s`, `usemtl`, `mtllib` and other statements are skipped.  All geometries
// share the vertex array; geometry g covers triangles [first[g], first[g + 1]).
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rtx.h"

namespace {

struct Cursor {
    const char* p;
    const char* end;
    bool at_eol() const { return p >= end || *p == '\n' || *p == '\r' || *p == '#'; }
    void skip_blank() {
        while (p < end && (*p == ' ' || *p == '\t')) ++p;
    }
    void next_line() {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
};

bool read_float(Cursor& c, float& out) {
    c.skip_blank();
    if (c.at_eol()) return false;
    char buf[64];
    size_t n = 0;
    while (c.p < c.end && n < sizeof(buf) - 1 && !isspace((unsigned char)*c.p)) buf[n++] = *c.p++;
    buf[n] = 0;
    char* e = nullptr;
    errno = 0;
    out = strtof(buf, &e);
    return e != buf && *e == 0 && errno == 0;
}

// One face corner "v", "v/vt", "v//vn" or "v/vt/vn": returns the vertex reference.
bool read_corner(Cursor& c, long& v) {
    c.skip_blank();
    if (c.at_eol()) return false;
    char* e = nullptr;
    char buf[96];
    size_t n = 0;
    while (c.p < c.end && n < sizeof(buf) - 1 && !isspace((unsigned char)*c.p)) buf[n++] = *c.p++;
    buf[n] = 0;
    errno = 0;
    v = strtol(buf, &e, 10);
    if (e == buf || errno != 0 || v == 0) return false;
    return *e == 0 || *e == '/';
}

int parse(const char* text, size_t len, rtx_obj_mesh* out) {
    std::vector<float> verts;
    std::vector<uint32_t> idx;
    std::vector<uint32_t> first;  // first triangle of each started geometry
    first.push_back(0);
    Cursor c{text, text + len};
    std::vector<long> poly;
    while (c.p < c.end) {
        c.skip_blank();
        if (c.at_eol()) {
            c.next_line();
            continue;
        }
        const char* kw = c.p;
        while (c.p < c.end && !isspace((unsigned char)*c.p)) ++c.p;
        const size_t kl = (size_t)(c.p - kw);
        if (kl == 1 && kw[0] == 'v') {
            float x, y, z;
            if (!read_float(c, x) || !read_float(c, y) || !read_float(c, z)) return RTX_ERROR_FORMAT;
            verts.push_back(x);
            verts.push_back(y);
            verts.push_back(z);
        } else if (kl == 1 && kw[0] == 'f') {
            poly.clear();
            long v;
            while (read_corner(c, v)) poly.push_back(v);
            c.skip_blank();
            if (!c.at_eol() || poly.size() < 3) return RTX_ERROR_FORMAT;
            const long nv = (long)(verts.size() / 3);
            for (long& r : poly) {
                r = r > 0 ? r - 1 : nv + r;
                if (r < 0 || r >= nv) return RTX_ERROR_FORMAT;
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k) {
                idx.push_back((uint32_t)poly[0]);
                idx.push_back((uint32_t)poly[k]);
                idx.push_back((uint32_t)poly[k + 1]);
            }
        } else if (kl == 1 && (kw[0] == 'o' || kw[0] == 'g')) {
            const uint32_t t = (uint32_t)(idx.size() / 3);
            if (first.back() != t) first.push_back(t);  // close the current geometry if it has triangles
        }
        c.next_line();
    }
    const uint32_t ntri = (uint32_t)(idx.size() / 3);
    if (first.back() != ntri || ntri == 0) first.push_back(ntri);
    if (verts.size() / 3 > 0xFFFFFFFFull || ntri > (1u << 30)) return RTX_ERROR_FORMAT;
    rtx_obj_mesh m{};
    m.vertex_count = (uint32_t)(verts.size() / 3);
    m.triangle_count = ntri;
    m.geometry_count = ntri ? (uint32_t)first.size() - 1 : 0;
    m.vertices = (float*)malloc(verts.size() * sizeof(float) + 4);
    m.indices = (uint32_t*)malloc(idx.size() * sizeof(uint32_t) + 4);
    m.geometry_first_triangle = (uint32_t*)malloc(first.size() * sizeof(uint32_t));
    if (!m.vertices || !m.indices || !m.geometry_first_triangle) {
        free(m.vertices);
        free(m.indices);
        free(m.geometry_first_triangle);
        return RTX_ERROR_OUT_OF_HOST_MEMORY;
    }
    if (!verts.empty()) memcpy(m.vertices, verts.data(), verts.size() * sizeof(float));
    if (!idx.empty()) memcpy(m.indices, idx.data(), idx.size() * sizeof(uint32_t));
    memcpy(m.geometry_first_triangle, first.data(), (m.geometry_count + 1) * sizeof(uint32_t));
    *out = m;
    return RTX_SUCCESS;
}

}  // namespace

extern "C" {

int rtx_obj_parse(const char* text, size_t length, rtx_obj_mesh* out_mesh) {
    if (!out_mesh || (length > 0 && !text)) return RTX_ERROR_INVALID_ARGUMENT;
    memset(out_mesh, 0, sizeof *out_mesh);
    try {
        return parse(text, length, out_mesh);
    } catch (const std::bad_alloc&) {
        return RTX_ERROR_OUT_OF_HOST_MEMORY;
    }
}

int rtx_obj_load(const char* path, rtx_obj_mesh* out_mesh) {
    if (!path || !out_mesh) return RTX_ERROR_INVALID_ARGUMENT;
    memset(out_mesh, 0, sizeof *out_mesh);
    FILE* f = fopen(path, "rb");
    if (!f) return RTX_ERROR_INVALID_ARGUMENT;
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t n;
    try {
        while ((n = fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    } catch (const std::bad_alloc&) {
        fclose(f);
        return RTX_ERROR_OUT_OF_HOST_MEMORY;
    }
    const bool err = ferror(f) != 0;
    fclose(f);
    if (err) return RTX_ERROR_INVALID_ARGUMENT;
    return rtx_obj_parse(buf.data(), buf.size(), out_mesh);
}

void rtx_obj_free(rtx_obj_mesh* mesh) {
    if (!mesh) return;
    free(mesh->vertices);
    free(mesh->indices);
    free(mesh->geometry_first_triangle);
    memset(mesh, 0, sizeof *mesh);
}

}  // extern "C"
