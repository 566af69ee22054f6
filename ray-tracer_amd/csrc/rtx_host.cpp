// Host side of librtx: errors, tree export, the .bvh file format, and the .scn scene
// front-end with the reference's sphere/box generators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <exception>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rtx_clip.hpp"
#include "rtx_internal.hpp"

namespace rtx {

static thread_local std::string g_err;

void set_error(const std::string& m) { g_err = m; }

int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    g_err = std::string("HIP error ") + hipGetErrorString(e) + " at " + what;
    return false;
}

namespace {

// ---- reference math for scene construction (src/math/vec3.hpp, quat.hpp:36-42) ----
struct Q { float x, y, z, w; };

f3 qrot(const Q& q, f3 v) {
    // Rodrigues form of the reference: t = 2 cross(q.xyz, v); v' = v + w t + cross(q.xyz, t)
    f3 u{q.x, q.y, q.z};
    f3 t = scl(2.0f, cross(u, v));
    return add(add(v, scl(q.w, t)), cross(u, t));
}

struct Mesh {
    std::vector<float> v;       // 6 floats per vertex
    std::vector<uint32_t> t;    // 3 per triangle
    uint32_t nv() const { return (uint32_t)(v.size() / 6); }
    void vert(float x, float y, float z, float nx, float ny, float nz) {
        v.insert(v.end(), {x, y, z, nx, ny, nz});
    }
    void tri(uint32_t a, uint32_t b, uint32_t c) { t.insert(t.end(), {a, b, c}); }
};

const float kPi = 3.14159265358979323846f;

// UV sphere: poles plus (rings-1) rings of `sectors` vertices (bvh_geo.cpp:22-112).
void add_sphere(Mesh& m, f3 c, float radius, int rings, int sectors) {
    const uint32_t off = m.nv();
    m.vert(c.x, c.y + radius, c.z, 0.f, 1.f, 0.f);
    for (int r = 1; r < rings; ++r) {
        float theta = kPi * (float)r / (float)rings;
        for (int s = 0; s < sectors; ++s) {
            float phi = 2.f * kPi * (float)s / (float)sectors;
            float nx = sinf(theta) * cosf(phi);
            float ny = cosf(theta);
            float nz = sinf(theta) * sinf(phi);
            m.vert(c.x + radius * nx, c.y + radius * ny, c.z + radius * nz, nx, ny, nz);
        }
    }
    const uint32_t south = m.nv();
    m.vert(c.x, c.y - radius, c.z, 0.f, -1.f, 0.f);
    auto ring = [&](int r, int s) { return off + 1u + (uint32_t)((r - 1) * sectors + s); };
    for (int s = 0; s < sectors; ++s) m.tri(off, ring(1, (s + 1) % sectors), ring(1, s));
    for (int r = 1; r < rings - 1; ++r) {
        for (int s = 0; s < sectors; ++s) {
            int sn = (s + 1) % sectors;
            uint32_t a = ring(r, s), b = ring(r, sn), cc = ring(r + 1, sn), d = ring(r + 1, s);
            m.tri(a, b, cc);
            m.tri(a, cc, d);
        }
    }
    for (int s = 0; s < sectors; ++s) m.tri(south, ring(rings - 1, (s + 1) % sectors), ring(rings - 1, s));
}

// Oriented box: 6 faces x 4 vertices, faces (0,1,2)+(0,2,3) (bvh_geo.cpp:114-192).
void add_box(Mesh& m, f3 c, f3 h, const Q& q) {
    const f3 normal[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const f3 corner[6][4] = {
        {{h.x, h.y, -h.z}, {h.x, -h.y, -h.z}, {h.x, -h.y, h.z}, {h.x, h.y, h.z}},
        {{-h.x, h.y, h.z}, {-h.x, -h.y, h.z}, {-h.x, -h.y, -h.z}, {-h.x, h.y, -h.z}},
        {{-h.x, h.y, -h.z}, {-h.x, h.y, h.z}, {h.x, h.y, h.z}, {h.x, h.y, -h.z}},
        {{-h.x, -h.y, h.z}, {-h.x, -h.y, -h.z}, {h.x, -h.y, -h.z}, {h.x, -h.y, h.z}},
        {{h.x, h.y, h.z}, {h.x, -h.y, h.z}, {-h.x, -h.y, h.z}, {-h.x, h.y, h.z}},
        {{-h.x, h.y, -h.z}, {-h.x, -h.y, -h.z}, {h.x, -h.y, -h.z}, {h.x, h.y, -h.z}},
    };
    for (int f = 0; f < 6; ++f) {
        uint32_t off = m.nv();
        f3 rn = qrot(q, normal[f]);
        for (int k = 0; k < 4; ++k) {
            f3 rc = qrot(q, corner[f][k]);
            m.vert(c.x + rc.x, c.y + rc.y, c.z + rc.z, rn.x, rn.y, rn.z);
        }
        m.tri(off, off + 1, off + 2);
        m.tri(off, off + 2, off + 3);
    }
}

// Index validation and normal renormalisation of initGeos (bvh_IO.cpp:72-121).
int init_geos(Mesh& m) {
    if (m.v.empty()) return fail(RTX_E_ARG, "no vertices loaded on initGeos()");
    if (m.t.empty()) return fail(RTX_E_ARG, "no triangles loaded on initGeos()");
    const uint32_t nv = m.nv();
    for (size_t i = 0; i < m.t.size(); ++i)
        if (m.t[i] >= nv) return fail(RTX_E_ARG, "triangle has out-of-bounds vertex index on initGeos()");
    for (uint32_t i = 0; i < nv; ++i) {
        float* n = &m.v[6 * i + 3];
        float len = std::hypot(n[0], n[1], n[2]);
        if (len > 1e-6f) {
            n[0] /= len; n[1] /= len; n[2] /= len;
        } else {
            n[0] = 0.f; n[1] = 1.f; n[2] = 0.f;
        }
    }
    return RTX_OK;
}

// ASCII PLY mesh for `mesh` lines, in place of the assimp import of loadMesh
// (bvh_IO.cpp:24-70): polygons are split into fans (0,i,i+1) as aiProcess_Triangulate does
// for convex faces; missing normals are the per-vertex average of the unit face normals
// (aiProcess_GenSmoothNormals without its angle limit). The transform is
// translate(t) * scale(s) * rotate(q) for positions and the rotation for normals, which
// initGeos then renormalises; for a uniform scale this is the inverse-transpose direction.
int load_ply(const std::string& path, const f3& t, float scale, const Q& q, Mesh& m) {
    std::ifstream f(path);
    if (!f.is_open()) return fail(RTX_E_IO, "failed to open mesh " + path + " on loadMesh()");
    f.seekg(0, std::ios::end);
    const uint64_t fsize = (uint64_t)std::max<std::streamoff>(0, (std::streamoff)f.tellg());
    f.seekg(0, std::ios::beg);
    std::string line, tok;
    std::getline(f, line);
    if (line.rfind("ply", 0) != 0) return fail(RTX_E_IO, "not a PLY file: " + path + " on loadMesh()");
    uint64_t nvert = 0, nface = 0;
    int cur = 0;   // 1 = vertex element, 2 = face element
    std::vector<std::string> vprops;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        if (!(is >> tok)) continue;
        if (tok == "format") {
            std::string fmt;
            is >> fmt;
            if (fmt != "ascii") return fail(RTX_E_IO, "only ASCII PLY is supported: " + path + " on loadMesh()");
        } else if (tok == "element") {
            std::string name;
            uint64_t n = 0;
            is >> name >> n;
            if (name == "vertex") { nvert = n; cur = 1; }
            else if (name == "face") { nface = n; cur = 2; }
            else cur = 0;
        } else if (tok == "property" && cur == 1) {
            std::string type, name;
            is >> type >> name;
            vprops.push_back(name);
        } else if (tok == "end_header") {
            break;
        }
    }
    int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1;
    for (int i = 0; i < (int)vprops.size(); ++i) {
        const std::string& n = vprops[i];
        if (n == "x") ix = i; else if (n == "y") iy = i; else if (n == "z") iz = i;
        else if (n == "nx") inx = i; else if (n == "ny") iny = i; else if (n == "nz") inz = i;
    }
    if (ix < 0 || iy < 0 || iz < 0) return fail(RTX_E_IO, "PLY without x/y/z: " + path + " on loadMesh()");
    // every value takes at least two bytes ("0 "): counts the file cannot hold are refused
    // before anything is sized from them
    if (nvert > fsize / (2 * vprops.size()) || nface > fsize / 2 || nvert > 0xFFFFFFFFull)
        return fail(RTX_E_IO, "PLY element counts exceed the file size: " + path + " on loadMesh()");
    const bool has_n = inx >= 0 && iny >= 0 && inz >= 0;
    // rotation matrix of toMat4 (src/math/quat.hpp:44-59)
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z, xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    const float R[3][3] = {{1.f - 2.f * (yy + zz), 2.f * (xy - wz), 2.f * (xz + wy)},
                           {2.f * (xy + wz), 1.f - 2.f * (xx + zz), 2.f * (yz - wx)},
                           {2.f * (xz - wy), 2.f * (yz + wx), 1.f - 2.f * (xx + yy)}};
    auto rot = [&](f3 v) {
        return f3{R[0][0] * v.x + R[0][1] * v.y + R[0][2] * v.z, R[1][0] * v.x + R[1][1] * v.y + R[1][2] * v.z,
                  R[2][0] * v.x + R[2][1] * v.y + R[2][2] * v.z};
    };
    const uint32_t off = m.nv();
    std::vector<f3> pos(nvert), nrm(nvert, f3{0.f, 0.f, 0.f});
    std::vector<double> val(vprops.size());
    for (uint64_t i = 0; i < nvert; ++i) {
        for (size_t k = 0; k < vprops.size(); ++k)
            if (!(f >> val[k])) return fail(RTX_E_IO, "truncated PLY vertex list: " + path + " on loadMesh()");
        pos[i] = f3{(float)val[ix], (float)val[iy], (float)val[iz]};
        if (has_n) nrm[i] = f3{(float)val[inx], (float)val[iny], (float)val[inz]};
    }
    std::vector<uint32_t> tri;
    for (uint64_t i = 0; i < nface; ++i) {
        uint32_t c = 0;
        if (!(f >> c)) return fail(RTX_E_IO, "truncated PLY face list: " + path + " on loadMesh()");
        if (c > 65536 || c > fsize / 2) return fail(RTX_E_IO, "PLY face with too many vertices: " + path + " on loadMesh()");
        std::vector<uint32_t> idx(c);
        for (uint32_t k = 0; k < c; ++k) {
            if (!(f >> idx[k])) return fail(RTX_E_IO, "truncated PLY face list: " + path + " on loadMesh()");
            if (idx[k] >= nvert) return fail(RTX_E_ARG, "PLY face index out of range: " + path + " on loadMesh()");
        }
        for (uint32_t k = 1; k + 1 < c; ++k) tri.insert(tri.end(), {idx[0], idx[k], idx[k + 1]});
    }
    if (!has_n) {
        for (size_t i = 0; i < tri.size(); i += 3) {
            f3 a = pos[tri[i]], b = pos[tri[i + 1]], c = pos[tri[i + 2]];
            f3 fn = cross(sub(b, a), sub(c, a));
            float len = std::sqrt(dot(fn, fn));
            if (!(len > 0.f)) continue;
            fn = scl(1.f / len, fn);
            for (int k = 0; k < 3; ++k) nrm[tri[i + k]] = add(nrm[tri[i + k]], fn);
        }
    }
    for (uint64_t i = 0; i < nvert; ++i) {
        f3 p = rot(pos[i]);
        f3 n = rot(nrm[i]);
        m.vert(t.x + scale * p.x, t.y + scale * p.y, t.z + scale * p.z, n.x, n.y, n.z);
    }
    for (uint32_t v : tri) m.t.push_back(off + v);
    return RTX_OK;
}

// Geometry lines of a .scn file (readScene, bvh_IO.cpp:125-234).
int read_scene(const std::string& path, Mesh& m) {
    if (path.empty()) return fail(RTX_E_ARG, "empty scene path on readScene()");
    std::ifstream file(path);
    if (!file.is_open()) return fail(RTX_E_IO, "failed to open scene file " + path + " on readScene()");
    std::string line;
    while (std::getline(file, line)) {
        std::istringstream iss(line);
        std::string tok;
        if (!(iss >> tok)) continue;
        if (tok == "sphere") {
            float cx, cy, cz, r;
            if (!(iss >> cx >> cy >> cz >> r)) {
                std::fprintf(stderr, "Error : malformed sphere line on readScene()\n");
                continue;
            }
            int rings = 16, sectors = 32, a, b;
            if (iss >> a >> b) { rings = a; sectors = b; }
            if (rings < 2 || sectors < 3) return fail(RTX_E_ARG, "sphere needs rings >= 2 and sectors >= 3");
            add_sphere(m, {cx, cy, cz}, r, rings, sectors);
            continue;
        }
        if (tok == "box") {
            float cx, cy, cz, hx, hy, hz, qx, qy, qz, qw;
            std::string ot;
            if (!(iss >> cx >> cy >> cz >> hx >> hy >> hz >> ot >> qx >> qy >> qz >> qw)) {
                std::fprintf(stderr, "Error : malformed box line on readScene()\n");
                continue;
            }
            if (ot != "q") { qx = 0.f; qy = 0.f; qz = 0.f; qw = 1.f; }
            add_box(m, {cx, cy, cz}, {hx, hy, hz}, Q{qx, qy, qz, qw});
            continue;
        }
        if (tok == "mesh") {
            std::string mp, ot;
            float tx, ty, tz, sc, qx = 0.f, qy = 0.f, qz = 0.f, qw = 1.f;
            if (!(iss >> mp >> tx >> ty >> tz >> sc >> ot)) continue;
            if (ot == "q") iss >> qx >> qy >> qz >> qw;
            // the reference opens the path as given; fall back to the scene file's directory
            std::string full = mp;
            if (!std::ifstream(full).good()) {
                auto slash = path.find_last_of('/');
                if (slash != std::string::npos) full = path.substr(0, slash + 1) + mp;
            }
            auto dot_ = mp.rfind('.');
            std::string ext = dot_ == std::string::npos ? "" : mp.substr(dot_);
            if (ext != ".ply")
                return fail(RTX_E_IO, "mesh format not supported (ASCII .ply only): " + mp + " on readScene()");
            int rc = load_ply(full, {tx, ty, tz}, sc, Q{qx, qy, qz, qw}, m);
            if (rc) return rc;
            continue;
        }
    }
    return RTX_OK;
}

std::string bvh_path_for(const std::string& scn) {
    auto dot = scn.rfind('.');
    return (dot != std::string::npos ? scn.substr(0, dot) : scn) + ".bvh";
}

}  // namespace
}  // namespace rtx

using namespace rtx;

extern "C" const char* rtx_last_error(void) { return g_err.c_str(); }

extern "C" int rtx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int rtx_bvh_get_info(const rtx_bvh* b, rtx_bvh_info* info) {
    if (!b || !info) return fail(RTX_E_ARG, "null argument on rtx_bvh_get_info()");
    info->node_count = b->nn;
    info->tri_count = b->ntri;
    info->vert_count = b->nv;
    info->input_tri_count = b->nt;
    info->internal_count = b->ninternal;
    info->leaf_count = b->nleaves;
    info->max_depth = b->max_depth;
    info->tri_index_size = b->tindex_size;
    info->sorted_tris_size = b->sorted_size;
    info->levels = b->levels;
    for (int k = 0; k < 3; ++k) { info->scene_min[k] = b->scene[k]; info->scene_max[k] = b->scene[3 + k]; }
    return RTX_OK;
}

extern "C" int rtx_bvh_export(const rtx_bvh* b, void* nodes48, uint32_t* tris3, uint32_t* prims) {
    if (!b) return fail(RTX_E_ARG, "null argument on rtx_bvh_export()");
    if (nodes48) RTX_HIP(hipMemcpy(nodes48, b->nodes48, sizeof(Node48) * b->nn, hipMemcpyDeviceToHost));
    if (tris3) RTX_HIP(hipMemcpy(tris3, b->tris3, 12ull * b->ntri, hipMemcpyDeviceToHost));
    if (prims) RTX_HIP(hipMemcpy(prims, b->prims, 4ull * b->ntri, hipMemcpyDeviceToHost));
    return RTX_OK;
}

extern "C" int rtx_debug_clip_slab(const float* v9, const float* region6, int axis, float* out6) {
    if (!v9 || !region6 || !out6 || axis < 0 || axis > 2) return fail(RTX_E_ARG, "bad argument on rtx_debug_clip_slab()");
    const Box r{{region6[0], region6[1], region6[2]}, {region6[3], region6[4], region6[5]}};
    Box b;
    if (!clip_tri_slab(f3{v9[0], v9[1], v9[2]}, f3{v9[3], v9[4], v9[5]}, f3{v9[6], v9[7], v9[8]}, r, axis, b)) return 0;
    out6[0] = b.lo.x; out6[1] = b.lo.y; out6[2] = b.lo.z;
    out6[3] = b.hi.x; out6[4] = b.hi.y; out6[5] = b.hi.z;
    return 1;
}

extern "C" int rtx_bvh_morton(const rtx_bvh* b, uint64_t* codes, uint32_t* sorted) {
    if (!b) return fail(RTX_E_ARG, "null argument on rtx_bvh_morton()");
    if (!b->codes || !b->sorted) return fail(RTX_E_STATE, "tree was loaded, not built, on rtx_bvh_morton()");
    if (codes) RTX_HIP(hipMemcpy(codes, b->codes, 8ull * b->nt, hipMemcpyDeviceToHost));
    if (sorted) RTX_HIP(hipMemcpy(sorted, b->sorted, 4ull * b->nt, hipMemcpyDeviceToHost));
    return RTX_OK;
}

extern "C" int rtx_bvh_export_bfs(const rtx_bvh* b, void* nodes48, uint32_t* count) {
    if (!b) return fail(RTX_E_ARG, "null argument on rtx_bvh_export_bfs()");
    if (!b->bfs48) return fail(RTX_E_STATE, "tree was loaded, not built, on rtx_bvh_export_bfs()");
    if (count) *count = b->nn;
    if (nodes48) RTX_HIP(hipMemcpy(nodes48, b->bfs48, sizeof(Node48) * b->nn, hipMemcpyDeviceToHost));
    return RTX_OK;
}

// ---- .bvh files (writeBakedData, bvh_IO.cpp:236-275; loadBVH, bvh_traverse.cpp:182-235) ----
static const uint32_t kBvhVersion = 1;

extern "C" int rtx_save(const rtx_bvh* b, const char* path) {
    if (!b || !path) return fail(RTX_E_ARG, "null argument on rtx_save()");
    std::vector<Node48> nodes(b->nn);
    std::vector<uint32_t> tris(3ull * b->ntri);
    std::vector<float> verts(6ull * b->nv);
    int rc = rtx_bvh_export(b, nodes.data(), tris.data(), nullptr);
    if (rc) return rc;
    RTX_HIP(hipMemcpy(verts.data(), b->verts, 24ull * b->nv, hipMemcpyDeviceToHost));
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) return fail(RTX_E_IO, std::string("failed to open output file ") + path + " on rtx_save()");
    uint32_t hdr[4] = {kBvhVersion, b->nn, b->ntri, b->nv};
    out.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    out.write(reinterpret_cast<const char*>(nodes.data()), sizeof(Node48) * nodes.size());
    out.write(reinterpret_cast<const char*>(tris.data()), 4 * tris.size());
    out.write(reinterpret_cast<const char*>(verts.data()), 4 * verts.size());
    if (!out.good()) return fail(RTX_E_IO, std::string("write error for ") + path + " on rtx_save()");
    return RTX_OK;
}

extern "C" int rtx_load(const char* path, rtx_bvh** outp) {
    if (!path || !outp) return fail(RTX_E_ARG, "null argument on rtx_load()");
    *outp = nullptr;
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) return fail(RTX_E_IO, std::string("failed to open ") + path + " on loadBVH()");
    uint32_t hdr[4] = {0, 0, 0, 0};
    in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    if (!in.good()) return fail(RTX_E_IO, std::string("failed to read header from ") + path + " on loadBVH()");
    if (hdr[0] != kBvhVersion) return fail(RTX_E_VERSION, std::string("version mismatch in ") + path + " on loadBVH()");
    // Reject counts the file cannot hold before allocating for them.
    in.seekg(0, std::ios::end);
    uint64_t fsize = (uint64_t)in.tellg();
    uint64_t need = 16ull + 48ull * hdr[1] + 12ull * hdr[2] + 24ull * hdr[3];
    if (need > fsize) return fail(RTX_E_IO, std::string("failed to read payload from ") + path + " on loadBVH()");
    in.seekg(16, std::ios::beg);
    std::vector<Node48> nodes(hdr[1]);
    std::vector<uint32_t> tris(3ull * hdr[2]);
    std::vector<float> verts(6ull * hdr[3]);
    in.read(reinterpret_cast<char*>(nodes.data()), 48ull * hdr[1]);
    in.read(reinterpret_cast<char*>(tris.data()), 12ull * hdr[2]);
    in.read(reinterpret_cast<char*>(verts.data()), 24ull * hdr[3]);
    if (!in.good()) return fail(RTX_E_IO, std::string("failed to read payload from ") + path + " on loadBVH()");
    rtx_bvh* b = new rtx_bvh;
    (void)hipGetDevice(&b->device);
    int rc = upload_reference_tree(b, nodes.data(), hdr[1], tris.data(), hdr[2], verts.data(), hdr[3]);
    if (rc) {
        rtx_bvh_free(b);
        return rc;
    }
    *outp = b;
    return RTX_OK;
}

extern "C" int rtx_scene_load(const char* scn, float** verts6, uint32_t* nv, uint32_t** tris, uint32_t* nt) {
    if (!scn || !verts6 || !nv || !tris || !nt) return fail(RTX_E_ARG, "null argument on rtx_scene_load()");
    Mesh m;
    int rc;
    try {   // no exception crosses the C ABI: a scene too large for host memory is an error
        rc = read_scene(scn, m);
    } catch (const std::exception& e) {
        return fail(RTX_E_IO, std::string("scene could not be read (") + e.what() + ") on rtx_scene_load()");
    }
    if (rc) return rc;
    rc = init_geos(m);
    if (rc) return rc;
    *verts6 = (float*)std::malloc(4 * m.v.size());
    *tris = (uint32_t*)std::malloc(4 * m.t.size());
    if (!*verts6 || !*tris) return fail(RTX_E_ARG, "host allocation failed on rtx_scene_load()");
    std::memcpy(*verts6, m.v.data(), 4 * m.v.size());
    std::memcpy(*tris, m.t.data(), 4 * m.t.size());
    *nv = m.nv();
    *nt = (uint32_t)(m.t.size() / 3);
    return RTX_OK;
}

extern "C" void rtx_free_host(void* p) { std::free(p); }

// ---- Radiance .hdr (RGBE) -------------------------------------------------------------
extern "C" int rtx_hdr_write(const char* path, int w, int h, const float* rgb, int flip_y) {
    if (!path || !rgb || w <= 0 || h <= 0) return fail(RTX_E_ARG, "bad argument on rtx_hdr_write()");
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) return fail(RTX_E_IO, std::string("failed to open ") + path + " on rtx_hdr_write()");
    out << "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " << h << " +X " << w << "\n";
    std::vector<uint8_t> row(4 * (size_t)w);
    for (int y = 0; y < h; ++y) {
        const int src = flip_y ? h - 1 - y : y;
        for (int x = 0; x < w; ++x) {
            const float* p = rgb + 3 * ((size_t)src * w + x);
            float v = p[0] > p[1] ? p[0] : p[1];
            v = v > p[2] ? v : p[2];
            uint8_t* o = &row[4 * (size_t)x];
            if (!(v >= 1e-32f)) { o[0] = o[1] = o[2] = o[3] = 0; continue; }
            int e;
            const float m = std::frexp(v, &e);      // v = m * 2^e, m in [0.5, 1)
            const float f = m * 256.0f / v;
            o[0] = (uint8_t)(p[0] > 0.f ? p[0] * f : 0.f);
            o[1] = (uint8_t)(p[1] > 0.f ? p[1] * f : 0.f);
            o[2] = (uint8_t)(p[2] > 0.f ? p[2] * f : 0.f);
            o[3] = (uint8_t)(e + 128);
        }
        out.write(reinterpret_cast<const char*>(row.data()), (std::streamsize)row.size());
    }
    if (!out.good()) return fail(RTX_E_IO, std::string("write error for ") + path + " on rtx_hdr_write()");
    return RTX_OK;
}

extern "C" int rtx_hdr_read(const char* path, int* wp, int* hp, float** rgbp) {
    if (!path || !wp || !hp || !rgbp) return fail(RTX_E_ARG, "null argument on rtx_hdr_read()");
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) return fail(RTX_E_IO, std::string("failed to open ") + path + " on rtx_hdr_read()");
    std::string line;
    std::getline(in, line);
    if (line.rfind("#?", 0) != 0) return fail(RTX_E_IO, std::string("not a Radiance file: ") + path);
    while (std::getline(in, line) && !line.empty()) {
        if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe")
            return fail(RTX_E_IO, std::string("unsupported Radiance format in ") + path);
    }
    std::getline(in, line);
    char ys[8], xs[8];
    int h = 0, w = 0;
    if (std::sscanf(line.c_str(), "%7s %d %7s %d", ys, &h, xs, &w) != 4 || std::string(ys) != "-Y" ||
        std::string(xs) != "+X" || w <= 0 || h <= 0 || (uint64_t)w * h > (1ull << 28))
        return fail(RTX_E_IO, std::string("unsupported Radiance resolution line in ") + path);
    float* rgb = (float*)std::malloc(12ull * w * h);
    if (!rgb) return fail(RTX_E_ARG, "host allocation failed on rtx_hdr_read()");
    std::vector<uint8_t> row(4 * (size_t)w);
    auto bad = [&](const char* what) {
        std::free(rgb);
        return fail(RTX_E_IO, std::string(what) + " in " + path + " on rtx_hdr_read()");
    };
    for (int y = 0; y < h; ++y) {
        uint8_t hdr4[4];
        if (!in.read(reinterpret_cast<char*>(hdr4), 4)) return bad("truncated scanline");
        const bool rle = w >= 8 && w < 32768 && hdr4[0] == 2 && hdr4[1] == 2 && ((hdr4[2] << 8) | hdr4[3]) == w;
        if (!rle) {
            if (hdr4[0] == 1 && hdr4[1] == 1 && hdr4[2] == 1) return bad("old-style run-length encoding");
            std::memcpy(row.data(), hdr4, 4);
            if (w > 1 && !in.read(reinterpret_cast<char*>(row.data() + 4), 4 * ((std::streamsize)w - 1)))
                return bad("truncated scanline");
        } else {
            // channel-planar runs: count > 128 repeats the next byte (count - 128) times
            std::vector<uint8_t> plane((size_t)w);
            for (int c = 0; c < 4; ++c) {
                int x = 0;
                while (x < w) {
                    int cnt = in.get();
                    if (cnt == EOF) return bad("truncated run");
                    if (cnt > 128) {
                        cnt -= 128;
                        int v = in.get();
                        if (v == EOF || x + cnt > w) return bad("bad run");
                        std::memset(&plane[x], v, cnt);
                    } else {
                        if (cnt == 0 || x + cnt > w) return bad("bad run");
                        if (!in.read(reinterpret_cast<char*>(&plane[x]), cnt)) return bad("truncated run");
                    }
                    x += cnt;
                }
                for (int i = 0; i < w; ++i) row[4 * (size_t)i + c] = plane[i];
            }
        }
        for (int x = 0; x < w; ++x) {
            const uint8_t* q = &row[4 * (size_t)x];
            float* o = rgb + 3 * ((size_t)y * w + x);
            if (q[3] == 0) { o[0] = o[1] = o[2] = 0.f; continue; }
            const float f = std::ldexp(1.0f, (int)q[3] - (128 + 8));
            o[0] = q[0] * f; o[1] = q[1] * f; o[2] = q[2] * f;
        }
    }
    *wp = w;
    *hp = h;
    *rgbp = rgb;
    return RTX_OK;
}

extern "C" int rtx_bake_scene(const char* scn) {
    float* v = nullptr;
    uint32_t* t = nullptr;
    uint32_t nv = 0, nt = 0;
    int rc = rtx_scene_load(scn, &v, &nv, &t, &nt);
    rtx_bvh* b = nullptr;
    if (rc == RTX_OK) rc = rtx_build(v, nv, t, nt, &b);
    if (rc == RTX_OK) rc = rtx_save(b, bvh_path_for(scn).c_str());
    rtx_bvh_free(b);
    std::free(v);
    std::free(t);
    return rc;
}

static int geo_out(const Mesh& m, float* verts6, uint32_t* nv, uint32_t* tris, uint32_t* nt) {
    if (nv) *nv = m.nv();
    if (nt) *nt = (uint32_t)(m.t.size() / 3);
    if (verts6) std::memcpy(verts6, m.v.data(), 4 * m.v.size());
    if (tris) std::memcpy(tris, m.t.data(), 4 * m.t.size());
    return RTX_OK;
}

extern "C" int rtx_geo_sphere(const float c[3], float radius, int rings, int sectors, float* verts6, uint32_t* nv,
                              uint32_t* tris, uint32_t* nt) {
    if (!c || rings < 2 || sectors < 3) return fail(RTX_E_ARG, "bad sphere arguments on rtx_geo_sphere()");
    Mesh m;
    add_sphere(m, {c[0], c[1], c[2]}, radius, rings, sectors);
    return geo_out(m, verts6, nv, tris, nt);
}

extern "C" int rtx_geo_box(const float c[3], const float h[3], const float q[4], float* verts6, uint32_t* nv,
                           uint32_t* tris, uint32_t* nt) {
    if (!c || !h || !q) return fail(RTX_E_ARG, "null argument on rtx_geo_box()");
    Mesh m;
    add_box(m, {c[0], c[1], c[2]}, {h[0], h[1], h[2]}, Q{q[0], q[1], q[2], q[3]});
    return geo_out(m, verts6, nv, tris, nt);
}
