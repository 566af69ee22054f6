// Harness that compiles the reference's own builder and tracer sources (as they lie
// under the reference checkout) into oracle/_ref/libref.so, so the CPU oracle can be
// pinned against the real thing. TEST INFRASTRUCTURE ONLY; never part of the product.
//
// bvh_bake.cpp and bvh_traverse.cpp are #included so their file-static pipeline steps
// (buildCentroidCache, reorderNodes, reorderTriangles) and traversal state (t_nodes,
// t_tris, t_verts) are reachable. AABB.cpp, bvh_morton.cpp, bvh_bfs.cpp, bvh_sah.cpp and
// bvh_geo.cpp are compiled unchanged beside this file. bvh_IO.cpp needs assimp, which
// this image lacks, so it is not compiled: the harness fills the baking state in memory
// and never calls bakeBVH(). The three bvh_IO.cpp entry points bakeBVH() would reach are
// defined below only to satisfy the linker; each reports an error and returns false.
#include "bvh_bake.cpp"
#include "bvh_traverse.cpp"
#include "bvh_geo.hpp"
#include "bvh_morton.hpp"
#include "bvh_bfs.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace bvh {
bool readScene() { std::fprintf(stderr, "ref harness: readScene unavailable\n"); return false; }
bool initGeos() { std::fprintf(stderr, "ref harness: initGeos unavailable\n"); return false; }
bool writeBakedData() { std::fprintf(stderr, "ref harness: writeBakedData unavailable\n"); return false; }
}  // namespace bvh

static_assert(sizeof(BVHNode) == 48, "BVHNode layout");
static_assert(sizeof(Triangle) == 12, "Triangle layout");
static_assert(sizeof(Vertex) == 24, "Vertex layout");

namespace {
struct Result {
    std::vector<BVHNode> bfs_nodes, nodes;
    std::vector<Triangle> tris;
    std::vector<uint32_t> prims, sorted;
    std::vector<uint64_t> codes;
    float scene[6];
    uint32_t tindex_size = 0, sorted_size = 0;
};
Result g_res;
}  // namespace

extern "C" {

int ref_build(const float* verts6, uint32_t nv, const uint32_t* tris, uint32_t nt) {
    using namespace bvh;
    finishBake();
    g_res = Result{};
    s_vertices.resize(nv);
    std::memcpy(s_vertices.data(), verts6, sizeof(Vertex) * nv);
    s_triangles.resize(nt);
    std::memcpy(s_triangles.data(), tris, sizeof(Triangle) * nt);
    if (!sortTrisByMorton()) return -1;
    g_res.codes.resize(nt);
    for (uint32_t i = 0; i < nt; ++i) g_res.codes[i] = s_mortonData[i].mortonCode;
    g_res.sorted.assign(s_sortedTris.begin(), s_sortedTris.begin() + nt);
    g_res.scene[0] = s_sceneAABB.min.x; g_res.scene[1] = s_sceneAABB.min.y; g_res.scene[2] = s_sceneAABB.min.z;
    g_res.scene[3] = s_sceneAABB.max.x; g_res.scene[4] = s_sceneAABB.max.y; g_res.scene[5] = s_sceneAABB.max.z;
    if (!buildCentroidCache()) return -2;
    if (!initRoot()) return -3;
    if (!bfsLoop()) return -4;
    g_res.bfs_nodes.assign(s_nodes.begin(), s_nodes.begin() + s_totalNodeCount);
    g_res.tindex_size = s_triIndexSize;
    g_res.sorted_size = s_sortedTrisSize;
    if (!reorderNodes()) return -5;
    // Record the original triangle id of every reordered slot, walking leaves in the
    // same order reorderTriangles() does.
    for (uint32_t i = 0; i < s_totalNodeCount; ++i) {
        if (!s_nodes[i].isLeaf) continue;
        for (uint32_t k = 0; k < s_nodes[i].triSize; ++k)
            g_res.prims.push_back(s_sortedTris[s_triIndex[s_nodes[i].beginTriIndex + k]]);
    }
    if (!reorderTriangles()) return -6;
    g_res.nodes.assign(s_nodes.begin(), s_nodes.begin() + s_totalNodeCount);
    g_res.tris = s_triangles;
    finishBake();
    return 0;
}

void ref_counts(uint32_t* c) {
    c[0] = (uint32_t)g_res.nodes.size(); c[1] = (uint32_t)g_res.tris.size();
    c[2] = g_res.tindex_size; c[3] = g_res.sorted_size;
}
void ref_nodes(void* out) { std::memcpy(out, g_res.nodes.data(), g_res.nodes.size() * sizeof(BVHNode)); }
void ref_bfs_nodes(void* out) { std::memcpy(out, g_res.bfs_nodes.data(), g_res.bfs_nodes.size() * sizeof(BVHNode)); }
void ref_tris(uint32_t* out) { std::memcpy(out, g_res.tris.data(), g_res.tris.size() * sizeof(Triangle)); }
void ref_prims(uint32_t* out) { std::memcpy(out, g_res.prims.data(), g_res.prims.size() * sizeof(uint32_t)); }
void ref_sorted(uint32_t* out) { std::memcpy(out, g_res.sorted.data(), g_res.sorted.size() * sizeof(uint32_t)); }
void ref_codes(uint64_t* out) { std::memcpy(out, g_res.codes.data(), g_res.codes.size() * sizeof(uint64_t)); }
void ref_scene_box(float* out) { std::memcpy(out, g_res.scene, sizeof(g_res.scene)); }

// Load reference-layout arrays into the tracer's state and run bvh::trace per ray.
// rays: 8 floats {ox,oy,oz,tmin,dx,dy,dz,tmax}. hits: {t,u,v,triIndex|0xFFFFFFFF}.
// nodes == NULL traces against the arrays loaded by the previous call.
int ref_trace(const void* nodes, uint32_t nn, const uint32_t* tris, uint32_t ntri,
              const float* verts6, uint32_t nv, const float* rays, size_t n,
              float* hits4, float* normals3) {
    using namespace bvh;
    if (nodes) {
        t_nodes.resize(nn);
        std::memcpy(t_nodes.data(), nodes, sizeof(BVHNode) * nn);
        t_tris.resize(ntri);
        std::memcpy(t_tris.data(), tris, sizeof(Triangle) * ntri);
        t_verts.resize(nv);
        std::memcpy(t_verts.data(), verts6, sizeof(Vertex) * nv);
    }
    for (size_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray;
        ray.origin = math::vec3(r[0], r[1], r[2]);
        ray.tmin = r[3];
        ray.dir = math::vec3(r[4], r[5], r[6]);
        ray.tmax = r[7];
        Hit h = trace(ray);
        hits4[4 * i] = h.t; hits4[4 * i + 1] = h.u; hits4[4 * i + 2] = h.v;
        uint32_t id = h.hit ? h.triIndex : 0xFFFFFFFFu;
        std::memcpy(&hits4[4 * i + 3], &id, 4);
        if (normals3) { normals3[3 * i] = h.normal.x; normals3[3 * i + 1] = h.normal.y; normals3[3 * i + 2] = h.normal.z; }
    }
    return 0;
}

// bvh_geo.cpp generators into fresh buffers (used to pin the scene front-end).
int ref_geo(int kind, const float* args, float* verts6, uint32_t* nv, uint32_t* tris, uint32_t* nt) {
    using namespace bvh;
    s_vertices.clear();
    s_triangles.clear();
    if (kind == 0) {
        addSphere(math::vec3(args[0], args[1], args[2]), args[3], (int)args[4], (int)args[5]);
    } else {
        addBox(math::vec3(args[0], args[1], args[2]), math::vec3(args[3], args[4], args[5]),
               math::quat(args[9], args[6], args[7], args[8]));
    }
    *nv = (uint32_t)s_vertices.size();
    *nt = (uint32_t)s_triangles.size();
    if (verts6) std::memcpy(verts6, s_vertices.data(), sizeof(Vertex) * s_vertices.size());
    if (tris) std::memcpy(tris, s_triangles.data(), sizeof(Triangle) * s_triangles.size());
    s_vertices.clear();
    s_triangles.clear();
    return 0;
}

// clipTriangleToAABB (AABB.cpp:86-171): v 9 floats, region 6 floats, out 6 floats; returns valid().
int ref_clip(const float* v, const float* region, float* out) {
    AABB r;
    r.min = math::vec3(region[0], region[1], region[2]);
    r.max = math::vec3(region[3], region[4], region[5]);
    AABB b = clipTriangleToAABB(math::vec3(v[0], v[1], v[2]), math::vec3(v[3], v[4], v[5]),
                                math::vec3(v[6], v[7], v[8]), r);
    out[0] = b.min.x; out[1] = b.min.y; out[2] = b.min.z; out[3] = b.max.x; out[4] = b.max.y; out[5] = b.max.z;
    return b.valid() ? 1 : 0;
}

}  // extern "C"
