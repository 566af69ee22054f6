// Drop-in implementation of the reference's builder/tracer entry points on librtx.
//
// A program written against minsuk0519/ray-tracer compiles this file instead of
// src/bvh/{bvh_bake,bvh_morton,bvh_bfs,bvh_sah,bvh_traverse}.cpp and links librtx.so.
// The reference headers provide the types; this file provides:
//   bool bvh::bakeBVH(const std::string&)   (src/bvh/bvh.hpp:45-52)
//   bool bvh::loadBVH(const std::string&)   (src/bvh/bvh_traverse.hpp:28)
//   bvh::Hit bvh::trace(const bvh::Ray&)    (src/bvh/bvh_traverse.hpp:29)
// with the reference's bool + "Error : ... on fn()" failure behaviour, plus a batched
// bvh::traceBatch for callers that can hand over many rays at once.
#define BAKING
#include "bvh.hpp"
#include "bvh_traverse.hpp"

#include <cstdio>
#include <vector>

#include "rtx.h"

namespace bvh {
namespace {
rtx_bvh* g_tree = nullptr;
}

bool bakeBVH(const std::string& scenePath) {
    if (rtx_bake_scene(scenePath.c_str()) != RTX_OK) {
        std::fprintf(stderr, "Error : %s on bakeBVH()\n", rtx_last_error());
        return false;
    }
    return true;
}

bool loadBVH(const std::string& path) {
    rtx_bvh* t = nullptr;
    if (rtx_load(path.c_str(), &t) != RTX_OK) {
        std::fprintf(stderr, "Error : %s on loadBVH()\n", rtx_last_error());
        return false;
    }
    rtx_bvh_free(g_tree);
    g_tree = t;
    rtx_bvh_info info;
    rtx_bvh_get_info(t, &info);
    std::fprintf(stdout, "BVH loaded: %s  (%u nodes, %u tris, %u verts)\n", path.c_str(), info.node_count,
                 info.tri_count, info.vert_count);
    return true;
}

// Traces n rays in one GPU launch; hits[i] is what trace(rays[i]) returns.
bool traceBatch(const Ray* rays, size_t n, Hit* hits) {
    if (!g_tree) {
        std::fprintf(stderr, "Error : BVH not loaded on trace()\n");
        for (size_t i = 0; i < n; ++i) hits[i] = Hit{};
        return false;
    }
    std::vector<rtx_ray> r(n);
    for (size_t i = 0; i < n; ++i) {
        r[i] = rtx_ray{{rays[i].origin.x, rays[i].origin.y, rays[i].origin.z}, rays[i].tmin,
                       {rays[i].dir.x, rays[i].dir.y, rays[i].dir.z}, rays[i].tmax};
    }
    std::vector<rtx_hit> h(n);
    std::vector<float> nrm(3 * n);
    if (rtx_trace_closest_host(g_tree, r.data(), n, h.data(), nrm.data()) != RTX_OK) {
        std::fprintf(stderr, "Error : %s on trace()\n", rtx_last_error());
        for (size_t i = 0; i < n; ++i) hits[i] = Hit{};
        return false;
    }
    for (size_t i = 0; i < n; ++i) {
        Hit o;
        if (h[i].tri != 0xFFFFFFFFu) {
            o.hit = true;
            o.t = h[i].t;
            o.u = h[i].u;
            o.v = h[i].v;
            o.triIndex = h[i].tri;
            o.normal = math::vec3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        }
        hits[i] = o;
    }
    return true;
}

Hit trace(const Ray& ray) {
    Hit h;
    traceBatch(&ray, 1, &h);
    return h;
}

}  // namespace bvh
