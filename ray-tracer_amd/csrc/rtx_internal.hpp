// Internal state of an rtx_bvh handle and helpers shared by the librtx sources.
#pragma once
#include <mutex>
#include <string>
#include <vector>

#include "rtx_common.hpp"
#include "../../include/rtx.h"

namespace rtx {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
bool hip_ok(hipError_t e, const char* what);

#define RTX_HIP(call)                                   \
    do {                                                \
        if (!::rtx::hip_ok((call), #call)) return RTX_E_HIP; \
    } while (0)

// Process-wide cache of device scratch blocks (rtx_build.hip): scratch_take returns a
// cached block of at least `bytes` (its size in `got`) or a new allocation, and orders the
// work later queued on `st` after the block's last user; scratch_give returns an idle block,
// scratch_give_after one that work queued on `st` may still use. Saves hipMalloc/hipFree
// (which synchronizes the device) on every build and render.
void* scratch_take(size_t bytes, size_t& got, hipStream_t st);
void scratch_give(void* p, size_t bytes);
void scratch_give_after(void* p, size_t bytes, hipStream_t st);
// Peak of the bytes requested by blocks taken and not yet given back, since the last reset.
uint64_t scratch_peak(bool reset);

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    return (T*)p;
}

}  // namespace rtx

struct rtx_bvh {
    int device = 0;
    uint32_t nv = 0, nt = 0;          // input vertices / triangles
    float* verts = nullptr;           // device copy of the 6-float vertices
    // products, DFS order
    rtx::Node48* nodes48 = nullptr;   // reference layout
    uint32_t nn = 0;
    uint32_t* tris3 = nullptr;        // reordered triangles (vertex indices)
    uint32_t* prims = nullptr;        // original triangle id per reordered triangle
    uint32_t ntri = 0;
    rtx::GNode* gnodes = nullptr;     // traversal layout, internal nodes only
    uint32_t ninternal = 0;
    rtx::WNode* wnodes = nullptr;     // 4-wide traversal layout, indexed like gnodes (entries
                                      // of binary nodes absorbed into a wide parent are unused)
    rtx::Leaf* leaves = nullptr;
    uint32_t nleaves = 0;
    rtx::GTri* gtris = nullptr;
    float root_box[6] = {0, 0, 0, 0, 0, 0};
    uint32_t root_ref = 0;
    uint32_t* work = nullptr;         // ray counter of the persistent trace kernel
    bool boxes_finite = false;        // every node box coordinate finite (fast slab test)
    // ray-sort scratch (grow-only)
    uint32_t* rs_keys[2] = {nullptr, nullptr};
    uint32_t* rs_vals[2] = {nullptr, nullptr};
    uint32_t* rs_hist = nullptr;
    uint32_t* rs_part = nullptr;
    uint32_t rs_cap = 0;
    // Trace calls that use the handle's scratch (ray counter, sort buffers) are ordered on
    // the GPU: each waits for the previous user's event on its own stream.
    std::mutex scratch_mu;
    hipEvent_t scratch_ev = nullptr;
    bool scratch_used = false;
    // device staging of rtx_trace_closest_host (grow-only, one caller at a time)
    std::mutex host_mu;
    void* host_rays = nullptr;
    void* host_hits = nullptr;
    float* host_nrm = nullptr;
    size_t host_cap = 0;
    // small rtx_trace_closest_host calls: pinned host buffers (hipHostMalloc) that the
    // traversal kernel reads and writes directly, and the stream those calls run on
    void* pin_rays = nullptr;
    void* pin_hits = nullptr;
    float* pin_nrm = nullptr;
    void* pin_rays_d = nullptr;       // their device addresses (hipHostGetDevicePointer, once)
    void* pin_hits_d = nullptr;
    void* pin_nrm_d = nullptr;
    size_t pin_cap = 0;
    hipStream_t host_stream = nullptr;
    // diagnostics
    uint64_t* codes = nullptr;        // Morton code per input triangle
    uint32_t* sorted = nullptr;       // s_sortedTris[0..nt)
    rtx::Node48* bfs48 = nullptr;     // node array in builder (FIFO) order
    uint32_t max_depth = 0, levels = 0, tindex_size = 0, sorted_size = 0;
    float scene[6] = {0, 0, 0, 0, 0, 0};
};

namespace rtx {
// Allocates the traversal products of `b` from host arrays in the reference layout
// (used by rtx_load).  Defined in rtx_build.hip.
int upload_reference_tree(rtx_bvh* b, const Node48* nodes, uint32_t nn, const uint32_t* tris3,
                          uint32_t ntri, const float* verts6, uint32_t nv);
// Collapses the binary traversal nodes (b->gnodes) into b->wnodes. Defined in rtx_build.hip.
int build_wide(rtx_bvh* b, hipStream_t st);
}  // namespace rtx
