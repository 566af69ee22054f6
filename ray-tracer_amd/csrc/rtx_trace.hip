// Closest-hit and any-hit traversal (src/bvh/bvh_traverse.cpp:22-178) for gfx950.
//
// One lane per ray. The node being visited lives in registers; the stack of pending
// (node, entry distance) pairs is a per-lane ring of kLdsDepth entries in LDS, with the
// oldest entries spilled to private memory when it fills, so any tree depth the builder
// can produce is handled (the reference's fixed stack[64] has no overflow check).
//
// Visiting order is the reference's: at an internal node both child boxes are tested
// against the current best t, the farther child is pushed and the nearer one visited
// (left first on a tie). A popped entry is skipped when its stored entry distance is
// >= best t, which equals the reference's re-test of the popped box (tNear does not
// depend on tmax, and tNear <= tFar held when it was pushed). Triangles are tested in
// leaf order with a strict t < best update, so exact-t ties resolve as in the reference.
#include <hip/hip_runtime.h>

#include "rtx_internal.hpp"

namespace rtx {
namespace {

constexpr int kTraceThreads = 256;
constexpr int kLdsDepth = 16;          // power of two
constexpr int kSpillDepth = 128;

struct TraceArgs {
    const GNode* nodes;
    const Leaf* leaves;
    const GTri* tris;
    const uint32_t* tris3;
    const float* verts;
    float rb[6];
    uint32_t root_ref;
};

// Map hardware workgroup ids (dealt round-robin over the 8 XCDs) to logical ids so that
// each XCD, with its own L2, works on one contiguous range of rays.
__device__ __forceinline__ uint32_t xcd_block(uint32_t g) {
    const uint32_t G = gridDim.x, per = G / 8, rem = G % 8, x = g % 8, j = g / 8;
    return x < rem ? x * (per + 1) + j : rem * (per + 1) + (x - rem) * per + j;
}

template <bool ANY, bool COUNT>
__global__ void __launch_bounds__(kTraceThreads) k_trace(TraceArgs A, const Ray32* __restrict__ rays, uint32_t n,
                                                         Hit16* __restrict__ hits, float* __restrict__ normals,
                                                         uint8_t* __restrict__ occl, uint32_t* __restrict__ counts) {
    __shared__ uint2 lst[kLdsDepth][kTraceThreads];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = xcd_block(blockIdx.x) * kTraceThreads + tid;
    if (i >= n) return;
    const float4 r0 = reinterpret_cast<const float4*>(rays)[2 * i];
    const float4 r1 = reinterpret_cast<const float4*>(rays)[2 * i + 1];
    const f3 o{r0.x, r0.y, r0.z}, d{r1.x, r1.y, r1.z};
    const float tmin = r0.w, tmax = r1.w;
    const float ix = 1.f / d.x, iy = 1.f / d.y, iz = 1.f / d.z;
    uint2 spill[kSpillDepth];
    uint32_t top = 0, lo = 0;
    float bestT = tmax, bu = 0.f, bv = 0.f;
    uint32_t best = kInvalid;
    uint32_t nI = 0, nT = 0;

    uint32_t ref = A.root_ref;
    bool live = slab(o.x, o.y, o.z, ix, iy, iz, tmin, A.rb[0], A.rb[1], A.rb[2], A.rb[3], A.rb[4], A.rb[5], bestT) < bestT;
    while (live) {
        if (ref & kLeafBit) {
            const Leaf lf = A.leaves[ref & ~kLeafBit];
            for (uint32_t k = lf.begin; k < lf.begin + lf.count; ++k) {
                const float4 a = A.tris[k].a, b = A.tris[k].b, c = A.tris[k].c;
                if (COUNT) ++nT;
                float u, v;
                float t = tri_test(o, d, tmin, tmax, f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, f3{b.z, b.w, c.x}, u, v);
                if (t < bestT) {
                    bestT = t; bu = u; bv = v; best = k;
                    if (ANY) break;
                }
            }
            if (ANY && best != kInvalid) break;
        } else {
            if (COUNT) ++nI;
            const GNode& g = A.nodes[ref];
            const float4 l0 = g.l0, l1 = g.l1, r1b = g.r1;
            const uint2 refs = *reinterpret_cast<const uint2*>(&g.lref);
            const float tl = slab(o.x, o.y, o.z, ix, iy, iz, tmin, l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, bestT);
            const float tr = slab(o.x, o.y, o.z, ix, iy, iz, tmin, l1.z, l1.w, r1b.x, r1b.y, r1b.z, r1b.w, bestT);
            uint32_t nearRef, farRef;
            float tn, tf;
            if (tl <= tr) { nearRef = refs.x; tn = tl; farRef = refs.y; tf = tr; }
            else { nearRef = refs.y; tn = tr; farRef = refs.x; tf = tl; }
            if (tn < bestT) {
                if (tf < bestT) {
                    if (top - lo == (uint32_t)kLdsDepth) { spill[lo] = lst[lo & (kLdsDepth - 1)][tid]; ++lo; }
                    lst[top & (kLdsDepth - 1)][tid] = make_uint2(farRef, __float_as_uint(tf));
                    ++top;
                }
                ref = nearRef;
                continue;
            }
        }
        // pop the next entry that can still hold a closer hit
        live = false;
        while (top > 0) {
            --top;
            uint2 e;
            if (top < lo) { --lo; e = spill[top]; }
            else e = lst[top & (kLdsDepth - 1)][tid];
            if (__uint_as_float(e.y) < bestT) { ref = e.x; live = true; break; }
        }
    }

    if (ANY) {
        occl[i] = best != kInvalid ? 1 : 0;
        return;
    }
    Hit16 h;
    h.t = best != kInvalid ? bestT : 0.f;
    h.u = bu;
    h.v = bv;
    h.tri = best;
    reinterpret_cast<float4*>(hits)[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.tri));
    if (COUNT) {
        counts[2 * i] = nI;
        counts[2 * i + 1] = nT;
    }
    if (normals) {
        f3 nn{0.f, 0.f, 0.f};
        if (best != kInvalid) {
            // Interpolated shading normal of the final hit (bvh_traverse.cpp:135-139).
            const float* V = A.verts;
            uint32_t i0 = A.tris3[3 * best], i1 = A.tris3[3 * best + 1], i2 = A.tris3[3 * best + 2];
            f3 n0{V[6 * i0 + 3], V[6 * i0 + 4], V[6 * i0 + 5]};
            f3 n1{V[6 * i1 + 3], V[6 * i1 + 4], V[6 * i1 + 5]};
            f3 n2{V[6 * i2 + 3], V[6 * i2 + 4], V[6 * i2 + 5]};
            float w = 1.f - bu - bv;
            nn = normalize(add(add(scl(w, n0), scl(bu, n1)), scl(bv, n2)));
        }
        normals[3 * i] = nn.x;
        normals[3 * i + 1] = nn.y;
        normals[3 * i + 2] = nn.z;
    }
}

TraceArgs make_args(const rtx_bvh* b) {
    TraceArgs A;
    A.nodes = b->gnodes;
    A.leaves = b->leaves;
    A.tris = b->gtris;
    A.tris3 = b->tris3;
    A.verts = b->verts;
    for (int k = 0; k < 6; ++k) A.rb[k] = b->root_box[k];
    A.root_ref = b->root_ref;
    return A;
}

int check(const rtx_bvh* b, const void* rays, size_t n, const void* out, const char* fn) {
    if (!b || !out || (!rays && n)) return fail(RTX_E_ARG, std::string("null argument on ") + fn);
    if (!b->gnodes || !b->leaves || !b->gtris) return fail(RTX_E_STATE, std::string("BVH not loaded on ") + fn);
    if (n > 0xFFFFFFFFull - kTraceThreads) return fail(RTX_E_LIMIT, std::string("too many rays on ") + fn);
    if (b->max_depth + 2 > (uint32_t)(kLdsDepth + kSpillDepth))
        return fail(RTX_E_LIMIT, std::string("tree deeper than the traversal stack on ") + fn);
    return RTX_OK;
}

template <bool ANY, bool COUNT>
int launch(const rtx_bvh* b, const rtx_ray* rays, size_t n, rtx_hit* hits, float* normals, uint8_t* occl,
           uint32_t* counts, hipStream_t s) {
    if (n == 0) return RTX_OK;
    uint32_t blocks = (uint32_t)((n + kTraceThreads - 1) / kTraceThreads);
    k_trace<ANY, COUNT><<<blocks, kTraceThreads, 0, s>>>(make_args(b), reinterpret_cast<const Ray32*>(rays),
                                                         (uint32_t)n, reinterpret_cast<Hit16*>(hits), normals, occl,
                                                         counts);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

}  // namespace
}  // namespace rtx

using namespace rtx;

extern "C" int rtx_trace_closest(const rtx_bvh* b, const rtx_ray* rays, size_t n, rtx_hit* hits, float* normals,
                                 void* stream) {
    int rc = check(b, rays, n, hits, "rtx_trace_closest()");
    if (rc) return rc;
    return launch<false, false>(b, rays, n, hits, normals, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int rtx_trace_any(const rtx_bvh* b, const rtx_ray* rays, size_t n, uint8_t* occluded, void* stream) {
    int rc = check(b, rays, n, occluded, "rtx_trace_any()");
    if (rc) return rc;
    return launch<true, false>(b, rays, n, nullptr, nullptr, occluded, nullptr, (hipStream_t)stream);
}

extern "C" int rtx_trace_counts(const rtx_bvh* b, const rtx_ray* rays, size_t n, uint32_t* counts, void* stream) {
    int rc = check(b, rays, n, counts, "rtx_trace_counts()");
    if (rc) return rc;
    rtx_hit* tmp = dalloc<rtx_hit>(n);
    if (!tmp) return fail(RTX_E_HIP, "device allocation failed on rtx_trace_counts()");
    rc = launch<false, true>(b, rays, n, tmp, nullptr, nullptr, counts, (hipStream_t)stream);
    if (rc == RTX_OK) rc = hip_ok(hipStreamSynchronize((hipStream_t)stream), "sync") ? RTX_OK : RTX_E_HIP;
    (void)hipFree(tmp);
    return rc;
}

extern "C" int rtx_trace_closest_host(const rtx_bvh* b, const rtx_ray* rays, size_t n, rtx_hit* hits,
                                      float* normals) {
    int rc = check(b, rays, n, hits, "rtx_trace_closest_host()");
    if (rc) return rc;
    if (n == 0) return RTX_OK;
    rtx_ray* dr = dalloc<rtx_ray>(n);
    rtx_hit* dh = dalloc<rtx_hit>(n);
    float* dn = normals ? dalloc<float>(3 * n) : nullptr;
    rc = (!dr || !dh || (normals && !dn)) ? fail(RTX_E_HIP, "device allocation failed on rtx_trace_closest_host()")
                                          : RTX_OK;
    if (rc == RTX_OK && !hip_ok(hipMemcpy(dr, rays, sizeof(rtx_ray) * n, hipMemcpyHostToDevice), "copy")) rc = RTX_E_HIP;
    if (rc == RTX_OK) rc = launch<false, false>(b, dr, n, dh, dn, nullptr, nullptr, nullptr);
    if (rc == RTX_OK && !hip_ok(hipMemcpy(hits, dh, sizeof(rtx_hit) * n, hipMemcpyDeviceToHost), "copy")) rc = RTX_E_HIP;
    if (rc == RTX_OK && normals && !hip_ok(hipMemcpy(normals, dn, 12 * n, hipMemcpyDeviceToHost), "copy")) rc = RTX_E_HIP;
    if (dr) (void)hipFree(dr);
    if (dh) (void)hipFree(dh);
    if (dn) (void)hipFree(dn);
    return rc;
}
