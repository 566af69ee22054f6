// Closest-hit and any-hit traversal (src/bvh/bvh_traverse.cpp:22-178) for gfx950.
//
// One lane per ray. The node being visited lives in registers; the stack of pending
// (node, entry distance) pairs is a per-lane ring of 8-16 entries in LDS, with the
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

#include <algorithm>

#include "rtx_trace_impl.hpp"

namespace rtx {
namespace {

template <bool ANY, bool COUNT, int DEPTH, int LOOP>
__global__ void __launch_bounds__(kTraceThreads) k_trace(TraceArgs A, const Ray32* __restrict__ rays, uint32_t n,
                                                         Hit16* __restrict__ hits, float* __restrict__ normals,
                                                         uint8_t* __restrict__ occl, uint32_t* __restrict__ counts) {
    __shared__ uint2 lst[DEPTH][kTraceThreads];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = xcd_block(blockIdx.x) * kTraceThreads + tid;
    // out-of-range lanes still run (with a ray that misses) so wave-wide votes see them
    const bool in = i < n;
    float4 r0 = make_float4(0.f, 0.f, 0.f, 1.f), r1 = make_float4(0.f, 0.f, 1.f, 0.f);
    if (in) {
        r0 = reinterpret_cast<const float4*>(rays)[2 * i];
        r1 = reinterpret_cast<const float4*>(rays)[2 * i + 1];
    }
    const TraceOut h = trace_ray<ANY, COUNT, DEPTH, LOOP>(A, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, r0.w, r1.w,
                                                          lst, tid);
    if (!in) return;
    if (ANY) {
        occl[i] = h.tri != kInvalid ? 1 : 0;
        return;
    }
    reinterpret_cast<float4*>(hits)[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.tri));
    if (COUNT) {
        counts[2 * i] = h.nI;
        counts[2 * i + 1] = h.nT;
    }
    if (normals) {
        f3 nn{0.f, 0.f, 0.f};
        if (h.tri != kInvalid) nn = hit_normal(A, h.tri, h.u, h.v);
        normals[3 * i] = nn.x;
        normals[3 * i + 1] = nn.y;
        normals[3 * i + 2] = nn.z;
    }
}

// Persistent variant (tuning trace_loop = 2): a fixed grid of waves; when at least
// kRefill lanes of a wave are idle, one atomic on `work` hands them the next rays. Each
// ray is still traversed exactly as trace_ray's while-while loop does.
constexpr int kRefill = 16;

template <bool ANY, bool COUNT, int DEPTH>
__global__ void __launch_bounds__(kTraceThreads) k_trace_persist(TraceArgs A, const Ray32* __restrict__ rays,
                                                                 uint32_t n, Hit16* __restrict__ hits,
                                                                 float* __restrict__ normals,
                                                                 uint8_t* __restrict__ occl,
                                                                 uint32_t* __restrict__ counts, uint32_t* work) {
    __shared__ uint2 lst[DEPTH][kTraceThreads];
    const uint32_t tid = threadIdx.x;
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    Walker<DEPTH> w;
    w.live = false;
    w.best = kInvalid;
    uint32_t id = kInvalid;
    bool more = true;
    for (;;) {
        if (!w.live && id != kInvalid) {
            if (ANY) {
                occl[id] = w.best != kInvalid ? 1 : 0;
            } else {
                reinterpret_cast<float4*>(hits)[id] =
                    make_float4(w.best != kInvalid ? w.bestT : 0.f, w.bu, w.bv, __uint_as_float(w.best));
                if (COUNT) { counts[2 * id] = w.nI; counts[2 * id + 1] = w.nT; }
                if (normals) {
                    f3 nn{0.f, 0.f, 0.f};
                    if (w.best != kInvalid) nn = hit_normal(A, w.best, w.bu, w.bv);
                    normals[3 * id] = nn.x; normals[3 * id + 1] = nn.y; normals[3 * id + 2] = nn.z;
                }
            }
            id = kInvalid;
        }
        const uint64_t idle = __ballot(!w.live);
        if (more && __popcll(idle) >= kRefill) {
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t base = 0;
            if (__lane_id() == leader) base = atomicAdd(work, (uint32_t)__popcll(idle));
            base = __shfl(base, leader);
            if (base + __popcll(idle) >= n) more = false;
            if (!w.live) {
                const uint32_t k = base + __popcll(idle & lt);
                if (k < n) {
                    id = k;
                    const float4 r0 = reinterpret_cast<const float4*>(rays)[2 * k];
                    const float4 r1 = reinterpret_cast<const float4*>(rays)[2 * k + 1];
                    w.start(A, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, r0.w, r1.w);
                    continue;   // a ray that misses the root is written at the loop top
                }
            }
        }
        if (!__any(w.live)) {
            if (!more && !__any(id != kInvalid)) break;
            continue;
        }
        while (w.live && !(w.ref & kLeafBit)) w.template node_step<COUNT>(A, lst, tid);
        if (w.live) {
            if (w.template leaf_step<ANY, COUNT>(A)) w.live = false;
            else w.pop(lst, tid);
        }
    }
}

__global__ void k_zero_work(uint32_t* w) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *w = 0u;
}

int check(const rtx_bvh* b, const void* rays, size_t n, const void* out, const char* fn) {
    if (!b || !out || (!rays && n)) return fail(RTX_E_ARG, std::string("null argument on ") + fn);
    if (!b->gnodes || !b->leaves || !b->gtris) return fail(RTX_E_STATE, std::string("BVH not loaded on ") + fn);
    if (n > (1ull << 31)) return fail(RTX_E_LIMIT, std::string("more than 2^31 rays on ") + fn);
    if (b->max_depth > kMaxTraceDepth)
        return fail(RTX_E_LIMIT, std::string("tree deeper than the traversal stack on ") + fn);
    return RTX_OK;
}

template <bool ANY, bool COUNT>
int launch(const rtx_bvh* b, const rtx_ray* rays, size_t n, rtx_hit* hits, float* normals, uint8_t* occl,
           uint32_t* counts, hipStream_t s) {
    if (n == 0) return RTX_OK;
    uint32_t blocks = (uint32_t)((n + kTraceThreads - 1) / kTraceThreads);
    const TraceArgs A = make_trace_args(b);
    if (g_tuning.trace_loop == 2 && b->work) {
        const uint32_t pb = std::min<uint32_t>(blocks, g_tuning.grid_blocks);
        k_zero_work<<<1, 64, 0, s>>>(b->work);
        with_trace_variant([&](auto D, auto) {
            k_trace_persist<ANY, COUNT, decltype(D)::value><<<pb, kTraceThreads, 0, s>>>(
                A, reinterpret_cast<const Ray32*>(rays), (uint32_t)n, reinterpret_cast<Hit16*>(hits), normals, occl,
                counts, b->work);
        });
        RTX_HIP(hipGetLastError());
        return RTX_OK;
    }
    with_trace_variant([&](auto D, auto L) {
        k_trace<ANY, COUNT, decltype(D)::value, decltype(L)::value><<<blocks, kTraceThreads, 0, s>>>(
            A, reinterpret_cast<const Ray32*>(rays), (uint32_t)n, reinterpret_cast<Hit16*>(hits), normals, occl,
            counts);
    });
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

}  // namespace

rtx_tuning g_tuning = {1u << 22, 2048, 256, 8, 1};

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
