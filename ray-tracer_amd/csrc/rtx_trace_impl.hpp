// Traversal device code shared by the trace entry points and the path tracer.
// See rtx_trace.hip for the visiting-order argument.
#pragma once
#include <type_traits>

#include "rtx_internal.hpp"

namespace rtx {

constexpr int kTraceThreads = 256;
constexpr int kLdsDepthMax = 16;       // LDS ring entries per lane: 8, 12 or 16 (a tuning choice)
constexpr int kSpillDepth = 128;
constexpr int kTraceTriBatch = 4;   // triangles of a leaf loaded together by the trace kernels

struct TraceArgs {
    const GNode* nodes;
    const Leaf* leaves;
    const GTri* tris;
    const uint32_t* tris3;
    const float* verts;
    float rb[6];
    uint32_t root_ref;
    uint32_t finite;    // every node box coordinate is finite
    uint32_t wait;      // while-while: lanes still at internal nodes the leaf phase need not wait for
};

extern rtx_tuning g_tuning;

// Calls f(integral_constant<DEPTH>, integral_constant<LOOP>) for LDS depth d (8, 12, 16)
// and loop shape l (0, or 1 for the while-while loop; the persistent trace kernel is
// dispatched by its caller and reuses the while-while loop's depth instantiation).
template <class F>
void with_trace_variant(uint32_t d, uint32_t l, F&& f) {
    using I8 = std::integral_constant<int, 8>;
    using I12 = std::integral_constant<int, 12>;
    using I16 = std::integral_constant<int, 16>;
    using L0 = std::integral_constant<int, 0>;
    using L1 = std::integral_constant<int, 1>;
    if (d == 8) { if (l) f(I8{}, L1{}); else f(I8{}, L0{}); }
    else if (d == 12) { if (l) f(I12{}, L1{}); else f(I12{}, L0{}); }
    else { if (l) f(I16{}, L1{}); else f(I16{}, L0{}); }
}

// Deepest tree every variant can traverse (smallest LDS ring + spill).
constexpr uint32_t kMaxTraceDepth = 8 + kSpillDepth - 2;

inline TraceArgs make_trace_args(const rtx_bvh* b) {
    TraceArgs A;
    A.nodes = b->gnodes;
    A.leaves = b->leaves;
    A.tris = b->gtris;
    A.tris3 = b->tris3;
    A.verts = b->verts;
    for (int k = 0; k < 6; ++k) A.rb[k] = b->root_box[k];
    A.root_ref = b->root_ref;
    A.finite = b->boxes_finite ? 1u : 0u;
    A.wait = g_tuning.wait_lanes;
    return A;
}

// Map hardware workgroup ids (dealt round-robin over the 8 XCDs) to logical ids so that
// each XCD, with its own L2, works on one contiguous range of work.
__device__ __forceinline__ uint32_t xcd_block(uint32_t g) {
    const uint32_t G = gridDim.x, per = G / 8, rem = G % 8, x = g % 8, j = g / 8;
    return x < rem ? x * (per + 1) + j : rem * (per + 1) + (x - rem) * per + j;
}

struct TraceOut {
    float t, u, v;
    uint32_t tri;       // kInvalid on a miss
    uint32_t nI, nT;    // internal visits, triangle tests
};

// Traversal state of one lane. lst is the block's LDS stack, DEPTH entries per lane in
// column tid; entries beyond that spill (oldest first) to private memory. TB triangles of
// a leaf are loaded together.
template <int DEPTH, int TB = kTraceTriBatch>
struct Walker {
    f3 o, d;
    float ix, iy, iz, tmin, tmax;
    float bestT, bu, bv;
    uint32_t best, ref, top, lo, nI, nT;
    bool live, fast;
    uint2* spill;   // kSpillDepth private entries; kept outside the struct so that the
                    // fields above stay in registers

    __device__ __forceinline__ uint32_t slot(uint32_t e) const {
        return DEPTH == 16 || DEPTH == 8 ? (e & (DEPTH - 1)) : (e % DEPTH);
    }
    __device__ __forceinline__ void start(const TraceArgs& A, f3 o_, f3 d_, float tmin_, float tmax_) {
        o = o_; d = d_; tmin = tmin_; tmax = tmax_;
        ix = 1.f / d.x; iy = 1.f / d.y; iz = 1.f / d.z;
        bestT = tmax; bu = 0.f; bv = 0.f;
        best = kInvalid; top = 0; lo = 0; nI = 0; nT = 0;
        ref = A.root_ref;
        fast = A.finite && __builtin_isfinite(ix) && __builtin_isfinite(iy) && __builtin_isfinite(iz) &&
               ix != 0.f && iy != 0.f && iz != 0.f && __builtin_isfinite(o.x) && __builtin_isfinite(o.y) &&
               __builtin_isfinite(o.z) && !__builtin_isnan(tmin) && !__builtin_isnan(tmax);
        live = slab(o.x, o.y, o.z, ix, iy, iz, tmin, A.rb[0], A.rb[1], A.rb[2], A.rb[3], A.rb[4], A.rb[5], bestT) < bestT;
    }
    __device__ __forceinline__ void push(uint2 (*lst)[kTraceThreads], uint32_t tid, uint32_t r, float t) {
        if (top - lo == (uint32_t)DEPTH) { spill[lo] = lst[slot(lo)][tid]; ++lo; }
        lst[slot(top)][tid] = make_uint2(r, __float_as_uint(t));
        ++top;
    }
    // pop the next entry that can still hold a closer hit
    __device__ __forceinline__ void pop(uint2 (*lst)[kTraceThreads], uint32_t tid) {
        live = false;
        while (top > 0) {
            --top;
            uint2 e;
            if (top < lo) { --lo; e = spill[top]; }
            else e = lst[slot(top)][tid];
            if (__uint_as_float(e.y) < bestT) { ref = e.x; live = true; break; }
        }
    }
    // one internal node: test both children, push the far one, continue with the near one
    template <bool COUNT>
    __device__ __forceinline__ void node_step(const TraceArgs& A, uint2 (*lst)[kTraceThreads], uint32_t tid) {
        if (COUNT) ++nI;
        const GNode& g = A.nodes[ref];
        const float4 l0 = g.l0, l1 = g.l1, r1b = g.r1;
        const uint2 refs = *reinterpret_cast<const uint2*>(&g.lref);
        float tl, tr;
        if (__all(fast)) {   // wave vote: a scalar branch, both forms give the same values
            tl = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, bestT);
            tr = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, l1.z, l1.w, r1b.x, r1b.y, r1b.z, r1b.w, bestT);
        } else {
            tl = slab(o.x, o.y, o.z, ix, iy, iz, tmin, l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, bestT);
            tr = slab(o.x, o.y, o.z, ix, iy, iz, tmin, l1.z, l1.w, r1b.x, r1b.y, r1b.z, r1b.w, bestT);
        }
        uint32_t nearRef, farRef;
        float tn, tf;
        if (tl <= tr) { nearRef = refs.x; tn = tl; farRef = refs.y; tf = tr; }
        else { nearRef = refs.y; tn = tr; farRef = refs.x; tf = tl; }
        if (tn < bestT) {
            if (tf < bestT) push(lst, tid, farRef, tf);
            ref = nearRef;
        } else {
            pop(lst, tid);
        }
    }
    // Node phase of the while-while loop: node steps for the lanes at internal nodes until
    // none is left, or until at most `wait` are left while some lane already waits at a
    // leaf (those lanes then continue in a later phase). Every vote is wave-uniform; each
    // lane's own sequence of visits is unchanged.
    // `wait` must be the same on every lane (a kernel argument or a uniform field).
    template <bool COUNT>
    __device__ __forceinline__ void node_phase(const TraceArgs& A, uint2 (*lst)[kTraceThreads], uint32_t tid,
                                               uint32_t wait) {
        for (;;) {
            const bool need = live && !(ref & kLeafBit);
            const uint64_t nm = __ballot(need);
            if (nm == 0) break;
            if ((uint32_t)__popcll(nm) <= wait && __ballot(live && (ref & kLeafBit)) != 0) break;
            if (need) node_step<COUNT>(A, lst, tid);
        }
    }
    // All triangles of the current leaf, in order; true when an any-hit query is answered.
    // Triangles are loaded TB at a time before any of them is tested, so a leaf costs
    // ceil(count / TB) memory round trips instead of count. Loads past the
    // leaf's end re-read its last triangle (always in bounds) and are not tested.
    template <bool ANY, bool COUNT>
    __device__ __forceinline__ bool leaf_step(const TraceArgs& A) {
        const Leaf lf = A.leaves[ref & ~kLeafBit];
        const uint32_t end = lf.begin + lf.count;
        for (uint32_t k = lf.begin; k < end; k += TB) {
            float4 ta[TB], tb[TB];
            float tc[TB];
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                const uint32_t kk = min(k + (uint32_t)j, end - 1u);
                ta[j] = A.tris[kk].a;
                tb[j] = A.tris[kk].b;
                tc[j] = A.tris[kk].c.x;
            }
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                if (k + (uint32_t)j < end) {
                    if (COUNT) ++nT;
                    float u, v;
                    const float4 a = ta[j], b = tb[j];
                    const float t = tri_test(o, d, tmin, tmax, f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y},
                                             f3{b.z, b.w, tc[j]}, u, v);
                    if (t < bestT) {
                        bestT = t; bu = u; bv = v; best = k + (uint32_t)j;
                        if (ANY) return true;
                    }
                }
            }
        }
        return false;
    }
};

// One ray against one tree. LOOP 0 visits nodes and leaves in one loop; LOOP 1
// ("while-while") lets a lane that reaches a leaf wait until the other lanes of the wave
// are at leaves too (all but A.wait of them), so triangle tests run with the wave mostly
// converged. Both visit the same nodes in the same order per ray.
template <bool ANY, bool COUNT, int DEPTH, int LOOP, int TB = kTraceTriBatch>
__device__ __forceinline__ TraceOut trace_ray(const TraceArgs& A, f3 o, f3 d, float tmin, float tmax,
                                              uint2 (*lst)[kTraceThreads], uint32_t tid, uint32_t wait) {
    Walker<DEPTH, TB> w;
    uint2 spill[kSpillDepth];
    w.spill = spill;
    w.start(A, o, d, tmin, tmax);
    if (LOOP == 0) {
        while (w.live) {
            if (w.ref & kLeafBit) {
                if (w.template leaf_step<ANY, COUNT>(A)) break;
                w.pop(lst, tid);
            } else {
                w.template node_step<COUNT>(A, lst, tid);
            }
        }
    } else {
        for (;;) {
            w.template node_phase<COUNT>(A, lst, tid, wait);
            if (!__any(w.live)) break;
            if (w.live && (w.ref & kLeafBit)) {
                if (w.template leaf_step<ANY, COUNT>(A)) w.live = false;
                else w.pop(lst, tid);
            }
        }
    }
    TraceOut r;
    r.t = w.best != kInvalid ? w.bestT : 0.f;
    r.u = w.bu;
    r.v = w.bv;
    r.tri = w.best;
    r.nI = w.nI;
    r.nT = w.nT;
    return r;
}

// Interpolated shading normal of a hit (bvh_traverse.cpp:135-139).
__device__ __forceinline__ f3 hit_normal(const TraceArgs& A, uint32_t tri, float u, float v) {
    const float* V = A.verts;
    uint32_t i0 = A.tris3[3 * tri], i1 = A.tris3[3 * tri + 1], i2 = A.tris3[3 * tri + 2];
    f3 n0{V[6 * i0 + 3], V[6 * i0 + 4], V[6 * i0 + 5]};
    f3 n1{V[6 * i1 + 3], V[6 * i1 + 4], V[6 * i1 + 5]};
    f3 n2{V[6 * i2 + 3], V[6 * i2 + 4], V[6 * i2 + 5]};
    float w = 1.f - u - v;
    return normalize(add(add(scl(w, n0), scl(u, n1)), scl(v, n2)));
}

}  // namespace rtx
