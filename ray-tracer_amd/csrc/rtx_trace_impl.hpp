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
    const WNode* wnodes;   // 4-wide layout of the same tree (trace entry points, trace_tree 1)
    const Leaf* leaves;
    const GTri* tris;
    const uint32_t* tris3;
    const float* verts;
    float rb[6];
    uint32_t root_ref;
    uint32_t finite;    // every node box coordinate is finite
    uint32_t wait;      // while-while: lanes still at internal nodes the leaf phase need not wait for
    uint32_t refill;    // persistent trace kernel: idle lanes of a wave that trigger a refill
};

extern rtx_tuning g_tuning;

// Loop shapes of trace_ray: one loop over nodes and leaves; the while-while loop; the
// while-while loop with cooperative leaf phases.
constexpr int kLoopSingle = 0, kLoopWhileWhile = 1, kLoopCoop = 2;

// Trees a trace can walk: the binary tree (the reference's visiting order, so hits equal
// the reference's including which triangle wins an exact-t tie), or its 4-wide collapse
// (same t for every ray; at an exact-t tie between triangles of different leaves another
// of the tied triangles may be reported).
constexpr int kTreeBinary = 0, kTreeWide = 1;

// Order-preserving 32-bit image of a float t (t not NaN; -0 and +0 map alike).
__device__ __forceinline__ uint32_t ord_t(float t) {
    const uint32_t b = t == 0.f ? 0u : __float_as_uint(t);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Calls f(integral_constant<DEPTH>, integral_constant<LOOP>) for LDS depth d (8, 12, 16)
// and loop shape l (kLoopSingle, kLoopWhileWhile or kLoopCoop).
template <class F>
void with_trace_variant(uint32_t d, uint32_t l, F&& f) {
    auto by_loop = [&](auto D) {
        if (l == (uint32_t)kLoopCoop) f(D, std::integral_constant<int, kLoopCoop>{});
        else if (l == (uint32_t)kLoopWhileWhile) f(D, std::integral_constant<int, kLoopWhileWhile>{});
        else f(D, std::integral_constant<int, kLoopSingle>{});
    };
    if (d == 8) by_loop(std::integral_constant<int, 8>{});
    else if (d == 12) by_loop(std::integral_constant<int, 12>{});
    else by_loop(std::integral_constant<int, 16>{});
}

// Deepest tree every variant can traverse (smallest LDS ring + spill).
constexpr uint32_t kMaxTraceDepth = 8 + kSpillDepth - 2;
// The wide walk pushes up to three entries per wide node on the path and a wide node
// spans at least one binary level, so at most 3 * max_depth entries are ever pending; trees
// deeper than this take the binary walk.
constexpr uint32_t kMaxWideTraceDepth = (8 + kSpillDepth - 2) / 3;

inline TraceArgs make_trace_args(const rtx_bvh* b) {
    TraceArgs A;
    A.nodes = b->gnodes;
    A.wnodes = b->wnodes;
    A.leaves = b->leaves;
    A.tris = b->gtris;
    A.tris3 = b->tris3;
    A.verts = b->verts;
    for (int k = 0; k < 6; ++k) A.rb[k] = b->root_box[k];
    A.root_ref = b->root_ref;
    A.finite = b->boxes_finite ? 1u : 0u;
    A.wait = g_tuning.wait_lanes;
    A.refill = g_tuning.refill_lanes;
    return A;
}

// Map hardware workgroup ids (dealt round-robin over the 8 XCDs) to logical ids so that
// each XCD, with its own L2, works on one contiguous range of work.
__device__ __forceinline__ uint32_t xcd_block(uint32_t g) {
    const uint32_t G = gridDim.x, per = G / 8, rem = G % 8, x = g % 8, j = g / 8;
    return x < rem ? x * (per + 1) + j : rem * (per + 1) + (x - rem) * per + j;
}

// LDS of the cooperative leaf phase, per wave of 64 lanes.
// 576 bytes per wave: with a 12-entry stack ring a 256-thread block needs 26,880 bytes of
// LDS, so six blocks fit in a CU's 160 KiB (32-bit marks made it 27,648 and five blocks).
struct CoopLds {
    unsigned long long mins[64];  // per owner lane: smallest (t, triangle index) key
    uint8_t mark[64];             // lane whose pairs start at this position of the round
};
constexpr uint8_t kNoLane = 0xFF;

// Loads through global-memory pointers that the compiler cannot prove global (pointers
// read from memory, such as the renderer's per-scene descriptors) would otherwise be flat
// loads, which wait on both the vector-memory and the LDS counters. These load through an
// explicit global address-space pointer.
typedef float gfloat4 __attribute__((ext_vector_type(4)));
typedef uint32_t guint2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 ld_g(const float4* p) {
    const gfloat4 v = *(__attribute__((address_space(1))) const gfloat4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ld_g(const uint2* p) {
    const guint2 v = *(__attribute__((address_space(1))) const guint2*)p;
    return make_uint2(v.x, v.y);
}
typedef uint32_t guint4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 ld_g(const uint4* p) {
    const guint4 v = *(__attribute__((address_space(1))) const guint4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float ld_g(const float* p) { return *(__attribute__((address_space(1))) const float*)p; }
__device__ __forceinline__ uint32_t ld_g(const uint32_t* p) {
    return *(__attribute__((address_space(1))) const uint32_t*)p;
}

// Inclusive sum over the 64 lanes of a wave, lane order, every lane active: four row shifts
// within each 16-lane row, then row 0's and row 1's last lanes broadcast to the rows above
// (DPP, no LDS round trips; checked against the serial sum on gfx950).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += __builtin_amdgcn_update_dpp(0u, x, 0x111, 0xf, 0xf, false);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0u, x, 0x112, 0xf, 0xf, false);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0u, x, 0x114, 0xf, 0xf, false);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0u, x, 0x118, 0xf, 0xf, false);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0u, x, 0x142, 0xa, 0xf, false);   // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0u, x, 0x143, 0xc, 0xf, false);   // row_bcast:31 into rows 2, 3
    return x;
}

// Triangle range (begin, count) of a leaf reference: packed in the reference itself, or
// read from the leaf table.
__device__ __forceinline__ uint2 leaf_range(const Leaf* leaves, uint32_t ref) {
    if (ref & kInlineLeaf) return make_uint2(ref & ((1u << 25) - 1u), (ref >> 25) & 31u);
    return ld_g(reinterpret_cast<const uint2*>(leaves + (ref & ~kLeafBit)));
}

// LDS words shared by the lanes of one wave: relaxed wave-scope atomics, so the compiler
// keeps every access (a lane reads what another lane wrote) and still emits plain ds_*
// instructions (a volatile pointer would turn them into flat accesses). A wave's LDS
// operations complete in issue order.
template <class T>
__device__ __forceinline__ void wave_st(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
template <class T>
__device__ __forceinline__ T wave_ld(T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// The 36 bytes of a traversal triangle a test reads, loaded before the test starts: the
// empty asm takes the values as operands, so the compiler cannot sink the v0 load behind
// the back-face test (a second dependent memory round trip for the lanes that pass it).
__device__ __forceinline__ void ld_tri(const GTri* p, float4& a, float4& b, float& c) {
    const gtri_f4 x = ((const gtri_f4_global*)p->v)[0], y = ((const gtri_f4_global*)p->v)[1];
    a = make_float4(x.x, x.y, x.z, x.w);
    b = make_float4(y.x, y.y, y.z, y.w);
    c = ((const float_global*)p->v)[8];
}
__device__ __forceinline__ void load_tri(const GTri* p, float4& a, float4& b, float& c) {
    ld_tri(p, a, b, c);
    __asm__ volatile("" : : "v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(b.x), "v"(b.y), "v"(b.z), "v"(b.w), "v"(c));
}

// Orders (t, triangle index) as the in-order strict-less-than scan of a leaf does: smaller
// t first, then the earlier triangle. -0 and +0 compare equal, as they do as floats. t is
// never NaN here (a NaN t fails t < best).
__device__ __forceinline__ unsigned long long coop_key(float t, uint32_t k) {
    return ((unsigned long long)ord_t(t) << 32) | k;
}
// The float of an ord_t image (a zero comes back as +0).
__device__ __forceinline__ float t_of_ord(uint32_t b) {
    return __uint_as_float((b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b);
}

// tri_test's accept/reject decision and its t up to the sign of a zero t, for the dealt
// tests of the cooperative leaf phase (the owner re-tests the winner with tri_test for the
// exact t, u and v). The dots leave out the +0 w lane: x + 0 differs from x only for
// x = -0, so a dot, and every product of it, changes at most the sign of a zero, which no
// comparison below, nor coop_key, sees. The t > tmax rejection is left to the caller's
// t < best test: best starts at tmax and only decreases, so it implies t < tmax.
__device__ __forceinline__ float tri_test_t(f3 o, f3 d, float tmin, f3 v0, f3 e1, f3 e2) {
    auto dot3 = [](f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; };
    const f3 h = cross(d, e2);
    const float a = dot3(e1, h);
    if (a >= 0) return kInf;
    const float f = recip(a);
    const f3 s = sub(o, v0);
    const float u = f * dot3(s, h);
    if (u < 0.f || u > 1.f) return kInf;
    const f3 q = cross(s, e1);
    const float v = f * dot3(d, q);
    if (v < 0.f || u + v > 1.f) return kInf;
    const float t = f * dot3(e2, q);
    return t < tmin ? kInf : t;
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
    float ix, iy, iz, tmin;   // tmax lives on as the starting best t (bestT <= tmax always)
    float bestT, bu, bv;
    uint32_t best, ref, top, lo, nI, nT;
    bool fast;
    // a finished (or never started) ray: ref = kInvalid, which has kLeafBit set
    __device__ __forceinline__ bool alive() const { return ref != kInvalid; }
    __device__ __forceinline__ void kill() { ref = kInvalid; }
    uint2* spill;   // kSpillDepth private entries; kept outside the struct so that the
                    // fields above stay in registers

    // Ring slot of stack entry e. For the 12-entry ring, e / 12 = (e * 171) >> 11 exactly for
    // every e < 515 (entries never exceed the ring plus the spill array), computed with two
    // full-rate 24-bit multiplies instead of the 32-bit multiply-high of e % 12 (1% faster).
    __device__ __forceinline__ uint32_t slot(uint32_t e) const {
        static_assert(kSpillDepth + kLdsDepthMax < 515, "the ring-slot identity holds below 515");
        if (DEPTH == 12) return e - __umul24(__umul24(e, 171u) >> 11, 12u);
        return DEPTH == 16 || DEPTH == 8 ? (e & (DEPTH - 1)) : (e % DEPTH);
    }
    __device__ __forceinline__ void start(const TraceArgs& A, f3 o_, f3 d_, float tmin_, float tmax_) {
        o = o_; d = d_; tmin = tmin_;
        ix = recip(d.x); iy = recip(d.y); iz = recip(d.z);
        bestT = tmax_; bu = 0.f; bv = 0.f;
        best = kInvalid; top = 0; lo = 0; nI = 0; nT = 0;
        fast = A.finite && __builtin_isfinite(ix) && __builtin_isfinite(iy) && __builtin_isfinite(iz) &&
               ix != 0.f && iy != 0.f && iz != 0.f && __builtin_isfinite(o.x) && __builtin_isfinite(o.y) &&
               __builtin_isfinite(o.z) && !__builtin_isnan(tmin) && !__builtin_isnan(tmax_);
        ref = slab(o.x, o.y, o.z, ix, iy, iz, tmin, A.rb[0], A.rb[1], A.rb[2], A.rb[3], A.rb[4], A.rb[5], bestT) < bestT
                  ? A.root_ref : kInvalid;
    }
    __device__ __forceinline__ void push(uint2 (*lst)[kTraceThreads], uint32_t tid, uint32_t r, float t) {
        if (top - lo == (uint32_t)DEPTH) { spill[lo] = lst[slot(lo)][tid]; ++lo; }
        lst[slot(top)][tid] = make_uint2(r, __float_as_uint(t));
        ++top;
    }
    // pop the next entry that can still hold a closer hit
    __device__ __forceinline__ void pop(uint2 (*lst)[kTraceThreads], uint32_t tid) {
        ref = kInvalid;
        while (top > 0) {
            --top;
            uint2 e;
            if (top < lo) {
                --lo;
                e = spill[top];
            } else {
                // an atomic read: the compiler cannot merge it with the private read above
                // into one flat load through a selected pointer
                const uint64_t x = wave_ld(reinterpret_cast<uint64_t*>(&lst[slot(top)][tid]));
                e = make_uint2((uint32_t)x, (uint32_t)(x >> 32));
            }
            if (__uint_as_float(e.y) < bestT) { ref = e.x; break; }
        }
    }
    // one internal node: test both children, push the far one, continue with the near one
    template <bool COUNT>
    __device__ __forceinline__ void node_step(const TraceArgs& A, uint2 (*lst)[kTraceThreads], uint32_t tid,
                                              bool wfast) {
        if (COUNT) ++nI;
        const GNode* g = A.nodes + ref;
        const float4 l0 = ld_g(&g->l0), l1 = ld_g(&g->l1), r1b = ld_g(&g->r1);
        const uint2 refs = ld_g(reinterpret_cast<const uint2*>(&g->lref));
        float tl, tr;
        if (wfast) {   // wave vote (__all(fast), uniform): a scalar branch, both forms give the same values
            tl = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, bestT);
            tr = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, l1.z, l1.w, r1b.x, r1b.y, r1b.z, r1b.w, bestT);
        } else {
            // Opaque copies of the origin: the compiler cannot share this branch's products
            // with the fast branch, so the fast branch computes them in its own block, where
            // they are known canonical and its min/max need no canonicalizing moves.
            float ox = o.x, oy = o.y, oz = o.z;
            __asm__ volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));
            tl = slab(ox, oy, oz, ix, iy, iz, tmin, l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, bestT);
            tr = slab(ox, oy, oz, ix, iy, iz, tmin, l1.z, l1.w, r1b.x, r1b.y, r1b.z, r1b.w, bestT);
        }
        const bool lf = tl <= tr;
        const uint32_t nearRef = lf ? refs.x : refs.y, farRef = lf ? refs.y : refs.x;
        const float tn = lf ? tl : tr, tf = lf ? tr : tl;
        const bool go = tn < bestT, dopush = go && tf < bestT;
        if (dopush) {
            if (top - lo == (uint32_t)DEPTH) { spill[lo] = lst[slot(lo)][tid]; ++lo; }
            lst[slot(top)][tid] = make_uint2(farRef, __float_as_uint(tf));
        }
        top += dopush ? 1u : 0u;
        ref = go ? nearRef : ref;
        if (!go) pop(lst, tid);
    }
    // One wide node: test the (up to) four child boxes, continue with the nearest child
    // the ray enters before its best t, and push the others farthest first, each with its
    // exact entry distance (a popped entry is skipped once that distance is >= best t).
    // Children are ordered by 32-bit keys: the order-preserving image of the entry distance
    // with the slot index in its two low bits, so equal distances keep the binary tree's
    // left-to-right order (and distances within 4 ulps of each other go by slot); the order
    // only decides which subtree is searched first, never which boxes are searched.
    template <bool COUNT>
    __device__ __forceinline__ void wide_step(const TraceArgs& A, uint2 (*lst)[kTraceThreads], uint32_t tid) {
        if (COUNT) ++nI;
        const WNode* g = A.wnodes + ref;
        const float4 lx = ld_g(&g->lx), ly = ld_g(&g->ly), lz = ld_g(&g->lz);
        const float4 hx = ld_g(&g->hx), hy = ld_g(&g->hy), hz = ld_g(&g->hz);
        const uint4 rf = ld_g(&g->ref);
        float t0, t1, t2, t3;
        if (__all(fast)) {   // wave vote: a scalar branch, both forms give the same values
            t0 = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, bestT);
            t1 = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, bestT);
            t2 = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, bestT);
            t3 = slab_fast(o.x, o.y, o.z, ix, iy, iz, tmin, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, bestT);
        } else {
            float ox = o.x, oy = o.y, oz = o.z;
            __asm__ volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));
            t0 = slab(ox, oy, oz, ix, iy, iz, tmin, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, bestT);
            t1 = slab(ox, oy, oz, ix, iy, iz, tmin, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, bestT);
            t2 = slab(ox, oy, oz, ix, iy, iz, tmin, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, bestT);
            t3 = slab(ox, oy, oz, ix, iy, iz, tmin, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, bestT);
        }
        // a child is entered when its box is hit before best t; empty slots never are
        uint32_t k0 = (t0 < bestT && rf.x != kInvalid) ? ((ord_t(t0) & ~3u) | 0u) : ~0u;
        uint32_t k1 = (t1 < bestT && rf.y != kInvalid) ? ((ord_t(t1) & ~3u) | 1u) : ~0u;
        uint32_t k2 = (t2 < bestT && rf.z != kInvalid) ? ((ord_t(t2) & ~3u) | 2u) : ~0u;
        uint32_t k3 = (t3 < bestT && rf.w != kInvalid) ? ((ord_t(t3) & ~3u) | 3u) : ~0u;
        const uint32_t m = (k0 != ~0u) + (k1 != ~0u) + (k2 != ~0u) + (k3 != ~0u);
        if (m == 0) {
            pop(lst, tid);
            return;
        }
        // Sorting network on 4 keys; each key carries its child's entry distance and ref, so
        // no child is looked up by a lane-varying slot afterwards (the compiler would index
        // a private array in scratch memory for that).
        uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
        auto ce = [](uint32_t& ka, float& ta, uint32_t& ra, uint32_t& kb, float& tb, uint32_t& rb) {
            const bool sw = kb < ka;
            const uint32_t k = sw ? kb : ka, r = sw ? rb : ra;
            const float t = sw ? tb : ta;
            kb = sw ? ka : kb; rb = sw ? ra : rb; tb = sw ? ta : tb;
            ka = k; ra = r; ta = t;
        };
        ce(k0, t0, r0, k1, t1, r1);
        ce(k2, t2, r2, k3, t3, r3);
        ce(k0, t0, r0, k2, t2, r2);
        ce(k1, t1, r1, k3, t3, r3);
        ce(k1, t1, r1, k2, t2, r2);
        if (m > 3) push(lst, tid, r3, t3);
        if (m > 2) push(lst, tid, r2, t2);
        if (m > 1) push(lst, tid, r1, t1);
        ref = r0;
    }
    // Node phase of the while-while loop: node steps for the lanes at internal nodes until
    // none is left, or until at most `wait` are left while some lane already waits at a
    // leaf (those lanes then continue in a later phase). Every vote is wave-uniform; each
    // lane's own sequence of visits is unchanged.
    // `wait` must be the same on every lane (a kernel argument or a uniform field).
    template <bool COUNT, int TREE = kTreeBinary>
    __device__ __forceinline__ void node_phase(const TraceArgs& A, uint2 (*lst)[kTraceThreads], uint32_t tid,
                                               uint32_t wait) {
        const bool wfast = __all(fast);   // every lane's fast flag is fixed while it traces a ray
        for (;;) {
            const bool need = !(ref & kLeafBit);
            const uint64_t nm = __ballot(need);
            if (nm == 0) break;
            if ((uint32_t)__popcll(nm) <= wait && __ballot(alive() && (ref & kLeafBit)) != 0) break;
            if (need) {
                if (TREE != kTreeBinary) wide_step<COUNT>(A, lst, tid);
                else node_step<COUNT>(A, lst, tid, wfast);
            }
        }
    }
    // All triangles of the current leaf, in order; true when an any-hit query is answered.
    // Triangles are loaded TB at a time before any of them is tested, so a leaf costs
    // ceil(count / TB) memory round trips instead of count. Loads past the
    // leaf's end re-read its last triangle (always in bounds) and are not tested.
    template <bool ANY, bool COUNT>
    __device__ __forceinline__ bool leaf_step(const TraceArgs& A) {
        const uint2 lf = leaf_range(A.leaves, ref);   // begin, count
        const uint32_t end = lf.x + lf.y;
        for (uint32_t k = lf.x; k < end; k += TB) {
            float4 ta[TB], tb[TB];
            float tc[TB];
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                const uint32_t kk = min(k + (uint32_t)j, end - 1u);
                ld_tri(A.tris + kk, ta[j], tb[j], tc[j]);
            }
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                if (k + (uint32_t)j < end) {
                    if (COUNT) ++nT;
                    float u, v;
                    const float4 a = ta[j], b = tb[j];
                    // t < bestT below implies t <= tmax (bestT starts at tmax and only decreases)
                    const float t = tri_test(o, d, tmin, kInf, f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y},
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
    // Cooperative leaf phase (closest hit). The (ray, triangle) pairs of every lane waiting
    // at a leaf are dealt over the whole wave, 64 per round: lane r's pairs take positions
    // [start_r, start_r + count_r) of the wave's exclusive scan of leaf sizes, so each round
    // tests 64 pairs whatever the leaf sizes, and neighbouring lanes load neighbouring
    // triangles. A pair lane finds its owner from the round's segment starts (one LDS mark
    // per owner and a ballot), takes the owner's ray by shuffle and tests its triangle; an
    // LDS atomic min of coop_key(t, k) per owner keeps, among t < the owner's best t, the
    // smallest t and then the earliest triangle: what testing the leaf in order with a
    // strict t < best update keeps. The owner keeps that triangle and its t as its best; u
    // and v come from finish(), once per ray. Any-hit (ANY) keeps the smallest accepted triangle
    // index instead: the first triangle in leaf order that the in-order scan would stop at.
    // Returns true on the owner lanes whose any-hit query is answered. The caller keeps all
    // 64 lanes active and the control flow uniform. MIXED: lanes may trace different trees
    // (the batched renderer), so the triangle array is taken from the owner too.
    template <bool ANY, bool COUNT, bool MIXED>
    __device__ __forceinline__ bool leaf_coop(const TraceArgs& A, CoopLds* C) {
        const uint32_t lane = __lane_id();
        const bool at = alive() && (ref & kLeafBit);
        uint32_t c = 0, kb = 0;
        if (at) {
            const uint2 lf = leaf_range(A.leaves, ref);
            c = lf.y;
            kb = lf.x;
        }
        const uint32_t inc = wave_incl_scan(c);
        const uint32_t total = __builtin_amdgcn_readlane(inc, 63);   // uniform: a scalar loop bound
        const uint32_t start = inc - c;
        const uint32_t kbase = kb - start;   // triangle of position p is kbase + p (mod 2^32)
        if (COUNT && !ANY) nT += c;
        if (c) wave_st(&C->mins[lane], ~0ull);
        const uint64_t upto = (2ull << lane) - 1ull;   // lanes 0..lane
        for (uint32_t W = 0; W < total; W += 64) {
            wave_st(&C->mark[lane], kNoLane);
            __builtin_amdgcn_wave_barrier();
            // segments do not overlap, so at most one owner marks each position; the
            // segment holding position W marks position 0
            if (c && start < W + 64u && start + c > W) wave_st(&C->mark[(start > W ? start : W) - W], (uint8_t)lane);
            __builtin_amdgcn_wave_barrier();
            const uint32_t m = wave_ld(&C->mark[lane]);
            const uint64_t starts = __ballot(m != kNoLane);
            const int src = 63 - __clzll((long long)(starts & upto));
            const uint32_t r = __shfl(m, src);
            const float pox = __shfl(o.x, r), poy = __shfl(o.y, r), poz = __shfl(o.z, r);
            const float pdx = __shfl(d.x, r), pdy = __shfl(d.y, r), pdz = __shfl(d.z, r);
            const float ptmin = __shfl(tmin, r), pbest = __shfl(bestT, r);
            const uint32_t k = __shfl(kbase, r) + W + lane;
            const GTri* tris = A.tris;
            if (MIXED) {
                const uint64_t tp = reinterpret_cast<uint64_t>(A.tris);
                const uint32_t lo32 = __shfl((uint32_t)tp, r), hi32 = __shfl((uint32_t)(tp >> 32), r);
                tris = reinterpret_cast<const GTri*>(((uint64_t)hi32 << 32) | lo32);
            }
            if (W + lane < total) {
                float4 a, b;
                float cz;
                load_tri(tris + k, a, b, cz);
                const float t = tri_test_t(f3{pox, poy, poz}, f3{pdx, pdy, pdz}, ptmin, f3{a.x, a.y, a.z},
                                           f3{a.w, b.x, b.y}, f3{b.z, b.w, cz});
                if (t < pbest) atomicMin(&C->mins[r], ANY ? (unsigned long long)k : coop_key(t, k));
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (ANY) {
            bool hit = false;
            if (c) {
                const unsigned long long mk = wave_ld(&C->mins[lane]);
                hit = mk != ~0ull;
                if (hit) best = (uint32_t)mk;
                if (COUNT) nT += hit ? (uint32_t)mk - kb + 1u : c;   // tests up to the first accepted
            }
            return hit;
        }
        if (c) {
            // The winner and its t (up to the sign of a zero, which no later comparison
            // sees); finish() tests the final winner again for the exact t, u and v.
            const unsigned long long mk = wave_ld(&C->mins[lane]);
            if (mk != ~0ull) {
                best = (uint32_t)mk;
                bestT = t_of_ord((uint32_t)(mk >> 32));
            }
        }
        return false;
    }
    // Exact t, u and v of the closest hit. The cooperative leaf phases keep only each
    // winner's triangle and its t, so the ray's final winner is tested once more here with
    // tri_test: the inputs of the test that made it the winner, so the reference's bits (a
    // winner of leaf_step is tested again to the same bits). It passed t < best <= tmax then.
    // A is the lane's own tree.
    __device__ __forceinline__ void finish(const TraceArgs& A) {
        if (best == kInvalid) return;
        float4 a, b;
        float cz;
        load_tri(A.tris + best, a, b, cz);
        float u, v;
        bestT = tri_test(o, d, tmin, kInf, f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, f3{b.z, b.w, cz}, u, v);
        bu = u;
        bv = v;
    }
};

// One ray against one tree. LOOP 0 visits nodes and leaves in one loop; LOOP 1
// ("while-while") lets a lane that reaches a leaf wait until the other lanes of the wave
// are at leaves too (all but A.wait of them), so triangle tests run with the wave mostly
// converged. LOOP 2 is LOOP 1 with cooperative leaf phases (leaf_coop). All visit the
// same nodes in the same order per ray and return the same hit.
template <bool ANY, bool COUNT, int DEPTH, int LOOP, int TB = kTraceTriBatch, bool MIXED = false,
          int TREE = kTreeBinary>
__device__ __forceinline__ TraceOut trace_ray(const TraceArgs& A, f3 o, f3 d, float tmin, float tmax,
                                              uint2 (*lst)[kTraceThreads], uint32_t tid, uint32_t wait,
                                              CoopLds* coop) {
    Walker<DEPTH, TB> w;
    uint2 spill[kSpillDepth];
    w.spill = spill;
    w.start(A, o, d, tmin, tmax);
    if (LOOP == 0) {
        while (w.alive()) {
            if (w.ref & kLeafBit) {
                if (w.template leaf_step<ANY, COUNT>(A)) break;
                w.pop(lst, tid);
            } else if (TREE != kTreeBinary) {
                w.template wide_step<COUNT>(A, lst, tid);
            } else {
                w.template node_step<COUNT>(A, lst, tid, __all(w.fast));
            }
        }
    } else {
        for (;;) {
            w.template node_phase<COUNT, TREE>(A, lst, tid, wait);
            if (!__any(w.alive())) break;
            if (LOOP == kLoopCoop && __builtin_amdgcn_read_exec() == ~0ull) {
                const bool at = w.alive() && (w.ref & kLeafBit);
                if (w.template leaf_coop<ANY, COUNT, MIXED>(A, coop)) w.kill();
                else if (at) w.pop(lst, tid);
            } else if (w.alive() && (w.ref & kLeafBit)) {
                if (w.template leaf_step<ANY, COUNT>(A)) w.kill();
                else w.pop(lst, tid);
            }
        }
    }
    if (!ANY) w.finish(A);
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
    uint32_t i0 = ld_g(A.tris3 + 3 * tri), i1 = ld_g(A.tris3 + 3 * tri + 1), i2 = ld_g(A.tris3 + 3 * tri + 2);
    f3 n0{ld_g(V + 6 * i0 + 3), ld_g(V + 6 * i0 + 4), ld_g(V + 6 * i0 + 5)};
    f3 n1{ld_g(V + 6 * i1 + 3), ld_g(V + 6 * i1 + 4), ld_g(V + 6 * i1 + 5)};
    f3 n2{ld_g(V + 6 * i2 + 3), ld_g(V + 6 * i2 + 4), ld_g(V + 6 * i2 + 5)};
    float w = 1.f - u - v;
    return normalize(add(add(scl(w, n0), scl(u, n1)), scl(v, n2)));
}

}  // namespace rtx
