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
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "rtx_sort.hpp"
#include "rtx_trace_impl.hpp"

namespace rtx {
namespace {

template <bool ANY, bool COUNT, int DEPTH, int LOOP, int TREE>
__global__ void __launch_bounds__(kTraceThreads) k_trace(TraceArgs A, const Ray32* __restrict__ rays, uint32_t n,
                                                         Hit16* __restrict__ hits, float* __restrict__ normals,
                                                         uint8_t* __restrict__ occl, uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ perm) {
    __shared__ uint2 lst[DEPTH][kTraceThreads];
    __shared__ CoopLds coop[LOOP == kLoopCoop ? kTraceThreads / 64 : 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t k = xcd_block(blockIdx.x) * kTraceThreads + tid;
    // out-of-range lanes still run (with a ray that misses) so wave-wide votes see them
    const bool in = k < n;
    const uint32_t i = in && perm ? perm[k] : k;   // work item k traces ray i
    float4 r0 = make_float4(0.f, 0.f, 0.f, 1.f), r1 = make_float4(0.f, 0.f, 1.f, 0.f);
    if (in) {
        r0 = reinterpret_cast<const float4*>(rays)[2 * i];
        r1 = reinterpret_cast<const float4*>(rays)[2 * i + 1];
    }
    const TraceOut h = trace_ray<ANY, COUNT, DEPTH, LOOP, kTraceTriBatch, false, TREE>(
        A, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, r0.w, r1.w, lst, tid, A.wait,
        coop + (LOOP == kLoopCoop ? tid / 64 : 0));
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
// A.refill lanes of a wave are idle, one atomic on `work` hands them the next rays, so a
// wave no longer waits for its slowest ray. Every wave-level operation (ballot, leader
// election, shuffle, the exit test) sits in control flow that is uniform across the wave:
// no lane leaves the loop body early with `continue`, because the compiler may then run
// those lanes' next iteration on their own, and a refill seen by only part of the wave
// would split `more` and could elect a leader that never issues the atomic.
// The threshold is a tuning field. Measured (Mrays/s): incoherent closest hit 12: 891, 14: 935,
// 16: 940, 18: 939, 20: 935, 24: 930, 32: 925; coherent closest hit 18: 1696, 24: 1932,
// 32: 1955; incoherent any hit 18: 1205, 24: 1356, 32: 1377.

// The wide kernels are held to 6 waves per SIMD (80 VGPRs): at their natural 89 VGPRs only 5
// fit, and 6 measured faster (1M triangles, 4M incoherent rays: 4.31 against 4.62 ms).
template <bool ANY, bool COUNT, int DEPTH, int LOOP, int TREE>
__global__ void __launch_bounds__(kTraceThreads, TREE == kTreeBinary ? 1 : 6) k_trace_persist(TraceArgs A, const Ray32* __restrict__ rays,
                                                                 uint32_t n, Hit16* __restrict__ hits,
                                                                 float* __restrict__ normals,
                                                                 uint8_t* __restrict__ occl,
                                                                 uint32_t* __restrict__ counts, uint32_t* work,
                                                                 const uint32_t* __restrict__ perm) {
    __shared__ uint2 lst[DEPTH][kTraceThreads];
    __shared__ CoopLds coop[LOOP == kLoopCoop ? kTraceThreads / 64 : 1];
    const uint32_t tid = threadIdx.x;
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    Walker<DEPTH> w;
    uint2 spill[kSpillDepth];
    w.spill = spill;
    w.kill();
    w.best = kInvalid;
    uint32_t id = kInvalid;
    bool more = true;   // the same on every lane: set only from wave-wide values
    for (;;) {
        const uint64_t idle = __ballot(!w.alive());
        const uint32_t nidle = (uint32_t)__popcll(idle);
        const bool refill = more && nidle >= A.refill;   // wave-uniform
        const bool done = !more && nidle == 64u;        // wave-uniform: nothing left to trace
        // Finished rays are written when their lanes take new rays or the wave ends (they
        // idle until then in any case), so finish()'s test runs for many lanes at once.
        if ((refill || done) && !w.alive() && id != kInvalid) {
            if (ANY) {
                occl[id] = w.best != kInvalid ? 1 : 0;
            } else {
                w.finish(A);
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
        if (done) break;
        if (refill) {
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t base = 0;
            if (__lane_id() == leader) base = atomicAdd(work, nidle);
            base = __shfl(base, leader);
            more = base + nidle < n;
            if (!w.alive()) {
                const uint32_t k = base + (uint32_t)__popcll(idle & lt);
                if (k < n) {
                    id = perm ? perm[k] : k;
                    const float4 r0 = reinterpret_cast<const float4*>(rays)[2 * id];
                    const float4 r1 = reinterpret_cast<const float4*>(rays)[2 * id + 1];
                    w.start(A, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, r0.w, r1.w);
                }
            }
        }
        // A started ray that missed the root is written at the next refill or at the end.
        // Once the queue is empty, the leaf-phase wait scales with the lanes still tracing: a
        // wave's last rays no longer stop each other's node phases at every leaf one of them
        // reaches (4M incoherent rays: about 1% faster).
        const uint32_t wait = more ? A.wait : (A.wait * (uint32_t)__popcll(__ballot(w.alive()))) >> 6;
        w.template node_phase<COUNT, TREE>(A, lst, tid, wait);
        if (LOOP == kLoopCoop) {   // all lanes active: the loop is wave-uniform
            const bool at = w.alive() && (w.ref & kLeafBit);
            if (w.template leaf_coop<ANY, COUNT, false>(A, coop + tid / 64)) w.kill();
            else if (at) w.pop(lst, tid);
        } else if (w.alive() && (w.ref & kLeafBit)) {
            if (w.template leaf_step<ANY, COUNT>(A)) w.kill();
            else w.pop(lst, tid);
        }
    }
}

// One wave per ray, for calls of few rays (the per-ray bvh::trace shim, small batches),
// whose time is the ray's chain of dependent loads: the reference's depth-first walk
// (bvh_traverse.cpp:91-178) visits about 78 nodes one after another. Here the wave walks
// the tree breadth-first, a level per step, each lane testing the two children of one node
// of the level, and the leaves met at a level are tested together, their (ray, triangle)
// pairs dealt over the lanes. The chain shrinks to the tree's depth.
// The hit is still the reference's. Each node carries a path key: one bit per level, set
// when the node is its parent's far child (near = the smaller entry distance, the left one
// on a tie, as the reference orders them), so keys order leaves as the reference visits
// them. The reference keeps, among the triangles with the smallest t, the first it tests:
// the one with the smallest (key, triangle index). A subtree is dropped once its entry
// distance exceeds the best t, or equals it with a key after the best's: it can then hold no
// such triangle. (Where the two walks cull a child differently, that child's entry distance
// exceeds the smallest t, so no candidate depends on how its parent's children were ordered.)
// A level of more nodes or leaves than the LDS lists hold (or a tree deeper than the key's
// 64 bits) sends the ray through the usual one-lane walk instead.
constexpr uint32_t kWaveList = 512;       // nodes and leaves per level
constexpr uint32_t kWaveMaxDepth = 60;    // tree depth the 64-bit path keys cover
// Calls of at most this many rays take the wave kernel: on the 1M-triangle tree it is ahead
// of the persistent kernel up to about 25k rays (16,384 rays: 386 against 557 us per call;
// 32,768: 714 against 589).
constexpr size_t kWaveRayMax = 16384;
struct WaveLds {
    uint32_t ref[2][kWaveList];
    float tn[2][kWaveList];
    unsigned long long key[2][kWaveList];
    uint32_t lref[kWaveList];
    unsigned long long lkey[kWaveList];
    uint8_t mark[64];
};
// A candidate hit ordered as the reference ranks equal-t hits: t (ord_t image), path key,
// triangle index.
struct WaveCand {
    uint32_t ot;
    unsigned long long key;
    uint32_t k;
};
__device__ __forceinline__ bool cand_less(const WaveCand& a, const WaveCand& b) {
    return a.ot != b.ot ? a.ot < b.ot : a.key != b.key ? a.key < b.key : a.k < b.k;
}
__device__ __forceinline__ unsigned long long shfl64(unsigned long long x, int src) {
    const uint32_t lo = __shfl((uint32_t)x, src), hi = __shfl((uint32_t)(x >> 32), src);
    return ((unsigned long long)hi << 32) | lo;
}
// Smallest value over the wave: a DPP min scan (row shifts, then row broadcasts), read from
// the last lane.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#define RTX_DPP_MIN(ctrl, rmask) \
    x = min(x, (uint32_t)__builtin_amdgcn_update_dpp(~0u, x, ctrl, rmask, 0xf, false))
    RTX_DPP_MIN(0x111, 0xf);   // row_shr:1
    RTX_DPP_MIN(0x112, 0xf);   // row_shr:2
    RTX_DPP_MIN(0x114, 0xf);   // row_shr:4
    RTX_DPP_MIN(0x118, 0xf);   // row_shr:8
    RTX_DPP_MIN(0x142, 0xa);   // row_bcast:15 into rows 1, 3
    RTX_DPP_MIN(0x143, 0xc);   // row_bcast:31 into rows 2, 3
#undef RTX_DPP_MIN
    return __builtin_amdgcn_readlane(x, 63);
}
__device__ __forceinline__ WaveCand wave_min_cand(WaveCand c) {
    for (int o = 1; o < 64; o <<= 1) {
        const int src = (int)(__lane_id() ^ (uint32_t)o);
        WaveCand x;
        x.ot = __shfl(c.ot, src);
        x.key = shfl64(c.key, src);
        x.k = __shfl(c.k, src);
        if (cand_less(x, c)) c = x;
    }
    return c;
}

template <bool ANY>
__global__ void __launch_bounds__(64) k_trace_wave(TraceArgs A, const Ray32* __restrict__ rays,
                                                   Hit16* __restrict__ hits, float* __restrict__ normals,
                                                   uint8_t* __restrict__ occl, uint32_t max_depth) {
    // the lists, and in the same bytes the one-lane walk's stack when that is needed (it uses
    // column 0 of the 12-entry ring once the lists are abandoned)
    constexpr size_t kStackBytes = sizeof(uint2) * 12 * kTraceThreads;
    __shared__ __align__(16) unsigned char lds[sizeof(WaveLds) > kStackBytes ? sizeof(WaveLds) : kStackBytes];
    WaveLds& L = *reinterpret_cast<WaveLds*>(lds);
    auto lst = reinterpret_cast<uint2(*)[kTraceThreads]>(lds);
    const uint32_t lane = __lane_id(), r = blockIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull, upto = (2ull << lane) - 1ull;
    const float4 r0 = reinterpret_cast<const float4*>(rays)[2 * r];
    const float4 r1 = reinterpret_cast<const float4*>(rays)[2 * r + 1];
    const float ox = r0.x, oy = r0.y, oz = r0.z, tmin = r0.w, tmax = r1.w;
    const float ix = recip(r1.x), iy = recip(r1.y), iz = recip(r1.z);
    const f3 o{ox, oy, oz}, d{r1.x, r1.y, r1.z};
    // the best hit so far (wave-uniform); bestT is its t, or tmax before any hit
    bool found = false;
    WaveCand best{~0u, ~0ull, ~0u};
    float bestT = tmax;
    // entry distance tn and path key of a subtree that may still hold the reference's hit
    auto keep = [&](float tn, unsigned long long key) {
        return tn < bestT || (found && tn == bestT && key < best.key);
    };
    bool overflow = max_depth > kWaveMaxDepth;
    uint32_t ncur = 0, nleaf = 0, cur = 0, depth = 0;
    const float troot = slab(ox, oy, oz, ix, iy, iz, tmin, A.rb[0], A.rb[1], A.rb[2], A.rb[3], A.rb[4], A.rb[5], tmax);
    if (!overflow && troot < tmax) {
        if (lane == 0) {
            if (A.root_ref & kLeafBit) { L.lref[0] = A.root_ref; L.lkey[0] = 0ull; }
            else { L.ref[0][0] = A.root_ref; L.tn[0][0] = troot; L.key[0][0] = 0ull; }
        }
        if (A.root_ref & kLeafBit) nleaf = 1;
        else ncur = 1;
    }
    __builtin_amdgcn_wave_barrier();
    // A node of the level, read from the list, and its 64-byte record. The first 64 nodes of a
    // level are read before the leaves met at the previous level are tested, so their loads
    // are in flight together with the triangle loads.
    struct NodeIn {
        uint32_t ref;
        float tn;
        unsigned long long key;
        float4 l0, l1, rb;
        uint2 refs;
    };
    auto read_node = [&](uint32_t i) {
        NodeIn x{0u, kInf, 0ull, make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f),
                 make_float4(0.f, 0.f, 0.f, 0.f), make_uint2(kInvalid, kInvalid)};
        if (i < ncur) {
            x.ref = L.ref[cur][i];
            x.tn = L.tn[cur][i];
            x.key = L.key[cur][i];
            const GNode* g = A.nodes + x.ref;
            x.l0 = ld_g(&g->l0);
            x.l1 = ld_g(&g->l1);
            x.rb = ld_g(&g->r1);
            x.refs = ld_g(reinterpret_cast<const uint2*>(&g->lref));
        }
        return x;
    };
    for (;;) {
        const NodeIn first = read_node(lane);
        // the leaves met at this level: every (leaf, triangle) pair, 64 per round
        for (uint32_t lb = 0; lb < nleaf; lb += 64) {
            const uint32_t j = lb + lane;
            uint32_t c = 0, kb = 0;
            unsigned long long lk = 0ull;
            if (j < nleaf) {
                const uint2 lr = leaf_range(A.leaves, L.lref[j]);
                kb = lr.x;
                c = lr.y;
                lk = L.lkey[j];
            }
            const uint32_t inc = wave_incl_scan(c);
            const uint32_t total = __builtin_amdgcn_readlane(inc, 63);
            const uint32_t start = inc - c;
            WaveCand mine{~0u, ~0ull, ~0u};
            for (uint32_t W = 0; W < total; W += 64) {
                wave_st(&L.mark[lane], kNoLane);
                __builtin_amdgcn_wave_barrier();
                if (c && start < W + 64u && start + c > W) wave_st(&L.mark[(start > W ? start : W) - W], (uint8_t)lane);
                __builtin_amdgcn_wave_barrier();
                const uint32_t m = wave_ld(&L.mark[lane]);
                const uint64_t starts = __ballot(m != kNoLane);
                const int src = 63 - __clzll((long long)(starts & upto));
                const uint32_t ow = __shfl(m, src);
                const uint32_t okb = __shfl(kb, ow), ostart = __shfl(start, ow);
                const unsigned long long okey = shfl64(lk, ow);
                if (W + lane < total) {
                    const uint32_t k = okb + (W + lane - ostart);
                    float4 a, b;
                    float cz;
                    load_tri(A.tris + k, a, b, cz);
                    const float t = tri_test_t(o, d, tmin, f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, f3{b.z, b.w, cz});
                    const WaveCand x{ord_t(t), okey, k};
                    // only a hit that can still become (or tie) the best
                    if (t < tmax && (t < bestT || (found && t == bestT)) && cand_less(x, mine)) mine = x;
                }
                __builtin_amdgcn_wave_barrier();
            }
            // the wave's smallest t; usually one lane holds it, else the full (t, key, index) order
            const uint32_t mo = wave_min_u32(mine.ot);
            if (mo != ~0u) {
                const uint64_t tied = __ballot(mine.ot == mo);
                WaveCand w;
                if (__popcll(tied) == 1) {
                    const int h = __builtin_ctzll(tied);
                    w.ot = mo;
                    w.key = ((unsigned long long)__builtin_amdgcn_readlane((uint32_t)(mine.key >> 32), h) << 32) |
                            __builtin_amdgcn_readlane((uint32_t)mine.key, h);
                    w.k = __builtin_amdgcn_readlane(mine.k, h);
                } else {
                    w = wave_min_cand(mine);
                }
                if (!found || cand_less(w, best)) {
                    best = w;
                    found = true;
                    bestT = t_of_ord(w.ot);
                }
            }
        }
        if ((ANY && found) || ncur == 0 || overflow) break;
        // the nodes of this level: both children of each, kept ones into the next level's lists
        uint32_t nnext = 0;
        nleaf = 0;
        const unsigned long long bit = 1ull << (63u - depth);
        for (uint32_t base = 0; base < ncur; base += 64) {
            const uint32_t i = base + lane;
            const NodeIn x = base == 0 ? first : read_node(i);
            const bool act = i < ncur && keep(x.tn, x.key);
            const unsigned long long key = x.key;
            float tl = kInf, tr = kInf;
            const uint2 refs = x.refs;
            if (act) {
                tl = slab(ox, oy, oz, ix, iy, iz, tmin, x.l0.x, x.l0.y, x.l0.z, x.l0.w, x.l1.x, x.l1.y, bestT);
                tr = slab(ox, oy, oz, ix, iy, iz, tmin, x.l1.z, x.l1.w, x.rb.x, x.rb.y, x.rb.z, x.rb.w, bestT);
            }
            const bool lf = tl <= tr;
            const uint32_t nref = lf ? refs.x : refs.y, fref = lf ? refs.y : refs.x;
            const float ntn = lf ? tl : tr, ftn = lf ? tr : tl;
            const unsigned long long nkey = key, fkey = key | bit;
            const bool en = act && keep(ntn, nkey), ef = act && keep(ftn, fkey);
            const bool in_ = en && !(nref & kLeafBit), if_ = ef && !(fref & kLeafBit);
            const bool ln = en && (nref & kLeafBit), lf_ = ef && (fref & kLeafBit);
            const uint64_t bin = __ballot(in_), bif = __ballot(if_), bln = __ballot(ln), blf = __ballot(lf_);
            const uint32_t cin = (uint32_t)__popcll(bin), cif = (uint32_t)__popcll(bif);
            const uint32_t cln = (uint32_t)__popcll(bln), clf = (uint32_t)__popcll(blf);
            if (nnext + cin + cif > kWaveList || nleaf + cln + clf > kWaveList) {
                overflow = true;
                break;
            }
            const uint32_t nx = cur ^ 1u;
            if (in_) {
                const uint32_t q = nnext + (uint32_t)__popcll(bin & lt);
                L.ref[nx][q] = nref; L.tn[nx][q] = ntn; L.key[nx][q] = nkey;
            }
            if (if_) {
                const uint32_t q = nnext + cin + (uint32_t)__popcll(bif & lt);
                L.ref[nx][q] = fref; L.tn[nx][q] = ftn; L.key[nx][q] = fkey;
            }
            if (ln) {
                const uint32_t q = nleaf + (uint32_t)__popcll(bln & lt);
                L.lref[q] = nref; L.lkey[q] = nkey;
            }
            if (lf_) {
                const uint32_t q = nleaf + cln + (uint32_t)__popcll(blf & lt);
                L.lref[q] = fref; L.lkey[q] = fkey;
            }
            nnext += cin + cif;
            nleaf += cln + clf;
        }
        __builtin_amdgcn_wave_barrier();
        if (overflow) break;
        cur ^= 1u;
        ncur = nnext;
        ++depth;
    }
    if (overflow) {
        // the usual walk on one lane
        if (lane == 0) {
            const TraceOut h = trace_ray<ANY, false, 12, kLoopSingle, kTraceTriBatch, false, kTreeBinary>(
                A, o, d, tmin, tmax, lst, 0u, 0u, nullptr);
            if (ANY) {
                occl[r] = h.tri != kInvalid ? 1 : 0;
            } else {
                reinterpret_cast<float4*>(hits)[r] = make_float4(h.t, h.u, h.v, __uint_as_float(h.tri));
                if (normals) {
                    f3 nn{0.f, 0.f, 0.f};
                    if (h.tri != kInvalid) nn = hit_normal(A, h.tri, h.u, h.v);
                    normals[3 * r] = nn.x; normals[3 * r + 1] = nn.y; normals[3 * r + 2] = nn.z;
                }
            }
        }
        return;
    }
    if (lane != 0) return;
    if (ANY) {
        occl[r] = found ? 1 : 0;
        return;
    }
    // exact t, u and v of the winner: the reference's test on the reference's inputs
    float t = 0.f, u = 0.f, v = 0.f;
    if (found) {
        float4 a, b;
        float cz;
        load_tri(A.tris + best.k, a, b, cz);
        t = tri_test(o, d, tmin, kInf, f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, f3{b.z, b.w, cz}, u, v);
    }
    reinterpret_cast<float4*>(hits)[r] = make_float4(t, u, v, __uint_as_float(found ? best.k : kInvalid));
    if (normals) {
        f3 nn{0.f, 0.f, 0.f};
        if (found) nn = hit_normal(A, best.k, u, v);
        normals[3 * r] = nn.x; normals[3 * r + 1] = nn.y; normals[3 * r + 2] = nn.z;
    }
}

__global__ void k_zero_work(uint32_t* w) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *w = 0u;
}

// Ray order key: 30-bit Morton code of the origin inside the scene box (10 bits per axis).
// Rays with nearby keys share nodes and triangles in the caches and follow similar paths
// within a wave: the persistent kernel deals rays in key order from one counter, so at any
// time all XCDs work in one compact window of space. (Direction bits below the origin code, or a
// 48-bit origin+direction key, measured slower: the sort costs more and L2 reuse comes
// from the origins.)
__device__ __forceinline__ uint32_t spread10(uint32_t x) {   // bit b of x -> bit 3b
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ uint32_t quant(float x, float levels) {   // x in [0,1] -> [0, levels-1]
    const float f = x * levels;
    return f > 0.f ? (f < levels - 1.f ? (uint32_t)f : (uint32_t)(levels - 1.f)) : 0u;
}
__global__ void k_ray_keys(TraceArgs A, const Ray32* __restrict__ rays, uint32_t n, uint32_t* keys, uint32_t* vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = reinterpret_cast<const float4*>(rays)[2 * i];
    const float sx = A.rb[3] - A.rb[0], sy = A.rb[4] - A.rb[1], sz = A.rb[5] - A.rb[2];
    const float ux = sx > 0.f ? (r0.x - A.rb[0]) / sx : 0.f;
    const float uy = sy > 0.f ? (r0.y - A.rb[1]) / sy : 0.f;
    const float uz = sz > 0.f ? (r0.z - A.rb[2]) / sz : 0.f;
    const uint32_t key = (spread10(quant(ux, 1024.f)) << 2) | (spread10(quant(uy, 1024.f)) << 1) |
                         spread10(quant(uz, 1024.f));
    keys[i] = key;
    vals[i] = i;
}


constexpr uint32_t kSortMinRays = 1u << 16;
constexpr size_t kPinnedMaxRays = 4096;   // rtx_trace_closest_host calls up to this size use pinned buffers
constexpr uint32_t kPersistMinRays = 1u << 12;

// Event pairs of rtx_profile_begin/end.
struct Profile {
    std::mutex mu;
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};
Profile g_prof;

// Records the start event of a traversal kernel when profiling; returns the pair index.
int prof_start(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    if (!g_prof.on) return -1;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess) return -1;
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return -1; }
    (void)hipEventRecord(a, s);
    g_prof.ev.emplace_back(a, b);
    return (int)g_prof.ev.size() - 1;
}
void prof_stop(int k, hipStream_t s) {
    if (k < 0) return;
    std::lock_guard<std::mutex> lk(g_prof.mu);
    (void)hipEventRecord(g_prof.ev[k].second, s);
}

// Grow-only scratch of the ray sort, kept on the handle.
bool reserve_sort(rtx_bvh* b, uint32_t n) {
    if (b->rs_cap >= n) return true;
    void* ps[] = {b->rs_keys[0], b->rs_keys[1], b->rs_vals[0], b->rs_vals[1], b->rs_hist, b->rs_part};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    b->rs_keys[0] = dalloc<uint32_t>(n);
    b->rs_keys[1] = dalloc<uint32_t>(n);
    b->rs_vals[0] = dalloc<uint32_t>(n);
    b->rs_vals[1] = dalloc<uint32_t>(n);
    b->rs_hist = dalloc<uint32_t>(sort_hist_entries(n));
    b->rs_part = dalloc<uint32_t>(sort_part_entries(n));
    const bool ok = b->rs_keys[0] && b->rs_keys[1] && b->rs_vals[0] && b->rs_vals[1] && b->rs_hist && b->rs_part;
    b->rs_cap = ok ? n : 0;
    return ok;
}

int check(const rtx_bvh* b, const void* rays, size_t n, const void* out, const char* fn) {
    if (!b || ((!out || !rays) && n)) return fail(RTX_E_ARG, std::string("null argument on ") + fn);
    if (!b->gnodes || !b->leaves || !b->gtris) return fail(RTX_E_STATE, std::string("BVH not loaded on ") + fn);
    if (n > (1ull << 31)) return fail(RTX_E_LIMIT, std::string("more than 2^31 rays on ") + fn);
    if (b->max_depth > kMaxTraceDepth)
        return fail(RTX_E_LIMIT, std::string("tree deeper than the traversal stack on ") + fn);
    return RTX_OK;
}

// While-while loop shape of the trace kernels: cooperative leaf phases when tuned on.
uint32_t loop_shape() { return g_tuning.leaf_coop ? (uint32_t)kLoopCoop : (uint32_t)kLoopWhileWhile; }

// Device scratch of one trace call: the persistent kernel's ray counter and the ray
// sort's ping-pong buffers. Either the handle's (grow-only, ordered across calls with an
// event) or caller-provided memory that needs no allocation (rtx_trace_closest_noalloc).
struct TraceScratch {
    uint32_t* work;
    SortBuffers32 sort;
};

template <bool ANY, bool COUNT>
int launch(const rtx_bvh* b, const rtx_ray* rays, size_t n, rtx_hit* hits, float* normals, uint8_t* occl,
           uint32_t* counts, hipStream_t s, const TraceScratch* ext = nullptr) {
    if (n == 0) return RTX_OK;
    uint32_t blocks = (uint32_t)((n + kTraceThreads - 1) / kTraceThreads);
    const TraceArgs A = make_trace_args(b);
    const uint32_t* perm = nullptr;
    rtx_bvh* mb = const_cast<rtx_bvh*>(b);
    const bool sort = g_tuning.ray_sort && n >= kSortMinRays;
    // The handle's ray counter and sort buffers: a call that uses them waits (on its own
    // stream) for the previous such call, whatever stream that used.
    std::unique_lock<std::mutex> lk(mb->scratch_mu, std::defer_lock);
    // the persistent grid pays off only with rays for many waves; smaller counting calls take
    // the while-while kernel, which gives the same results
    const bool persist = g_tuning.trace_loop == 2 && n >= kPersistMinRays;
    // calls of a few rays: one wave per ray, breadth-first (not the counting launch, whose
    // per-ray visit counts are those of the depth-first walk)
    if (!COUNT && n <= kWaveRayMax && b->max_depth <= kWaveMaxDepth) {
        const int pk = ext ? -1 : prof_start(s);
        k_trace_wave<ANY><<<(uint32_t)n, 64, 0, s>>>(A, reinterpret_cast<const Ray32*>(rays),
                                                    reinterpret_cast<Hit16*>(hits), normals, occl, b->max_depth);
        prof_stop(pk, s);
        RTX_HIP(hipGetLastError());
        return RTX_OK;
    }
    const bool scratch = (sort || persist) && !ext;
    if (scratch) {
        lk.lock();
        if (!mb->scratch_ev) RTX_HIP(hipEventCreateWithFlags(&mb->scratch_ev, hipEventDisableTiming));
        if (mb->scratch_used) RTX_HIP(hipStreamWaitEvent(s, mb->scratch_ev, 0));
    }
    auto release = [&]() -> int {
        if (!scratch) return RTX_OK;
        RTX_HIP(hipEventRecord(mb->scratch_ev, s));
        mb->scratch_used = true;
        return RTX_OK;
    };
    uint32_t* work = ext ? ext->work : b->work;
    if (sort) {
        SortBuffers32 S;
        if (ext) {
            S = ext->sort;
        } else {
            if (!reserve_sort(mb, (uint32_t)n)) return fail(RTX_E_HIP, "device allocation failed for ray sorting");
            S = SortBuffers32{{mb->rs_keys[0], mb->rs_keys[1]}, {mb->rs_vals[0], mb->rs_vals[1]}, mb->rs_hist,
                              mb->rs_part, 0};
        }
        k_ray_keys<<<blocks, kTraceThreads, 0, s>>>(A, reinterpret_cast<const Ray32*>(rays), (uint32_t)n, S.keys[0],
                                                    S.vals[0]);
        // stable sort on the top 24 key bits (8 per axis): the lowest pass bought no
        // locality the traversal could measure, sorting on 16 bits lost some
        if (radix_sort_pairs(S, (uint32_t)n, 0xEu, s)) return fail(RTX_E_HIP, "ray sort failed");
        perm = S.vals[S.cur];
    }
    // the wide layout exists whenever the root is internal (a leaf root has one leaf to test);
    // trees too deep for its stack bound take the binary walk
    const int tree = b->wnodes && b->max_depth <= kMaxWideTraceDepth ? (int)g_tuning.trace_tree : kTreeBinary;
    if (persist) {
        const uint32_t pb = std::min<uint32_t>(blocks, g_tuning.trace_grid_blocks);
        k_zero_work<<<1, 64, 0, s>>>(work);
        const int pk = ext ? -1 : prof_start(s);
        with_trace_variant(g_tuning.trace_lds_depth, loop_shape(), [&](auto D, auto L) {
            constexpr int DD = decltype(D)::value, LL = decltype(L)::value;
            if (tree == kTreeWide)
                k_trace_persist<ANY, COUNT, DD, LL, kTreeWide><<<pb, kTraceThreads, 0, s>>>(
                    A, reinterpret_cast<const Ray32*>(rays), (uint32_t)n, reinterpret_cast<Hit16*>(hits), normals,
                    occl, counts, work, perm);
            else
                k_trace_persist<ANY, COUNT, DD, LL, kTreeBinary><<<pb, kTraceThreads, 0, s>>>(
                    A, reinterpret_cast<const Ray32*>(rays), (uint32_t)n, reinterpret_cast<Hit16*>(hits), normals,
                    occl, counts, work, perm);
        });
        prof_stop(pk, s);
        RTX_HIP(hipGetLastError());
        return release();
    }
    const int pk = ext ? -1 : prof_start(s);
    with_trace_variant(g_tuning.trace_lds_depth, g_tuning.trace_loop == 0 ? (uint32_t)kLoopSingle : loop_shape(),
                       [&](auto D, auto L) {
        constexpr int DD = decltype(D)::value, LL = decltype(L)::value;
        if (tree == kTreeWide)
            k_trace<ANY, COUNT, DD, LL, kTreeWide><<<blocks, kTraceThreads, 0, s>>>(
                A, reinterpret_cast<const Ray32*>(rays), (uint32_t)n, reinterpret_cast<Hit16*>(hits), normals, occl,
                counts, perm);
        else
            k_trace<ANY, COUNT, DD, LL, kTreeBinary><<<blocks, kTraceThreads, 0, s>>>(
                A, reinterpret_cast<const Ray32*>(rays), (uint32_t)n, reinterpret_cast<Hit16*>(hits), normals, occl,
                counts, perm);
    });
    prof_stop(pk, s);
    RTX_HIP(hipGetLastError());
    return release();
}

}  // namespace

// batch_paths 2^26: the wavefront's later bounces hold few paths, and with 2^23-path batches
// they left most of the GPU idle (the 132.7M-path terrain render: 116 ms at 2^23, 99 ms at
// 2^24, 87 ms at 2^25, 83.5 ms at 2^26; the path state is 100 bytes per path).
rtx_tuning g_tuning = {1u << 26, 8192, 1024, 12, 2, 1, 24, 2048, 8, 2, 16, 1, 1, 18, 0};

}  // namespace rtx

using namespace rtx;

extern "C" int rtx_profile_begin(void) {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    for (auto& e : g_prof.ev) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    g_prof.ev.clear();
    g_prof.on = true;
    return RTX_OK;
}

extern "C" int rtx_profile_end(double* kernel_ms, uint32_t* launches) {
    if (!kernel_ms || !launches) return fail(RTX_E_ARG, "null argument on rtx_profile_end()");
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.on = false;
    double total = 0.0;
    int rc = RTX_OK;
    for (auto& e : g_prof.ev) {
        float ms = 0.f;
        if (rc == RTX_OK && !(hip_ok(hipEventSynchronize(e.second), "hipEventSynchronize") &&
                              hip_ok(hipEventElapsedTime(&ms, e.first, e.second), "hipEventElapsedTime")))
            rc = RTX_E_HIP;
        total += ms;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    *kernel_ms = total;
    *launches = (uint32_t)g_prof.ev.size();
    g_prof.ev.clear();
    return rc;
}

extern "C" int rtx_trace_closest(const rtx_bvh* b, const rtx_ray* rays, size_t n, rtx_hit* hits, float* normals,
                                 void* stream) {
    int rc = check(b, rays, n, hits, "rtx_trace_closest()");
    if (rc) return rc;
    return launch<false, false>(b, rays, n, hits, normals, nullptr, nullptr, (hipStream_t)stream);
}

namespace rtx {
namespace {
// Scratch of rtx_trace_closest_noalloc: a file-scope device array, so the entry point needs
// no allocation. Each GPU the library runs on holds its own copy until the process ends.
// Layout in uint32 words: [0, 64) the persistent kernel's ray counter, then the sort's
// keys[0], keys[1], vals[0], vals[1] (kNoAllocMaxRays each), its per-tile digit counts and
// its digit totals.
constexpr uint32_t kNoAllocMaxRays = 1u << 21;
constexpr size_t kNoAllocHist = sort_hist_entries(kNoAllocMaxRays);
constexpr size_t kNoAllocPart = sort_part_entries(kNoAllocMaxRays);
constexpr size_t kNoAllocWords = 64 + 4ull * kNoAllocMaxRays + kNoAllocHist + kNoAllocPart;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

// Device address of `scratch` per device, resolved once per device with that device
// current. In host code `scratch` names a host-side stand-in: only these addresses may be
// passed to kernels or copies.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
uint32_t* g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(int device, uint32_t** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        RTX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_noalloc_ptr[device];
    return RTX_OK;
}
}  // namespace
}  // namespace rtx

extern "C" int rtx_trace_closest_noalloc(const rtx_bvh* b, const rtx_ray* rays, size_t n, rtx_hit* hits,
                                         float* normals, void* stream) {
    int rc = check(b, rays, n, hits, "rtx_trace_closest_noalloc()");
    if (rc) return rc;
    if (n > kNoAllocMaxRays)
        return fail(RTX_E_LIMIT, "more than 2097152 rays on rtx_trace_closest_noalloc() (its fixed scratch size)");
    if (n == 0) return RTX_OK;
    int dev = -1;
    RTX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(RTX_E_LIMIT, "device index out of range on rtx_trace_closest_noalloc()");
    if (dev != b->device)
        return fail(RTX_E_STATE, "rtx_trace_closest_noalloc() called with another device current than the BVH's");
    uint32_t* base = nullptr;
    rc = noalloc_scratch(dev, &base);
    if (rc) return rc;
    const size_t m = kNoAllocMaxRays;
    TraceScratch S;
    S.work = base;
    S.sort = SortBuffers32{{base + 64, base + 64 + m}, {base + 64 + 2 * m, base + 64 + 3 * m}, base + 64 + 4 * m,
                           base + 64 + 4 * m + kNoAllocHist, 0};
    if (sort_hist_entries((uint32_t)n) > kNoAllocHist || sort_part_entries((uint32_t)n) > kNoAllocPart)
        return fail(RTX_E_LIMIT, "ray sort scratch exceeds its fixed size on rtx_trace_closest_noalloc()");
    return launch<false, false>(b, rays, n, hits, normals, nullptr, nullptr, (hipStream_t)stream, &S);
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
    rtx_bvh* mb = const_cast<rtx_bvh*>(b);
    std::lock_guard<std::mutex> lk(mb->host_mu);
    if (!mb->host_stream) RTX_HIP(hipStreamCreateWithFlags(&mb->host_stream, hipStreamNonBlocking));
    hipStream_t st = mb->host_stream;
    // Small calls (the per-ray bvh::trace shim): the kernel reads the rays from and writes
    // the hits to pinned host memory kept on the handle, on the handle's own stream, so a
    // call is a host memcpy, one launch and one stream synchronization, with no DMA copies.
    if (n <= kPinnedMaxRays) {
        if (!mb->pin_rays) {   // allocated, and their device addresses taken, once per handle
            const size_t cap = kPinnedMaxRays;
            void *r = nullptr, *h = nullptr, *nr = nullptr, *dr = nullptr, *dh = nullptr, *dn = nullptr;
            const bool ok = hipHostMalloc(&r, sizeof(rtx_ray) * cap, hipHostMallocDefault) == hipSuccess &&
                            hipHostMalloc(&h, sizeof(rtx_hit) * cap, hipHostMallocDefault) == hipSuccess &&
                            hipHostMalloc(&nr, 12 * cap, hipHostMallocDefault) == hipSuccess &&
                            hipHostGetDevicePointer(&dr, r, 0) == hipSuccess &&
                            hipHostGetDevicePointer(&dh, h, 0) == hipSuccess &&
                            hipHostGetDevicePointer(&dn, nr, 0) == hipSuccess;
            if (!ok) {
                for (void* p : {r, h, nr})
                    if (p) (void)hipHostFree(p);
                (void)hipGetLastError();
                return fail(RTX_E_HIP, "pinned host allocation failed on rtx_trace_closest_host()");
            }
            mb->pin_rays = r; mb->pin_hits = h; mb->pin_nrm = static_cast<float*>(nr);
            mb->pin_rays_d = dr; mb->pin_hits_d = dh; mb->pin_nrm_d = dn;
            mb->pin_cap = cap;
        }
        std::memcpy(mb->pin_rays, rays, sizeof(rtx_ray) * n);
        rc = launch<false, false>(b, static_cast<const rtx_ray*>(mb->pin_rays_d), n,
                                  static_cast<rtx_hit*>(mb->pin_hits_d),
                                  normals ? static_cast<float*>(mb->pin_nrm_d) : nullptr, nullptr, nullptr, st);
        if (rc) return rc;
        RTX_HIP(hipStreamSynchronize(st));
        std::memcpy(hits, mb->pin_hits, sizeof(rtx_hit) * n);
        if (normals) std::memcpy(normals, mb->pin_nrm, 12 * n);
        return RTX_OK;
    }
    // Larger calls: device staging on the handle (grow-only) and copies on the stream.
    if (mb->host_cap < n) {
        (void)hipFree(mb->host_rays);
        (void)hipFree(mb->host_hits);
        (void)hipFree(mb->host_nrm);
        const size_t cap = std::max<size_t>(n, 1024);
        mb->host_rays = dalloc<rtx_ray>(cap);
        mb->host_hits = dalloc<rtx_hit>(cap);
        mb->host_nrm = dalloc<float>(3 * cap);
        mb->host_cap = (mb->host_rays && mb->host_hits && mb->host_nrm) ? cap : 0;
        if (!mb->host_cap) return fail(RTX_E_HIP, "device allocation failed on rtx_trace_closest_host()");
    }
    rtx_ray* dr = static_cast<rtx_ray*>(mb->host_rays);
    rtx_hit* dh = static_cast<rtx_hit*>(mb->host_hits);
    float* dn = normals ? mb->host_nrm : nullptr;
    RTX_HIP(hipMemcpyAsync(dr, rays, sizeof(rtx_ray) * n, hipMemcpyHostToDevice, st));
    rc = launch<false, false>(b, dr, n, dh, dn, nullptr, nullptr, st);
    if (rc) return rc;
    RTX_HIP(hipMemcpyAsync(hits, dh, sizeof(rtx_hit) * n, hipMemcpyDeviceToHost, st));
    if (normals) RTX_HIP(hipMemcpyAsync(normals, dn, 12 * n, hipMemcpyDeviceToHost, st));
    RTX_HIP(hipStreamSynchronize(st));
    return RTX_OK;
}

namespace rtx {
namespace {
// Compares recip(x) with the IEEE quotient 1.0f / x for the float bit patterns
// [first, first + count); mismatches are counted per wave (NaN inputs match any NaN).
__global__ void __launch_bounds__(256) k_recip_check(uint64_t first, uint64_t count, unsigned long long* out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0, lowest = ~0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint32_t bits = (uint32_t)(first + i);
        const float x = __uint_as_float(bits);
        const float q = 1.0f / x, r = recip(x);
        const bool both_nan = q != q && r != r;
        if (!both_nan && __float_as_uint(q) != __float_as_uint(r)) {
            ++bad;
            if (bits < lowest) lowest = bits;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        bad += __shfl_xor(bad, o);
        const unsigned long long w = __shfl_xor(lowest, o);
        lowest = w < lowest ? w : lowest;
    }
    if (__lane_id() == 0 && bad) {
        atomicAdd(&out[0], bad);
        atomicMin(&out[1], lowest);
    }
}
}  // namespace
}  // namespace rtx

extern "C" int rtx_debug_recip_check(uint64_t first, uint64_t count, uint64_t* mismatches, uint32_t* first_bad) {
    if (!mismatches || !first_bad || first > 0xFFFFFFFFull || count > (1ull << 32) - first)
        return fail(RTX_E_ARG, "bad argument on rtx_debug_recip_check()");
    unsigned long long* d = dalloc<unsigned long long>(2);
    if (!d) return fail(RTX_E_HIP, "device allocation failed on rtx_debug_recip_check()");
    const unsigned long long init[2] = {0ull, ~0ull};
    int rc = RTX_OK;
    unsigned long long h[2] = {0ull, ~0ull};
    if (!hip_ok(hipMemcpy(d, init, sizeof(init), hipMemcpyHostToDevice), "hipMemcpy")) rc = RTX_E_HIP;
    if (rc == RTX_OK && count) {
        k_recip_check<<<8192, 256>>>(first, count, d);
        if (!hip_ok(hipGetLastError(), "k_recip_check") ||
            !hip_ok(hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost), "hipMemcpy"))
            rc = RTX_E_HIP;
    }
    (void)hipFree(d);
    if (rc != RTX_OK) return rc;
    *mismatches = h[0];
    *first_bad = h[1] == ~0ull ? 0xFFFFFFFFu : (uint32_t)h[1];
    return RTX_OK;
}
