// GPU BVH builder: bit-exact with the reference's bakeBVH pipeline.
//
//   sortTrisByMorton  (src/bvh/bvh_morton.cpp:40-159)   -> k_tri_prep, k_scene_box_*, k_morton,
//                                                         stable 8-bit LSD radix (k_rs_*)
//   initRoot/bfsLoop  (src/bvh/bvh_bfs.cpp:50-134)      -> level-synchronous frontier:
//   trySAHSplit       (src/bvh/bvh_sah.cpp:24-460)         k_level_decide / k_level_budget / k_level_apply
//   reorderNodes/Tris (src/bvh/bvh_bake.cpp:75-149)     -> k_subtree_*, k_dfs_*, k_emit
//
// The reference processes one FIFO queue. A FIFO visits the tree level by level, left to
// right, so one launch per level over the level's contiguous node range reproduces the
// order in which nodes are numbered and in which the spatial-split budget (1.5 N triangle
// references) is consumed; k_level_budget resolves that budget in queue order.
// Node boxes are bit-identical: min/max extension keeps the last of equal extrema, which
// makes an ordered reduction (earlier, later) equal to the serial fold the reference runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <unordered_map>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "rtx_clip.hpp"
#include "rtx_internal.hpp"
#include "rtx_sort.hpp"

namespace rtx {

namespace {

constexpr uint32_t F_LEAF = 1u, F_SAH = 2u, F_PENDING = 4u;

struct BNode {
    f3 lo, hi;
    uint32_t a, b;       // children, or {first tri-index position, count}
    uint32_t flags;
    uint32_t pad;
};

enum : uint32_t { D_LEAF = 0, D_MORTON = 1, D_OBJECT = 2, D_SPATIAL = 3 };

struct Dec {
    uint32_t type;
    uint32_t mid;           // split position (Morton / object)
    uint32_t nl, ns, nr;    // spatial: left-only, straddling, right-only counts
    uint32_t fb_split;      // spatial: the object fallback would split
    int oAxis, oBin;
    float oCMin, oScale;
    float plane;
    int sAxis;
    int nb;
};

struct Counters {
    uint32_t total;      // s_totalNodeCount
    uint32_t tsize;      // s_triIndexSize
    uint32_t ssize;      // s_sortedTrisSize
    uint32_t err;
    uint32_t nwave;      // SAH nodes of this level queued for the wave evaluator
    uint32_t lvl_ls;     // the level being built: its first node and node count
    uint32_t lvl_cnt;
    uint32_t lvl_le;     // one past its last node (the next level's first)
    uint32_t fast;       // this level's budget was resolved by the multi-block pass
};

struct Ctx {
    const float* verts;
    const uint32_t* tris;
    const float4* tbox;      // 2 per input triangle (lo, hi)
    const float4* cent;      // centroid per input triangle
    const uint64_t* scodes;  // Morton codes in sorted order
    uint32_t* sorted;        // s_sortedTris (capacity cap)
    uint32_t* tindex;        // s_triIndex (capacity cap)
    BNode* nodes;
    Dec* dec;
    uint32_t* child;         // child base id per frontier item (kInvalid = none)
    uint32_t* toff;          // appended s_triIndex offset (accepted spatial)
    uint32_t* soff;          // appended s_sortedTris offset (accepted spatial)
    uint32_t* acc;           // spatial accepted flag
    uint32_t cap;
    uint32_t ncap;           // node array capacity
    Counters* ctr;
    uint32_t* wlist;         // frontier items for k_sah_wave
};

__device__ __forceinline__ f3 vpos(const float* v, uint32_t i) { return {v[6 * i], v[6 * i + 1], v[6 * i + 2]}; }

__device__ __forceinline__ Box load_tbox(const float4* tb, uint32_t t) {
    float4 a = tb[2 * t], b = tb[2 * t + 1];
    return Box{{a.x, a.y, a.z}, {b.x, b.y, b.z}};
}

__device__ __forceinline__ Box combine(const Box& early, const Box& late) {
    Box r = early;
    r.extend(late);
    return r;
}

// Ordered fold of triangle boxes over s_triIndex positions [b, e) (computeRangeAABB,
// bvh_bfs.cpp:14-28).
__device__ Box range_box(const Ctx& C, uint32_t b, uint32_t e) {
    Box r = Box::empty();
    for (uint32_t i = b; i < e; ++i) r.extend(load_tbox(C.tbox, C.sorted[C.tindex[i]]));
    return r;
}

// Sutherland-Hodgman clip of a triangle to a box, then its AABB (AABB.cpp:86-171), with
// the reference's vertex order and float operations. The polygon (at most 9 vertices
// after 6 planes) lives in a PolyStore:
//  - PolyRegs: small per-lane arrays, in private memory once indexed by lane-varying
//    positions (the serial SAH path for nodes over 64 triangles, and the fallback clip of
//    k_sah_wave: there 15 KB of LDS per wave held its occupancy to 6 waves per CU);
//  - PolyLds: a per-lane column of LDS (k_level_apply's child clipping). Output positions
//    differ between lanes, and indexing registers with a lane-varying index costs a loop
//    over the distinct indices in the wave, while LDS addresses are per lane.
constexpr int kPolyMax = 10;

struct PolyRegs {
    float v[2][3][kPolyMax];
    __device__ __forceinline__ float& at(int buf, int c, int i) { return v[buf][c][i]; }
};

template <int COLS>
struct PolyLdsT {
    float* base;   // poly_floats(COLS) per block, lane-minor: [buf][coord][vertex][COLS]
    __device__ __forceinline__ float& at(int buf, int c, int i) { return base[((buf * 3 + c) * kPolyMax + i) * COLS]; }
};
using PolyLds = PolyLdsT<64>;
constexpr int poly_floats(int cols) { return 2 * 3 * kPolyMax * cols; }
constexpr int kPolyLdsFloats = poly_floats(64);
// Clip pairs per chunk of k_sah_wave's spatial pass: one round of 64 lanes. (128-pair
// chunks took 1.5 KB more LDS per wave: 14 waves per CU instead of 17, and slower builds.)
constexpr int kClipChunk = 64;

template <class PS>
__device__ Box clip_tri_in(f3 v0, f3 v1, f3 v2, const Box& r, PS& P) {
    P.at(0, 0, 0) = v0.x; P.at(0, 1, 0) = v0.y; P.at(0, 2, 0) = v0.z;
    P.at(0, 0, 1) = v1.x; P.at(0, 1, 1) = v1.y; P.at(0, 2, 1) = v1.z;
    P.at(0, 0, 2) = v2.x; P.at(0, 1, 2) = v2.y; P.at(0, 2, 2) = v2.z;
    int n = 3, cur = 0;
    for (int axis = 0; axis < 3; ++axis) {
        for (int side = 0; side < 2; ++side) {
            const float plane = side == 0 ? comp(r.lo, axis) : comp(r.hi, axis);
            const int nx = cur ^ 1;
            int m = 0;
            for (int i = 0; i < n; ++i) {
                const int j = (i + 1 == n) ? 0 : i + 1;
                const float pi = P.at(cur, axis, i), pj = P.at(cur, axis, j);
                const bool ci = side == 0 ? (pi >= plane) : (pi <= plane);
                const bool cj = side == 0 ? (pj >= plane) : (pj <= plane);
                const float xi = P.at(cur, 0, i), yi = P.at(cur, 1, i), zi = P.at(cur, 2, i);
                if (ci) { P.at(nx, 0, m) = xi; P.at(nx, 1, m) = yi; P.at(nx, 2, m) = zi; ++m; }
                if (ci != cj) {
                    const float t = (plane - pi) / (pj - pi);
                    const float xj = P.at(cur, 0, j), yj = P.at(cur, 1, j), zj = P.at(cur, 2, j);
                    P.at(nx, 0, m) = xi + t * (xj - xi);
                    P.at(nx, 1, m) = yi + t * (yj - yi);
                    P.at(nx, 2, m) = zi + t * (zj - zi);
                    ++m;
                }
            }
            cur = nx;
            n = m;
            if (n == 0) return Box::empty();
        }
    }
    Box b = Box::empty();
    for (int i = 0; i < n; ++i) b.extend(f3{P.at(cur, 0, i), P.at(cur, 1, i), P.at(cur, 2, i)});
    return b;
}

__device__ Box clip_tri(f3 v0, f3 v1, f3 v2, const Box& r) {
    PolyRegs P;
    return clip_tri_in(v0, v1, v2, r, P);
}

__device__ __forceinline__ void tri_verts(const Ctx& C, uint32_t t, f3& a, f3& b, f3& c) {
    a = vpos(C.verts, C.tris[3 * t]);
    b = vpos(C.verts, C.tris[3 * t + 1]);
    c = vpos(C.verts, C.tris[3 * t + 2]);
}

// libstdc++ bidirectional std::partition (stl_algo.h:1487-1510) over s_triIndex.
template <class P>
__device__ uint32_t partition_idx(uint32_t* a, uint32_t first, uint32_t last, P pred) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (pred(a[first])) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!pred(a[last])) --last; else break;
        }
        uint32_t t = a[first]; a[first] = a[last]; a[last] = t;
        ++first;
    }
}

__device__ __forceinline__ float cent_axis(const Ctx& C, uint32_t si, int axis) {
    float4 c = C.cent[C.sorted[si]];
    return axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
}

__device__ __forceinline__ bool obj_pred(const Ctx& C, uint32_t si, int axis, int bin, float cmin, float scale, int nb) {
    return min(f2i_x86((cent_axis(C, si, axis) - cmin) * scale), nb - 1) < bin;
}

// Binned SAH evaluation of one node (bvh_sah.cpp:24-243), serial in one lane.
__device__ void sah_eval(const Ctx& C, uint32_t begin, uint32_t n, const Box& pbox, Dec& d) {
    const uint32_t end = begin + n;
    const int nb = clampi((int)n / 4, kBinsMin, kBinsMax);
    const float invA = 1.0f / pbox.half_area();
    float best = kCIsect * (float)n;
    int bAxis = -1, bBin = -1;
    float bCMin = 0.f, bScale = 0.f;
    Box bins[kBinsMax], lbox[kBinsMax];
    uint32_t cnt[kBinsMax], lcnt[kBinsMax], cnt2[kBinsMax];

    for (int axis = 0; axis < 3; ++axis) {
        float cmin = 3.40282347e+38f, cmax = -3.40282347e+38f;
        for (uint32_t i = begin; i < end; ++i) {
            float c = cent_axis(C, C.tindex[i], axis);
            cmin = std_min(cmin, c);
            cmax = std_max(cmax, c);
        }
        if (cmax - cmin < kCentroidEps) continue;
        float scale = (float)nb / (cmax - cmin);
        for (int b = 0; b < nb; ++b) { bins[b] = Box::empty(); cnt[b] = 0; }
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t t = C.sorted[C.tindex[i]];
            float4 c4 = C.cent[t];
            float c = axis == 0 ? c4.x : (axis == 1 ? c4.y : c4.z);
            int b = min(f2i_x86((c - cmin) * scale), nb - 1);
            cnt[b]++;
            bins[b].extend(load_tbox(C.tbox, t));
        }
        Box la = Box::empty();
        uint32_t lc = 0;
        for (int b = 0; b < nb - 1; ++b) { la.extend(bins[b]); lc += cnt[b]; lbox[b] = la; lcnt[b] = lc; }
        Box ra = Box::empty();
        uint32_t rc = 0;
        for (int b = nb - 1; b >= 1; --b) {
            ra.extend(bins[b]); rc += cnt[b];
            uint32_t lc2 = lcnt[b - 1];
            if (lc2 == 0 || rc == 0) continue;
            float cost = kCTrav + kCIsect * (lbox[b - 1].half_area() * (float)lc2 + ra.half_area() * (float)rc) * invA;
            if (cost < best) { best = cost; bAxis = axis; bBin = b; bCMin = cmin; bScale = scale; }
        }
    }
    d.oAxis = bAxis; d.oBin = bBin; d.oCMin = bCMin; d.oScale = bScale; d.nb = nb;

    bool spatial = false;
    float sMin = 0.f, sScale = 0.f;
    for (int axis = 0; axis < 3; ++axis) {
        float nmin = comp(pbox.lo, axis), nmax = comp(pbox.hi, axis);
        if (nmax - nmin < kSpatialEps) continue;
        float scale = (float)nb / (nmax - nmin);
        for (int b = 0; b < nb; ++b) { bins[b] = Box::empty(); cnt[b] = 0; cnt2[b] = 0; }
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t t = C.sorted[C.tindex[i]];
            Box tb = load_tbox(C.tbox, t);
            int b0 = clampi(f2i_x86((comp(tb.lo, axis) - nmin) * scale), 0, nb - 1);
            int b1 = clampi(f2i_x86((comp(tb.hi, axis) - nmin) * scale), 0, nb - 1);
            cnt[b0]++;
            cnt2[b1]++;
            if (b0 > b1) continue;
            f3 p0, p1, p2;
            tri_verts(C, t, p0, p1, p2);
            Box reg = pbox;
            for (int b = b0; b <= b1; ++b) {
                set_comp(reg.lo, axis, nmin + (float)b / scale);
                set_comp(reg.hi, axis, nmin + (float)(b + 1) / scale);
                Box cb = clip_tri(p0, p1, p2, reg);
                if (cb.valid()) bins[b].extend(cb);
            }
        }
        Box la = Box::empty();
        uint32_t lc = 0;
        for (int b = 0; b < nb - 1; ++b) { la.extend(bins[b]); lc += cnt[b]; lbox[b] = la; lcnt[b] = lc; }
        Box ra = Box::empty();
        uint32_t rc = 0;
        for (int b = nb - 1; b >= 1; --b) {
            ra.extend(bins[b]); rc += cnt2[b];
            uint32_t lc2 = lcnt[b - 1];
            if (lc2 == 0 || rc == 0) continue;
            float cost = kCTrav + kCIsect * (lbox[b - 1].half_area() * (float)lc2 + ra.half_area() * (float)rc) * invA;
            if (cost < best) { best = cost; bAxis = axis; bBin = b; spatial = true; sMin = nmin; sScale = scale; }
        }
    }

    if (bAxis == -1) { d.type = D_LEAF; return; }
    if (!spatial) {
        uint32_t mid = partition_idx(C.tindex, begin, end, [&](uint32_t si) {
            return obj_pred(C, si, bAxis, bBin, bCMin, bScale, nb);
        });
        if (mid == begin || mid == end) { d.type = D_LEAF; return; }
        d.type = D_OBJECT;
        d.mid = mid;
        return;
    }
    // Spatial winner: 3-way partition [left-only | straddling | right-only] (bvh_sah.cpp:296-333).
    const float plane = sMin + (float)bBin / sScale;
    const int ax = bAxis;
    uint32_t leftEnd = partition_idx(C.tindex, begin, end, [&](uint32_t si) {
        return comp(load_tbox(C.tbox, C.sorted[si]).hi, ax) <= plane;
    });
    uint32_t strEnd = partition_idx(C.tindex, leftEnd, end, [&](uint32_t si) {
        return comp(load_tbox(C.tbox, C.sorted[si]).lo, ax) < plane;
    });
    uint32_t nl = leftEnd - begin, ns = strEnd - leftEnd, nr = end - strEnd;
    d.plane = plane;
    d.sAxis = ax;
    d.nl = nl; d.ns = ns; d.nr = nr;
    if (nl + ns == 0 || ns + nr == 0) {
        // Degenerate spatial partition: object-split fallback (bvh_sah.cpp:335-388).
        if (d.oAxis == -1) { d.type = D_LEAF; return; }
        uint32_t mid = partition_idx(C.tindex, begin, end, [&](uint32_t si) {
            return obj_pred(C, si, d.oAxis, d.oBin, d.oCMin, d.oScale, nb);
        });
        if (mid == begin || mid == end) { d.type = D_LEAF; return; }
        d.type = D_OBJECT;
        d.mid = mid;
        return;
    }
    // Candidate; the budget decides in queue order. Precompute whether the fallback splits.
    uint32_t c = 0;
    if (d.oAxis != -1)
        for (uint32_t i = begin; i < end; ++i) c += obj_pred(C, C.tindex[i], d.oAxis, d.oBin, d.oCMin, d.oScale, nb) ? 1u : 0u;
    d.fb_split = (d.oAxis != -1 && c > 0 && c < n) ? 1u : 0u;
    d.type = D_SPATIAL;
}

// ---------------------------------------------------------------------------
// Kernels: triangle preparation, scene box, Morton codes
// ---------------------------------------------------------------------------
__global__ void k_check_indices(const uint32_t* tris, uint32_t n3, uint32_t nv, uint32_t* err) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3 && tris[i] >= nv) atomicOr(err, 1u);
}

__global__ void k_tri_prep(const float* verts, const uint32_t* tris, uint32_t nt, float4* tbox, float4* cent) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    f3 a = vpos(verts, tris[3 * t]), b = vpos(verts, tris[3 * t + 1]), c = vpos(verts, tris[3 * t + 2]);
    Box bx = Box::empty();
    bx.extend(a); bx.extend(b); bx.extend(c);
    tbox[2 * t] = make_float4(bx.lo.x, bx.lo.y, bx.lo.z, 0.f);
    tbox[2 * t + 1] = make_float4(bx.hi.x, bx.hi.y, bx.hi.z, 0.f);
    // Centroid exactly as computeMortonCodes / buildCentroidCache (bvh_morton.cpp:88-90, bvh_bake.cpp:51-55).
    cent[t] = make_float4(((a.x + b.x) + c.x) / 3.0f, ((a.y + b.y) + c.y) / 3.0f, ((a.z + b.z) + c.z) / 3.0f, 0.f);
}

constexpr int kRedThreads = 256;

// Ordered reduction of triangle boxes in input order (the scene AABB loop, bvh_morton.cpp:49-59).
__global__ void k_scene_box_partial(const float4* tbox, uint32_t nt, uint32_t chunk, float4* part) {
    __shared__ Box sh[kRedThreads];
    uint32_t blk_b = blockIdx.x * chunk * kRedThreads;
    uint32_t b = blk_b + threadIdx.x * chunk;
    Box acc = Box::empty();
    for (uint32_t i = b; i < b + chunk && i < nt; ++i) acc.extend(load_tbox(tbox, i));
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 1; s < kRedThreads; s <<= 1) {
        if ((threadIdx.x & (2 * s - 1)) == 0) sh[threadIdx.x] = combine(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = make_float4(sh[0].lo.x, sh[0].lo.y, sh[0].lo.z, 0.f);
        part[2 * blockIdx.x + 1] = make_float4(sh[0].hi.x, sh[0].hi.y, sh[0].hi.z, 0.f);
    }
}

__global__ void k_scene_box_final(const float4* part, uint32_t np, float* out6) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Box acc = Box::empty();
    for (uint32_t i = 0; i < np; ++i) acc.extend(load_tbox(part, i));
    out6[0] = acc.lo.x; out6[1] = acc.lo.y; out6[2] = acc.lo.z;
    out6[3] = acc.hi.x; out6[4] = acc.hi.y; out6[5] = acc.hi.z;
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint32_t quant(float c, float lo, float size) {
    uint32_t q = f2u_x86(((c - lo) * 2097151.0f) / size);
    return q > kMortonMax ? kMortonMax : q;
}

// Grid-stride over the triangles (launched with at most kMortonBlocks blocks). The OR and
// AND of all codes (to skip radix passes whose digit never varies) are reduced per block
// and taken with one atomic pair per block: an atomic pair per wave on the same two words
// serialized the kernel.
constexpr uint32_t kMortonBlocks = 1024;

__global__ void __launch_bounds__(256) k_morton(const float4* cent, uint32_t nt, const float* box6, uint64_t* codes,
                                                uint32_t* idx, unsigned long long* orand) {
    __shared__ unsigned long long so[4], sa[4];
    float sx = box6[3] - box6[0], sy = box6[4] - box6[1], sz = box6[5] - box6[2];
    if (sx == 0.f) sx = 1.f;
    if (sy == 0.f) sy = 1.f;
    if (sz == 0.f) sz = 1.f;
    unsigned long long o = 0ull, a = ~0ull;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        const float4 c = cent[t];
        const uint64_t code = (spread21(quant(c.x, box6[0], sx)) << 2) | (spread21(quant(c.y, box6[1], sy)) << 1) |
                              spread21(quant(c.z, box6[2], sz));
        codes[t] = code;
        idx[t] = t;
        o |= code;
        a &= code;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        o |= __shfl_xor(o, d);
        a &= __shfl_xor(a, d);
    }
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) { so[w] = o; sa[w] = a; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) { o |= so[i]; a &= sa[i]; }
        atomicOr(&orand[0], o);
        atomicAnd(&orand[1], a);
    }
}

// ---------------------------------------------------------------------------
// Level-synchronous top-down build (the FIFO of bvh_bfs.cpp:72-134)
// ---------------------------------------------------------------------------
__global__ void k_init_root(BNode* nodes, const float* box6, uint32_t nt, uint32_t* tindex, Counters* ctr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BNode r;
        r.lo = {box6[0], box6[1], box6[2]};
        r.hi = {box6[3], box6[4], box6[5]};
        r.a = 0;
        r.b = nt;
        r.flags = F_LEAF | (nt <= (uint32_t)kLbvhThreshold ? F_SAH : 0u);
        r.pad = 0;
        nodes[0] = r;
        ctr->total = 1;
        ctr->tsize = nt;
        ctr->ssize = nt;
        ctr->err = 0;
        ctr->nwave = 0;
        ctr->lvl_ls = 0;
        ctr->lvl_cnt = 0;
        ctr->lvl_le = 0;   // the first k_level_start makes the level [0, 1): the root
    }
    (void)tindex;
}

// Starts the next level on the device: its nodes are those the previous level appended,
// [previous end, total). The level kernels read this range, so the host can queue a level
// before it knows the range (its grids are sized by an upper bound and surplus blocks exit).
__global__ void k_level_start(Counters* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t le = c->total;
        c->lvl_ls = c->lvl_le;
        c->lvl_cnt = le - c->lvl_le;
        c->lvl_le = le;
        c->nwave = 0;
        c->fast = 0;
    }
}

__global__ void k_iota(uint32_t* a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void k_level_decide(Ctx C) {
    const uint32_t ls = C.ctr->lvl_ls, cnt = C.ctr->lvl_cnt;
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    const uint32_t ni = ls + k;
    BNode nd = C.nodes[ni];
    Dec d;
    d.type = D_LEAF;
    d.mid = 0; d.nl = d.ns = d.nr = 0; d.fb_split = 0;
    d.oAxis = -1; d.oBin = -1; d.oCMin = 0.f; d.oScale = 0.f; d.plane = 0.f; d.sAxis = -1; d.nb = 0;
    const uint32_t begin = nd.a, n = nd.b;
    if (n > kMaxLeaf) {
        bool sah = (nd.flags & F_SAH) != 0;
        if (!sah) {
            uint64_t first = C.scodes[begin], last = C.scodes[begin + n - 1];
            uint64_t x = first ^ last;
            if (x == 0) {
                sah = true;  // identical codes: SAH fallback (bvh_bfs.cpp:102-107)
            } else {
                uint64_t mask = 1ull << (63 - __clzll(x));
                uint32_t lo = begin, hi = begin + n - 1;   // code[lo] lacks the bit, code[hi] has it
                while (hi - lo > 1) {
                    uint32_t m = lo + (hi - lo) / 2;
                    if (C.scodes[m] & mask) hi = m; else lo = m;
                }
                d.type = D_MORTON;
                d.mid = hi;
            }
        }
        if (sah) {
            if (n <= 64) {
                C.wlist[atomicAdd(&C.ctr->nwave, 1u)] = k;   // one wave per node (k_sah_wave)
                return;
            }
            sah_eval(C, begin, n, Box{nd.lo, nd.hi}, d);
        }
    }
    C.dec[k] = d;
}

// Binned SAH of one node of at most 64 triangles per 64-lane block (bvh_sah.cpp:24-243).
// The triangles are staged in LDS in s_triIndex order. Lane (axis, bin) folds, in that
// order, the triangles falling into its bin -- object bins by centroid, spatial bins by
// clipping (AABB.cpp:86-171) -- so every bin box is the reference's serial fold. Each
// lane then prices split candidate `bin` of its axis; the winner is the minimum cost, and
// on equal cost the first one the reference's loops meet (lower axis, higher bin).
__global__ void __launch_bounds__(64) k_sah_wave(Ctx C) {
    __shared__ float tv[64][9];
    __shared__ float tb[64][6];
    __shared__ float tc[64][3];
    __shared__ uint32_t tsi[64];
    __shared__ float bbox[3 * 16][6];
    __shared__ uint32_t bc0[3 * 16], bc1[3 * 16];
    __shared__ uint32_t tsi_slot[64];   // rank exchange of the partitions
    __shared__ uint8_t obin[64][3];              // centroid bin per triangle and axis
    __shared__ uint8_t sb0[64][3], sb1[64][3];   // spatial entry / exit bin per triangle and axis
    __shared__ uint32_t tax[64][3];              // first clip pair of triangle i on axis a
    __shared__ float pbuf[kClipChunk][6];        // clipped boxes of one chunk of pairs
    __shared__ uint8_t powner[kClipChunk];       // triangle of each pair of the chunk
    if (blockIdx.x >= C.ctr->nwave) return;
    const uint32_t ls = C.ctr->lvl_ls;
    const uint32_t k = C.wlist[blockIdx.x];
    const uint32_t ni = ls + k;
    const BNode nd = C.nodes[ni];
    const uint32_t begin = nd.a, n = nd.b;
    const int lane = threadIdx.x;
    const int nb = clampi((int)n / 4, kBinsMin, kBinsMax);   // n <= 64 -> nb <= 16
    const Box pbox{nd.lo, nd.hi};
    if ((uint32_t)lane < n) {
        const uint32_t si = C.tindex[begin + lane];
        const uint32_t t = C.sorted[si];
        tsi[lane] = si;
        f3 a, b, c;
        tri_verts(C, t, a, b, c);
        tv[lane][0] = a.x; tv[lane][1] = a.y; tv[lane][2] = a.z;
        tv[lane][3] = b.x; tv[lane][4] = b.y; tv[lane][5] = b.z;
        tv[lane][6] = c.x; tv[lane][7] = c.y; tv[lane][8] = c.z;
        const Box bx = load_tbox(C.tbox, t);
        tb[lane][0] = bx.lo.x; tb[lane][1] = bx.lo.y; tb[lane][2] = bx.lo.z;
        tb[lane][3] = bx.hi.x; tb[lane][4] = bx.hi.y; tb[lane][5] = bx.hi.z;
        const float4 ce = C.cent[t];
        tc[lane][0] = ce.x; tc[lane][1] = ce.y; tc[lane][2] = ce.z;
    }
    __syncthreads();
    const float invA = 1.0f / pbox.half_area();
    const float leafCost = kCIsect * (float)n;
    const int axis = lane / nb, bin = lane % nb;
    const bool act = lane < 3 * nb;
    auto tbox_i = [&](int i) { return Box{{tb[i][0], tb[i][1], tb[i][2]}, {tb[i][3], tb[i][4], tb[i][5]}}; };
    auto save_bin = [&](int slot, const Box& b) {
        bbox[slot][0] = b.lo.x; bbox[slot][1] = b.lo.y; bbox[slot][2] = b.lo.z;
        bbox[slot][3] = b.hi.x; bbox[slot][4] = b.hi.y; bbox[slot][5] = b.hi.z;
    };
    auto bin_box = [&](int slot) {
        return Box{{bbox[slot][0], bbox[slot][1], bbox[slot][2]}, {bbox[slot][3], bbox[slot][4], bbox[slot][5]}};
    };
    // price candidate `bin` (> 0) of this lane's axis from the bin table; returns cost or +inf
    auto price = [&](bool spatialCounts) {
        if (!act || bin == 0) return kInf;
        Box la = Box::empty();
        uint32_t lc = 0;
        for (int b = 0; b < bin; ++b) { la.extend(bin_box(axis * nb + b)); lc += bc0[axis * nb + b]; }
        Box ra = Box::empty();
        uint32_t rc = 0;
        for (int b = nb - 1; b >= bin; --b) {
            ra.extend(bin_box(axis * nb + b));
            rc += spatialCounts ? bc1[axis * nb + b] : bc0[axis * nb + b];
        }
        if (lc == 0 || rc == 0) return kInf;
        return kCTrav + kCIsect * (la.half_area() * (float)lc + ra.half_area() * (float)rc) * invA;
    };
    // (cost, order) minimum over the wave; order = axis * 64 + (63 - bin)
    auto argmin = [&](float cost, int& wAxis, int& wBin) {
        float bc = cost;
        int key = (cost < kInf) ? axis * 64 + (63 - bin) : 0x7fffffff;
        for (int o = 32; o >= 1; o >>= 1) {
            float oc = __shfl_xor(bc, o);
            int ok = __shfl_xor(key, o);
            if (oc < bc || (oc == bc && ok < key)) { bc = oc; key = ok; }
        }
        if (key == 0x7fffffff) { wAxis = -1; wBin = -1; return kInf; }
        wAxis = key / 64;
        wBin = 63 - key % 64;
        return bc;
    };

    // ---- object split: centroid bins (bvh_sah.cpp:58-126)
    // Centroid range per axis: the reference folds std::min / std::max over the triangles
    // from +-FLT_MAX. Here lane i holds triangle i and an ordered tree reduction combines
    // (earlier, later) keeping the earlier value on ties, which is the fold's first
    // minimum; a NaN centroid never enters the fold, so it enters the tree as the identity.
    float cminA[3], cmaxA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float c = (uint32_t)lane < n ? tc[lane][a] : 0.f;
        const bool use = (uint32_t)lane < n && !__builtin_isnan(c);
        float lo = use ? c : kInf, hi = use ? c : -kInf;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const float olo = __shfl_down(lo, off), ohi = __shfl_down(hi, off);
            if ((lane & (2 * off - 1)) == 0) {
                lo = std_min(lo, olo);
                hi = std_max(hi, ohi);
            }
        }
        cminA[a] = std_min(3.40282347e+38f, __shfl(lo, 0));
        cmaxA[a] = std_max(-3.40282347e+38f, __shfl(hi, 0));
    }
    // centroid bin of each triangle on each axis (255: none, for a NaN centroid)
    if ((uint32_t)lane < n) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const bool ok = !(cmaxA[a] - cminA[a] < kCentroidEps);
            const float sc = ok ? (float)nb / (cmaxA[a] - cminA[a]) : 0.f;
            const int b = min(f2i_x86((tc[lane][a] - cminA[a]) * sc), nb - 1);
            obin[lane][a] = b >= 0 ? (uint8_t)b : (uint8_t)255;
        }
    }
    const float cmin = act ? (axis == 0 ? cminA[0] : (axis == 1 ? cminA[1] : cminA[2])) : 3.40282347e+38f;
    const float cmax = act ? (axis == 0 ? cmaxA[0] : (axis == 1 ? cmaxA[1] : cmaxA[2])) : -3.40282347e+38f;
    const bool axisOk = act && !(cmax - cmin < kCentroidEps);
    const float oscale = axisOk ? (float)nb / (cmax - cmin) : 0.f;
    __syncthreads();
    if (act) {
        Box bx = Box::empty();
        uint32_t cnt = 0;
        if (axisOk) {
            for (uint32_t i = 0; i < n; ++i) {
                if (obin[i][axis] == bin) { ++cnt; bx.extend(tbox_i(i)); }
            }
        }
        save_bin(lane, bx);
        bc0[lane] = cnt;
    }
    __syncthreads();
    int oAxis, oBin;
    float oCost = argmin(axisOk ? price(false) : kInf, oAxis, oBin);
    if (!(oCost < leafCost)) { oAxis = -1; oBin = -1; oCost = leafCost; }
    const int wl = oAxis >= 0 ? oAxis * nb + 1 : 0;          // a lane of the winning axis
    const float oCMin = __shfl(cmin, wl), oScale = __shfl(oscale, wl);
    __syncthreads();

    // ---- spatial split: node-extent bins with per-bin clipping (bvh_sah.cpp:157-243)
    // Each triangle is clipped to every bin it spans on each axis. Those (triangle, axis,
    // bin) pairs are listed in triangle order and dealt over all 64 lanes, kClipChunk at a
    // time, each lane clipping one pair into an LDS box; then each (axis, bin) lane folds
    // its boxes in triangle order, as the reference's loop over the node's triangles does.
    // (A lane per bin walking all triangles ran the clip for the whole wave at every
    // triangle while a few lanes needed it.)
    const float nmin = act ? comp(pbox.lo, axis) : 0.f, nmax = act ? comp(pbox.hi, axis) : 0.f;
    const bool sOk = act && !(nmax - nmin < kSpatialEps);
    const float sscale = sOk ? (float)nb / (nmax - nmin) : 0.f;
    float nminA[3], sscaleA[3];
    bool sOkA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {   // the same values the bin lanes of axis a compute
        nminA[a] = comp(pbox.lo, a);
        const float nmaxA = comp(pbox.hi, a);
        sOkA[a] = !(nmaxA - nminA[a] < kSpatialEps);
        sscaleA[a] = sOkA[a] ? (float)nb / (nmaxA - nminA[a]) : 0.f;
    }
    if (lane < 48) { bc0[lane] = 0; bc1[lane] = 0; }
    __syncthreads();
    uint32_t span[3] = {0, 0, 0};
    if ((uint32_t)lane < n) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!sOkA[a]) continue;
            const int b0 = clampi(f2i_x86((tb[lane][a] - nminA[a]) * sscaleA[a]), 0, nb - 1);
            const int b1 = clampi(f2i_x86((tb[lane][3 + a] - nminA[a]) * sscaleA[a]), 0, nb - 1);
            sb0[lane][a] = (uint8_t)b0;
            sb1[lane][a] = (uint8_t)b1;
            atomicAdd(&bc0[a * nb + b0], 1u);   // entry / exit counts: integer, any order
            atomicAdd(&bc1[a * nb + b1], 1u);
            span[a] = b0 <= b1 ? (uint32_t)(b1 - b0 + 1) : 0u;
        }
    }
    const uint32_t cnt = span[0] + span[1] + span[2];
    uint32_t inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    const uint32_t total = __shfl(inc, 63), tpre = inc - cnt;
    if ((uint32_t)lane < n) {
        tax[lane][0] = tpre;
        tax[lane][1] = tpre + span[0];
        tax[lane][2] = tpre + span[0] + span[1];
    }
    Box bx = Box::empty();
    for (uint32_t c0 = 0; c0 < total; c0 += kClipChunk) {
        if ((uint32_t)lane < n) {
            const uint32_t lo = max(tpre, c0), hi = min(tpre + cnt, c0 + (uint32_t)kClipChunk);
            for (uint32_t p = lo; p < hi; ++p) powner[p - c0] = (uint8_t)lane;
        }
        __syncthreads();
        for (int r = 0; r < kClipChunk; r += 64) {
            const uint32_t p = c0 + r + lane;
            if (p < total) {
                const uint32_t i = powner[r + lane];
                const int a = p >= tax[i][2] ? 2 : (p >= tax[i][1] ? 1 : 0);
                const int b = (int)sb0[i][a] + (int)(p - tax[i][a]);
                Box reg = pbox;
                set_comp(reg.lo, a, nminA[a] + (float)b / sscaleA[a]);
                set_comp(reg.hi, a, nminA[a] + (float)(b + 1) / sscaleA[a]);
                const f3 p0{tv[i][0], tv[i][1], tv[i][2]}, p1{tv[i][3], tv[i][4], tv[i][5]},
                    p2{tv[i][6], tv[i][7], tv[i][8]};
                Box cb;
                if (!clip_tri_slab(p0, p1, p2, reg, a, cb)) cb = clip_tri(p0, p1, p2, reg);
                pbuf[r + lane][0] = cb.lo.x; pbuf[r + lane][1] = cb.lo.y; pbuf[r + lane][2] = cb.lo.z;
                pbuf[r + lane][3] = cb.hi.x; pbuf[r + lane][4] = cb.hi.y; pbuf[r + lane][5] = cb.hi.z;
            }
        }
        __syncthreads();
        if (sOk) {
            // pairs are listed in triangle order: the chunk holds pairs of triangles
            // powner[0] .. powner[last] only
            const uint32_t i0 = powner[0], i1 = powner[min(total - c0, (uint32_t)kClipChunk) - 1];
            for (uint32_t i = i0; i <= i1; ++i) {
                const int b0 = sb0[i][axis], b1 = sb1[i][axis];
                if (b0 <= bin && bin <= b1) {
                    const uint32_t idx = tax[i][axis] + (uint32_t)(bin - b0);
                    if (idx >= c0 && idx < c0 + (uint32_t)kClipChunk) {
                        const uint32_t q = idx - c0;
                        const Box cb{{pbuf[q][0], pbuf[q][1], pbuf[q][2]}, {pbuf[q][3], pbuf[q][4], pbuf[q][5]}};
                        if (cb.valid()) bx.extend(cb);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (act) save_bin(lane, bx);
    __syncthreads();
    int sAxis, sBin;
    float sCost = argmin(sOk ? price(true) : kInf, sAxis, sBin);
    const bool spatial = sAxis >= 0 && sCost < oCost;
    const int ws = sAxis >= 0 ? sAxis * nb + 1 : 0;
    const float sMin = __shfl(nmin, ws), sScale = __shfl(sscale, ws);

    // ---- apply the partitions as std::partition orders them (bvh_sah.cpp:247-388). Lane l
    // holds the slot at position l; each partition is libstdc++'s two-ended algorithm
    // (stl_algo.h:1487-1510) over positions [first, last): with M elements satisfying pred,
    // the k-th failing element of [first, first+M) from the left swaps with the k-th
    // satisfying element of [first+M, last) from the right, found with ballots. Every value
    // below is the same on all lanes, so the control flow stays wave-uniform.
    Dec d;
    d.mid = 0; d.nl = d.ns = d.nr = 0; d.fb_split = 0;
    d.oAxis = oAxis; d.oBin = oBin; d.oCMin = oCMin; d.oScale = oScale; d.nb = nb;
    d.plane = 0.f; d.sAxis = -1;
    d.type = D_LEAF;
    uint32_t sv = (uint32_t)lane;
    auto opred = [&](uint32_t s_) { return min(f2i_x86((tc[s_][oAxis] - oCMin) * oScale), nb - 1) < oBin; };
    auto part = [&](uint32_t first, uint32_t last, auto pred) -> uint32_t {
        const bool inr = (uint32_t)lane >= first && (uint32_t)lane < last;
        const bool p = inr && pred(sv);
        const uint64_t below = (1ull << lane) - 1ull;
        const uint32_t M = (uint32_t)__popcll(__ballot(p));
        const uint64_t fl = __ballot(inr && !p && (uint32_t)lane < first + M);   // failing, left part
        const uint64_t tr = __ballot(inr && p && (uint32_t)lane >= first + M);   // satisfying, right part
        const bool isF = (fl >> lane) & 1ull, isT = (tr >> lane) & 1ull;
        const uint32_t rF = (uint32_t)__popcll(fl & below);                      // rank from the left
        const uint32_t rT = (uint32_t)__popcll(tr & ~below & ~(1ull << lane));   // rank from the right
        if (isT) tsi_slot[rT] = (uint32_t)lane;
        __syncthreads();
        uint32_t partner = (uint32_t)lane;
        if (isF) partner = tsi_slot[rF];
        __syncthreads();
        if (isF) tsi_slot[rF] = (uint32_t)lane;
        __syncthreads();
        if (isT) partner = tsi_slot[rT];
        __syncthreads();
        sv = __shfl(sv, (int)partner);
        return first + M;
    };
    if (oAxis == -1 && !spatial) {
        if (lane == 0) C.dec[k] = d;
        return;
    }
    if (!spatial) {
        uint32_t mid = part(0, n, opred);
        if (mid != 0 && mid != n) { d.type = D_OBJECT; d.mid = begin + mid; }
    } else {
        const float plane = sMin + (float)sBin / sScale;
        const int ax = sAxis;
        uint32_t le = part(0, n, [&](uint32_t s_) { return tb[s_][3 + ax] <= plane; });
        uint32_t se = part(le, n, [&](uint32_t s_) { return tb[s_][ax] < plane; });
        d.plane = plane; d.sAxis = ax;
        d.nl = le; d.ns = se - le; d.nr = n - se;
        if (d.nl + d.ns == 0 || d.ns + d.nr == 0) {
            if (oAxis != -1) {
                uint32_t mid = part(0, n, opred);
                if (mid != 0 && mid != n) { d.type = D_OBJECT; d.mid = begin + mid; }
            }
        } else {
            uint32_t c = 0;
            if (oAxis != -1) c = (uint32_t)__popcll(__ballot((uint32_t)lane < n && opred((uint32_t)lane)));
            d.fb_split = (oAxis != -1 && c > 0 && c < n) ? 1u : 0u;
            d.type = D_SPATIAL;
        }
    }
    if ((uint32_t)lane < n) C.tindex[begin + lane] = tsi[sv];
    if (lane == 0) C.dec[k] = d;
}

// Block-wide exclusive scan for 1024 threads (16 waves). Returns the exclusive prefix of
// v and the block total; ws needs 17 entries.
__device__ uint32_t block_scan(uint32_t v, uint32_t* ws, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t s = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(s, o);
        if (lane >= o) s += y;
    }
    if (lane == 63) ws[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 16; ++i) { uint32_t x = ws[i]; ws[i] = run; run += x; }
        ws[16] = run;
    }
    __syncthreads();
    uint32_t ex = ws[w] + s - v;
    total = ws[16];
    __syncthreads();
    return ex;
}

// Multi-block form of k_level_budget for wide levels. When every spatial candidate of the
// level fits the budget, each 1024-node chunk of k_level_budget takes its scan path, so
// the result is the exclusive scan over the whole level: per-chunk sums (k_budget_sums),
// a scan of the chunk sums with the budget check (k_budget_scan), and the per-chunk scans
// (k_budget_apply). Otherwise `fast` stays 0 and k_level_budget runs as for any level.
__device__ __forceinline__ void budget_needs(const Ctx& C, uint32_t k, uint32_t cnt, bool& sp, uint32_t& nt,
                                             uint32_t& ns, uint32_t& nch) {
    sp = false; nt = ns = nch = 0;
    if (k >= cnt) return;
    const Dec d = C.dec[k];
    sp = d.type == D_SPATIAL;
    nt = sp ? d.nr + d.ns : 0u;
    ns = sp ? d.ns : 0u;
    nch = (d.type == D_MORTON || d.type == D_OBJECT || sp) ? 2u : 0u;   // every spatial split accepted
}

__global__ void __launch_bounds__(1024) k_budget_sums(Ctx C, uint32_t* part) {
    __shared__ uint32_t ws[17];
    const uint32_t cnt = C.ctr->lvl_cnt;
    if (blockIdx.x * 1024u >= cnt) return;   // the whole block
    bool sp;
    uint32_t nt, ns, nch, tt, ts, tn;
    budget_needs(C, blockIdx.x * 1024u + threadIdx.x, cnt, sp, nt, ns, nch);
    block_scan(nt, ws, tt);
    block_scan(ns, ws, ts);
    block_scan(nch, ws, tn);
    if (threadIdx.x == 0) { part[3 * blockIdx.x] = tt; part[3 * blockIdx.x + 1] = ts; part[3 * blockIdx.x + 2] = tn; }
}

__global__ void k_budget_scan(Ctx C, uint32_t* part) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Counters* c = C.ctr;
    const uint32_t nb = (c->lvl_cnt + 1023u) / 1024u;
    uint64_t t = c->tsize, sz = c->ssize, n = c->total;
    for (uint32_t b = 0; b < nb; ++b) { t += part[3 * b]; sz += part[3 * b + 1]; }
    if (t > C.cap || sz > C.cap) return;   // some spatial split may not fit: k_level_budget decides
    uint32_t bt = c->tsize, bs = c->ssize, bn = c->total;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t xt = part[3 * b], xs = part[3 * b + 1], xn = part[3 * b + 2];
        part[3 * b] = bt; part[3 * b + 1] = bs; part[3 * b + 2] = bn;
        bt += xt; bs += xs; bn += xn;
    }
    (void)n;
    c->tsize = bt;
    c->ssize = bs;
    c->total = bn;
    c->fast = 1;
}

__global__ void __launch_bounds__(1024) k_budget_apply(Ctx C, const uint32_t* part) {
    __shared__ uint32_t ws[17];
    if (!C.ctr->fast) return;
    const uint32_t cnt = C.ctr->lvl_cnt;
    if (blockIdx.x * 1024u >= cnt) return;
    const uint32_t k = blockIdx.x * 1024u + threadIdx.x;
    bool sp;
    uint32_t nt, ns, nch, tt, ts, tn;
    budget_needs(C, k, cnt, sp, nt, ns, nch);
    const uint32_t et = block_scan(nt, ws, tt), es = block_scan(ns, ws, ts), en = block_scan(nch, ws, tn);
    if (k < cnt) {
        if (sp) {
            C.toff[k] = part[3 * blockIdx.x] + et;
            C.soff[k] = part[3 * blockIdx.x + 1] + es;
        }
        C.acc[k] = sp ? 1u : 0u;
        C.child[k] = nch ? part[3 * blockIdx.x + 2] + en : kInvalid;
    }
}

// Queue-order resolution of the spatial-split budget (bvh_sah.cpp:390-394) and child
// numbering (s_totalNodeCount++ in FIFO order). One block of 1024 threads walks the
// level in chunks; a chunk whose spatial candidates all fit is accepted with a scan,
// otherwise lane 0 runs the greedy in order (a rejected candidate costs nothing).
__global__ void __launch_bounds__(1024) k_level_budget(Ctx C) {
    if (C.ctr->fast) return;   // resolved by k_budget_sums / k_budget_scan / k_budget_apply
    const uint32_t cnt = C.ctr->lvl_cnt;
    __shared__ uint32_t ws[17];
    __shared__ uint32_t base[3];     // total nodes, s_triIndexSize, s_sortedTrisSize
    __shared__ uint32_t slow;
    const int tid = threadIdx.x;
    if (tid == 0) { base[0] = C.ctr->total; base[1] = C.ctr->tsize; base[2] = C.ctr->ssize; }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < cnt; c0 += 1024) {
        const uint32_t k = c0 + tid;
        const bool in = k < cnt;
        Dec d;
        d.type = D_LEAF; d.nr = d.ns = 0; d.fb_split = 0;
        if (in) d = C.dec[k];
        const bool sp = in && d.type == D_SPATIAL;
        const uint32_t need_t = sp ? d.nr + d.ns : 0u, need_s = sp ? d.ns : 0u;
        uint32_t tot_t, tot_s;
        uint32_t ex_t = block_scan(need_t, ws, tot_t);
        uint32_t ex_s = block_scan(need_s, ws, tot_s);
        if (tid == 0) slow = (base[1] + tot_t > C.cap || base[2] + tot_s > C.cap) ? 1u : 0u;
        __syncthreads();
        uint32_t accepted = 0;
        if (!slow) {
            if (sp) {
                accepted = 1;
                C.toff[k] = base[1] + ex_t;
                C.soff[k] = base[2] + ex_s;
            }
            __syncthreads();
            if (tid == 0) { base[1] += tot_t; base[2] += tot_s; }
        } else {
            if (tid == 0) {
                uint32_t ts = base[1], ss = base[2];
                const uint32_t e = min(cnt, c0 + 1024);
                for (uint32_t q = c0; q < e; ++q) {
                    Dec dq = C.dec[q];
                    uint32_t ok = 0;
                    if (dq.type == D_SPATIAL && !(ts + dq.nr + dq.ns > C.cap || ss + dq.ns > C.cap)) {
                        ok = 1;
                        C.toff[q] = ts;
                        C.soff[q] = ss;
                        ts += dq.nr + dq.ns;
                        ss += dq.ns;
                    }
                    C.acc[q] = ok;
                }
                base[1] = ts;
                base[2] = ss;
            }
            __syncthreads();
            if (sp) accepted = C.acc[k];
        }
        __syncthreads();
        uint32_t nch = 0;
        if (in) {
            if (d.type == D_MORTON || d.type == D_OBJECT) nch = 2;
            else if (d.type == D_SPATIAL) nch = (accepted || d.fb_split) ? 2u : 0u;
        }
        uint32_t tot_n;
        uint32_t ex_n = block_scan(nch, ws, tot_n);
        if (in) {
            C.acc[k] = accepted;
            C.child[k] = nch ? base[0] + ex_n : kInvalid;
        }
        __syncthreads();
        if (tid == 0) base[0] += tot_n;
        __syncthreads();
    }
    if (tid == 0) { C.ctr->total = base[0]; C.ctr->tsize = base[1]; C.ctr->ssize = base[2]; }
}

__device__ void set_child(const Ctx& C, uint32_t id, const Box& b, uint32_t begin, uint32_t n, uint32_t flags) {
    BNode c;
    c.lo = b.lo;
    c.hi = b.hi;
    c.a = begin;
    c.b = n;
    c.flags = flags;
    c.pad = 0;
    C.nodes[id] = c;
}

__device__ void object_children(const Ctx& C, uint32_t ni, uint32_t begin, uint32_t mid, uint32_t end, uint32_t c) {
    set_child(C, c, range_box(C, begin, mid), begin, mid - begin, F_LEAF | F_SAH);
    set_child(C, c + 1, range_box(C, mid, end), mid, end - mid, F_LEAF | F_SAH);
    BNode& p = C.nodes[ni];
    p.flags &= ~F_LEAF;
    p.a = c;
    p.b = c + 1;
}

// Ordered combine of the 64 lanes' boxes (lane order): lane 0 receives
// combine(combine(b0, b1), combine(b2, b3)) ... . The extension keeps the last of equal
// extrema, so this tree equals the serial fold over the same order. Empty boxes are
// identities.
__device__ __forceinline__ Box wave_box(Box b) {
    const int lane = __lane_id();
    for (int off = 1; off < 64; off <<= 1) {
        Box o;
        o.lo.x = __shfl_down(b.lo.x, off); o.lo.y = __shfl_down(b.lo.y, off); o.lo.z = __shfl_down(b.lo.z, off);
        o.hi.x = __shfl_down(b.hi.x, off); o.hi.y = __shfl_down(b.hi.y, off); o.hi.z = __shfl_down(b.hi.z, off);
        if ((lane & (2 * off - 1)) == 0 && lane + off < 64) b = combine(b, o);
    }
    return b;   // valid in lane 0
}

// Ordered box of triangles at s_triIndex positions [b, e), one wave, 64 at a time
// (computeRangeAABB, bvh_bfs.cpp:14-28). Result in every lane.
__device__ Box wave_range_box(const Ctx& C, uint32_t b, uint32_t e) {
    Box acc = Box::empty();
    for (uint32_t c0 = b; c0 < e; c0 += 64) {
        const uint32_t i = c0 + __lane_id();
        Box x = Box::empty();
        if (i < e) x = load_tbox(C.tbox, C.sorted[C.tindex[i]]);
        acc = combine(acc, wave_box(x));
    }
    Box r;
    r.lo.x = __shfl(acc.lo.x, 0); r.lo.y = __shfl(acc.lo.y, 0); r.lo.z = __shfl(acc.lo.z, 0);
    r.hi.x = __shfl(acc.hi.x, 0); r.hi.y = __shfl(acc.hi.y, 0); r.hi.z = __shfl(acc.hi.z, 0);
    return r;
}

// libstdc++'s two-ended std::partition (stl_algo.h:1487-1510) of s_triIndex[first, first+n)
// for n <= 64, one element per lane. With M elements satisfying pred, the algorithm swaps
// the k-th failing element of [0, M) (from the left) with the k-th satisfying element of
// [M, n) (from the right); every other element stays. Returns first + M.
template <class P>
__device__ uint32_t wave_partition(uint32_t* a, uint32_t first, uint32_t n, P pred, uint32_t* lds_lane) {
    const uint32_t lane = __lane_id();
    const bool in = lane < n;
    const uint32_t v = in ? a[first + lane] : 0u;
    const bool p = in && pred(v);
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t M = (uint32_t)__popcll(__ballot(p));
    const uint64_t fl = __ballot(in && !p && lane < M);     // failing, left of M
    const uint64_t tr = __ballot(in && p && lane >= M);     // satisfying, right of M
    const bool isF = (fl >> lane) & 1ull, isT = (tr >> lane) & 1ull;
    if (isT) lds_lane[__popcll(tr & ~below & ~(1ull << lane))] = lane;   // rank from the right
    __syncthreads();
    uint32_t partner = lane;
    if (isF) partner = lds_lane[__popcll(fl & below)];                   // rank from the left
    __syncthreads();
    if (isF) lds_lane[__popcll(fl & below)] = lane;
    __syncthreads();
    if (isT) partner = lds_lane[__popcll(tr & ~below & ~(1ull << lane))];
    const uint32_t nv = __shfl(v, (int)partner);
    if (in && partner != lane) a[first + lane] = nv;
    return first + M;
}

// Children of every frontier node that splits, one 64-lane wave per node: Morton splits
// (child boxes of ranges that will not be split by Morton bits again), object splits,
// and spatial splits (bvh_sah.cpp:247-460): right-only references appended,
// straddlers duplicated, both children clipped to the split plane.
__global__ void __launch_bounds__(64) k_level_apply(Ctx C) {
    __shared__ uint32_t lds_lane[64];
    __shared__ float poly_lds[kPolyLdsFloats];
    const uint32_t ls = C.ctr->lvl_ls, cnt = C.ctr->lvl_cnt;
    const uint32_t k = blockIdx.x;
    if (k >= cnt) return;
    const uint32_t lane = __lane_id();
    const uint32_t ni = ls + k;
    const Dec d = C.dec[k];
    const BNode nd = C.nodes[ni];
    const uint32_t begin = nd.a, n = nd.b, end = begin + n, c = C.child[k];
    if (d.type == D_LEAF || c == kInvalid) return;
    if (c + 1 >= C.ncap) {  // cannot happen for a valid build (nodes <= 2 * cap - 1)
        if (lane == 0) atomicOr(&C.ctr->err, 1u);
        return;
    }
    auto link = [&]() {
        if (lane != 0) return;
        BNode& p = C.nodes[ni];
        p.flags &= ~F_LEAF;
        p.a = c;
        p.b = c + 1;
    };
    auto object_split = [&](uint32_t mid) {
        const Box lb = wave_range_box(C, begin, mid), rb = wave_range_box(C, mid, end);
        if (lane == 0) {
            set_child(C, c, lb, begin, mid - begin, F_LEAF | F_SAH);
            set_child(C, c + 1, rb, mid, end - mid, F_LEAF | F_SAH);
        }
        link();
    };
    if (d.type == D_MORTON) {
        // Children of a Morton split (bvh_bfs.cpp:109-130). A child that will itself be
        // split by Morton bits gets its box from the bottom-up refit; the others now.
        const uint32_t cb[2] = {begin, d.mid}, ce[2] = {d.mid, end};
        for (int s2 = 0; s2 < 2; ++s2) {
            const uint32_t sz = ce[s2] - cb[s2];
            uint32_t fl = F_LEAF | (sz <= (uint32_t)kLbvhThreshold ? F_SAH : 0u);
            const bool direct = sz <= (uint32_t)kLbvhThreshold || C.scodes[cb[s2]] == C.scodes[ce[s2] - 1];
            Box b = Box::empty();
            if (direct) b = wave_range_box(C, cb[s2], ce[s2]);
            else fl |= F_PENDING;
            if (lane == 0) set_child(C, c + s2, b, cb[s2], sz, fl);
        }
        link();
        return;
    }
    if (d.type == D_OBJECT) { object_split(d.mid); return; }
    // D_SPATIAL
    if (!C.acc[k]) {
        // Budget exhausted: object-split fallback on the spatially partitioned range.
        auto pred = [&](uint32_t si) { return obj_pred(C, si, d.oAxis, d.oBin, d.oCMin, d.oScale, d.nb); };
        uint32_t mid;
        if (n <= 64) {
            mid = wave_partition(C.tindex, begin, n, pred, lds_lane);
        } else {
            if (lane == 0) lds_lane[0] = partition_idx(C.tindex, begin, end, pred);
            __syncthreads();
            mid = lds_lane[0];
        }
        __syncthreads();   // the partitioned s_triIndex is read by other lanes below
        object_split(mid);
        return;
    }
    // Accepted spatial split: append right-only references, duplicate straddlers
    // (bvh_sah.cpp:396-414), and clip children to the split plane (:427-445).
    const uint32_t tOff = C.toff[k], sOff = C.soff[k];
    const uint32_t leftEnd = begin + d.nl, strEnd = leftEnd + d.ns;
    const uint32_t lsize = d.nl + d.ns, rsize = d.nr + d.ns;
    const Box pbox{nd.lo, nd.hi};
    Box lclip = pbox, rclip = pbox;
    set_comp(lclip.hi, d.sAxis, d.plane);
    set_comp(rclip.lo, d.sAxis, d.plane);
    // Left child: positions [begin, begin + lsize), unchanged.
    Box lacc = Box::empty();
    for (uint32_t c0 = 0; c0 < lsize; c0 += 64) {
        const uint32_t q = c0 + lane;
        Box x = Box::empty();
        if (q < lsize) {
            f3 p0, p1, p2;
            tri_verts(C, C.sorted[C.tindex[begin + q]], p0, p1, p2);
            PolyLds P{poly_lds + lane};
            const Box cb = clip_tri_in(p0, p1, p2, lclip, P);
            if (cb.valid()) x = cb;
        }
        lacc = combine(lacc, wave_box(x));
    }
    // Right child, in its new order: the nr right-only references, then the ns straddlers
    // (their triangles read from the source positions, not from the copies written here).
    Box racc = Box::empty();
    for (uint32_t c0 = 0; c0 < rsize; c0 += 64) {
        const uint32_t q = c0 + lane;
        Box x = Box::empty();
        if (q < rsize) {
            uint32_t t;
            if (q < d.nr) {
                const uint32_t si = C.tindex[strEnd + q];
                C.tindex[tOff + q] = si;
                t = C.sorted[si];
            } else {
                const uint32_t qs = q - d.nr;
                const uint32_t old = C.tindex[leftEnd + qs];
                t = C.sorted[old];
                C.sorted[sOff + qs] = t;
                C.tindex[tOff + q] = sOff + qs;
            }
            f3 p0, p1, p2;
            tri_verts(C, t, p0, p1, p2);
            PolyLds P{poly_lds + lane};
            const Box cb = clip_tri_in(p0, p1, p2, rclip, P);
            if (cb.valid()) x = cb;
        }
        racc = combine(racc, wave_box(x));
    }
    if (lane == 0) {
        set_child(C, c, lacc, begin, lsize, F_LEAF | F_SAH);
        set_child(C, c + 1, racc, tOff, rsize, F_LEAF | F_SAH);
    }
    link();
}

// Bottom-up boxes for Morton-phase nodes: box(parent range) = ordered combine of the
// boxes of its two consecutive child ranges.
__global__ void k_refit(BNode* nodes, uint32_t ls, uint32_t cnt) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    BNode& n = nodes[ls + k];
    if (!(n.flags & F_PENDING) || (n.flags & F_LEAF)) return;
    const BNode& l = nodes[n.a];
    const BNode& r = nodes[n.b];
    Box b{l.lo, l.hi};
    b.extend(Box{r.lo, r.hi});
    n.lo = b.lo;
    n.hi = b.hi;
    n.flags &= ~F_PENDING;
}

struct Dfs {
    uint32_t *size, *tcnt, *icnt, *lcnt;
    uint32_t *dfs, *toff, *ioff, *loff;
};

__global__ void k_subtree(const BNode* nodes, uint32_t ls, uint32_t cnt, Dfs D) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    uint32_t i = ls + k;
    const BNode n = nodes[i];
    if (n.flags & F_LEAF) {
        D.size[i] = 1; D.tcnt[i] = n.b; D.icnt[i] = 0; D.lcnt[i] = 1;
    } else {
        D.size[i] = 1 + D.size[n.a] + D.size[n.b];
        D.tcnt[i] = D.tcnt[n.a] + D.tcnt[n.b];
        D.icnt[i] = 1 + D.icnt[n.a] + D.icnt[n.b];
        D.lcnt[i] = D.lcnt[n.a] + D.lcnt[n.b];
    }
}

// DFS pre-order, left first (reorderNodes, bvh_bake.cpp:75-118); leaves own consecutive
// triangle runs in that order (reorderTriangles, :120-149).
__global__ void k_dfs(const BNode* nodes, uint32_t ls, uint32_t cnt, Dfs D) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    uint32_t p = ls + k;
    if (p == 0) { D.dfs[0] = 0; D.toff[0] = 0; D.ioff[0] = 0; D.loff[0] = 0; }
    const BNode n = nodes[p];
    if (n.flags & F_LEAF) return;
    uint32_t dp = (p == 0) ? 0 : D.dfs[p], tp = (p == 0) ? 0 : D.toff[p];
    uint32_t ip = (p == 0) ? 0 : D.ioff[p], lp = (p == 0) ? 0 : D.loff[p];
    D.dfs[n.a] = dp + 1;
    D.dfs[n.b] = dp + 1 + D.size[n.a];
    D.toff[n.a] = tp;
    D.toff[n.b] = tp + D.tcnt[n.a];
    D.ioff[n.a] = ip + 1;
    D.ioff[n.b] = ip + 1 + D.icnt[n.a];
    D.loff[n.a] = lp;
    D.loff[n.b] = lp + D.lcnt[n.a];
}

struct Out {
    Node48* nodes48;
    Node48* bfs48;
    uint32_t* tris3;
    uint32_t* prims;
    GNode* gnodes;
    Leaf* leaves;
    GTri* gtris;
};

__device__ __forceinline__ Node48 to48(const BNode& n, uint32_t a, uint32_t b) {
    Node48 o;
    o.bmin[0] = n.lo.x; o.bmin[1] = n.lo.y; o.bmin[2] = n.lo.z; o.bmin[3] = 0.f;
    o.bmax[0] = n.hi.x; o.bmax[1] = n.hi.y; o.bmax[2] = n.hi.z; o.bmax[3] = 0.f;
    o.is_leaf = (n.flags & F_LEAF) ? 1 : 0;
    o.pad0[0] = o.pad0[1] = o.pad0[2] = 0;
    o.a = a;
    o.b = b;
    o.pad1 = 0;
    return o;
}

// Sets *flag when any node box coordinate is not finite (the traversal then keeps the
// NaN-exact slab test for every ray).
__global__ void k_boxes_nonfinite(const Node48* nodes, uint32_t nn, uint32_t* flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const Node48& n = nodes[i];
    bool ok = true;
    for (int k = 0; k < 3; ++k) ok = ok && __builtin_isfinite(n.bmin[k]) && __builtin_isfinite(n.bmax[k]);
    if (!ok) *flag = 1u;
}

__global__ void k_emit(Ctx C, uint32_t total, Dfs D, Out O) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const BNode n = C.nodes[i];
    const bool leaf = (n.flags & F_LEAF) != 0;
    O.bfs48[i] = to48(n, n.a, n.b);
    O.nodes48[D.dfs[i]] = to48(n, leaf ? D.toff[i] : D.dfs[n.a], leaf ? n.b : D.dfs[n.b]);
    if (!leaf) {
        const BNode l = C.nodes[n.a], r = C.nodes[n.b];
        GNode g;
        g.l0 = make_float4(l.lo.x, l.lo.y, l.lo.z, l.hi.x);
        g.l1 = make_float4(l.hi.y, l.hi.z, r.lo.x, r.lo.y);
        g.r1 = make_float4(r.lo.z, r.hi.x, r.hi.y, r.hi.z);
        g.lref = (l.flags & F_LEAF) ? leaf_ref(D.loff[n.a], D.toff[n.a], l.b) : D.ioff[n.a];
        g.rref = (r.flags & F_LEAF) ? leaf_ref(D.loff[n.b], D.toff[n.b], r.b) : D.ioff[n.b];
        g.pad0 = g.pad1 = 0;
        O.gnodes[D.ioff[i]] = g;
        return;
    }
    const uint32_t o0 = D.toff[i];
    O.leaves[D.loff[i]] = Leaf{o0, n.b};
    for (uint32_t k = 0; k < n.b; ++k) {
        uint32_t t = C.sorted[C.tindex[n.a + k]];
        uint32_t o = o0 + k;
        uint32_t i0 = C.tris[3 * t], i1 = C.tris[3 * t + 1], i2 = C.tris[3 * t + 2];
        O.tris3[3 * o] = i0; O.tris3[3 * o + 1] = i1; O.tris3[3 * o + 2] = i2;
        O.prims[o] = t;
        f3 v0 = vpos(C.verts, i0), v1 = vpos(C.verts, i1), v2 = vpos(C.verts, i2);
        store_gtri(O.gtris + o, v0, sub(v1, v0), sub(v2, v0));
    }
}

// ---------------------------------------------------------------------------
// Wide traversal layout: binary nodes collapsed four children at a time, top down. A wide
// node rooted at binary node i starts from i's two children and repeatedly replaces the
// internal child of largest surface area by its two children (in place, so the slots keep
// the binary tree's left-to-right order) until four slots are filled or only leaves are
// left; its internal slots root the wide nodes of the next level. Wide node indices are the
// binary internal indices of their roots, so no renumbering is needed. One launch per level:
// wl[i] holds the level of the wide node rooted at i (0xFF: absorbed or not yet reached).
// ---------------------------------------------------------------------------
__global__ void k_wide_init(uint8_t* wl, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wl[i] = i == 0 ? 0 : 0xFF;
}

__global__ void k_wide_level(const GNode* __restrict__ g, uint32_t n, uint32_t level, uint8_t* wl,
                             WNode* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || wl[i] != level) return;
    float lo[3][4], hi[3][4], area[4];
    uint32_t rf[4];
    auto put = [&](int s, const float* b, uint32_t r) {   // b: min xyz, max xyz
        for (int a = 0; a < 3; ++a) { lo[a][s] = b[a]; hi[a][s] = b[3 + a]; }
        const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        area[s] = (dx * dy + dy * dz) + dz * dx;
        rf[s] = r;
    };
    auto children = [&](const GNode& q, float* l6, float* r6) {
        l6[0] = q.l0.x; l6[1] = q.l0.y; l6[2] = q.l0.z; l6[3] = q.l0.w; l6[4] = q.l1.x; l6[5] = q.l1.y;
        r6[0] = q.l1.z; r6[1] = q.l1.w; r6[2] = q.r1.x; r6[3] = q.r1.y; r6[4] = q.r1.z; r6[5] = q.r1.w;
    };
    float l6[6], r6[6];
    GNode q = g[i];
    children(q, l6, r6);
    put(0, l6, q.lref);
    put(1, r6, q.rref);
    int m = 2;
    while (m < 4) {
        int e = -1;
        float best = -1.f;
        for (int s = 0; s < m; ++s)
            if (!(rf[s] & kLeafBit) && area[s] > best) { best = area[s]; e = s; }
        if (e < 0) break;
        q = g[rf[e]];
        children(q, l6, r6);
        for (int s = m; s > e + 1; --s) {   // open slot e + 1
            for (int a = 0; a < 3; ++a) { lo[a][s] = lo[a][s - 1]; hi[a][s] = hi[a][s - 1]; }
            area[s] = area[s - 1];
            rf[s] = rf[s - 1];
        }
        put(e, l6, q.lref);
        put(e + 1, r6, q.rref);
        ++m;
    }
    for (int s = m; s < 4; ++s) {
        for (int a = 0; a < 3; ++a) { lo[a][s] = 0.f; hi[a][s] = 0.f; }
        rf[s] = kInvalid;
    }
    WNode w;
    w.lx = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    w.ly = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    w.lz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    w.hx = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    w.hy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    w.hz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    w.ref = make_uint4(rf[0], rf[1], rf[2], rf[3]);
    w.pad = make_uint4(0, 0, 0, 0);
    out[i] = w;
    for (int s = 0; s < m; ++s)
        if (!(rf[s] & kLeafBit)) wl[rf[s]] = (uint8_t)(level + 1);
}

// Traversal layout from a reference-layout tree (a loaded .bvh file).
__global__ void k_from_ref(const Node48* nodes, uint32_t nn, const uint32_t* iidx, const uint32_t* lidx,
                           GNode* gnodes, Leaf* leaves) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const Node48 n = nodes[i];
    if (n.is_leaf) {
        leaves[lidx[i]] = Leaf{n.a, n.b};
        return;
    }
    const Node48 l = nodes[n.a], r = nodes[n.b];
    GNode g;
    g.l0 = make_float4(l.bmin[0], l.bmin[1], l.bmin[2], l.bmax[0]);
    g.l1 = make_float4(l.bmax[1], l.bmax[2], r.bmin[0], r.bmin[1]);
    g.r1 = make_float4(r.bmin[2], r.bmax[0], r.bmax[1], r.bmax[2]);
    g.lref = l.is_leaf ? leaf_ref(lidx[n.a], l.a, l.b) : iidx[n.a];
    g.rref = r.is_leaf ? leaf_ref(lidx[n.b], r.a, r.b) : iidx[n.b];
    g.pad0 = g.pad1 = 0;
    gnodes[iidx[i]] = g;
}

__global__ void k_gtris_from_ref(const uint32_t* tris3, const uint32_t* prims, uint32_t ntri, const float* verts,
                                 GTri* gtris) {
    uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= ntri) return;
    f3 v0 = vpos(verts, tris3[3 * o]), v1 = vpos(verts, tris3[3 * o + 1]), v2 = vpos(verts, tris3[3 * o + 2]);
    store_gtri(gtris + o, v0, sub(v1, v0), sub(v2, v0));
}

inline uint32_t nblk(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// Two pinned Counters slots and their events: the host reads one level's counters while
// the next level runs.
struct PinnedCounters {
    Counters* h = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    PinnedCounters() {
        if (hipHostMalloc((void**)&h, 2 * sizeof(Counters), hipHostMallocDefault) != hipSuccess) h = nullptr;
        for (auto& e : ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
    }
    bool ok() const { return h && ev[0] && ev[1]; }
    ~PinnedCounters() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (h) (void)hipHostFree(h);
    }
};

// Process-wide cache of device scratch blocks, per device. A build allocates about 380
// bytes of scratch per triangle in some twenty blocks; hipFree synchronizes the device and
// unmaps, and freeing them cost more than the build's kernels. A block goes back to the cache
// either idle, or with an event recorded on the stream whose queued work may still use it
// (a render returns its path state that way, without waiting for the GPU); whoever takes it
// next makes its own stream wait for that event. The next user takes the smallest cached
// block of at least the size it needs (at most twice that). At most kMaxCached bytes stay
// cached; beyond that blocks are freed. 16 GiB holds a build of 40M triangles or the path
// state of a 2^26-path render batch (6.7 GB) with its adjoint records, a few percent of the
// 288 GB of an MI355X.
struct ScratchPool {
    static constexpr size_t kMaxCached = size_t(16) << 30;
    struct Block {
        int dev;
        size_t bytes;
        void* p;
        hipEvent_t ev;   // recorded after the last work that used the block, or null
    };
    std::mutex mu;
    std::vector<Block> blocks;
    std::vector<std::pair<int, hipEvent_t>> spare;   // events to record again, per device
    size_t cached = 0;
    // bytes requested by the blocks handed out and not yet returned, and their peak since the
    // last reset (rtx_debug_scratch_peak)
    std::unordered_map<void*, size_t> live;
    size_t live_bytes = 0, peak = 0;
    void note_take(void* p, size_t bytes) {
        live[p] = bytes;
        live_bytes += bytes;
        peak = std::max(peak, live_bytes);
    }
    void* take(int dev, size_t bytes, size_t& got, hipStream_t st) {
        void* p = nullptr;
        hipEvent_t ev = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu);
            int bi = -1;
            for (int i = 0; i < (int)blocks.size(); ++i) {
                const Block& b = blocks[i];
                if (b.dev == dev && b.bytes >= bytes && b.bytes <= 2 * bytes + 4096 &&
                    (bi < 0 || b.bytes < blocks[bi].bytes))
                    bi = i;
            }
            if (bi >= 0) {
                p = blocks[bi].p;
                ev = blocks[bi].ev;
                got = blocks[bi].bytes;
                cached -= got;
                blocks.erase(blocks.begin() + bi);
                note_take(p, bytes);
            }
        }
        if (p) {
            if (ev) {
                // the wait takes the event's current record; the event can be recorded again
                if (hipStreamWaitEvent(st, ev, 0) != hipSuccess) (void)hipEventSynchronize(ev);
                std::lock_guard<std::mutex> lk(mu);
                spare.push_back({dev, ev});
            }
            return p;
        }
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        got = bytes;
        std::lock_guard<std::mutex> lk(mu);
        note_take(p, bytes);
        return p;
    }
    // pending: work queued on `after` may still use the block (otherwise it is idle)
    void give(int dev, size_t bytes, void* p, hipStream_t after, bool pending) {
        hipEvent_t ev = nullptr;
        bool keep = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = live.find(p);
            if (it != live.end()) {
                live_bytes -= it->second;
                live.erase(it);
            }
            keep = cached + bytes <= kMaxCached;
            if (keep) cached += bytes;   // reserved: the block is added below
            if (keep && pending)
                for (size_t i = 0; i < spare.size(); ++i)
                    if (spare[i].first == dev) {
                        ev = spare[i].second;
                        spare.erase(spare.begin() + i);
                        break;
                    }
        }
        if (!keep) {
            if (pending) (void)hipStreamSynchronize(after);
            (void)hipFree(p);
            return;
        }
        if (pending) {
            if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
            if (!ev || hipEventRecord(ev, after) != hipSuccess) {   // no event: wait here instead
                if (ev) (void)hipEventDestroy(ev);
                ev = nullptr;
                (void)hipStreamSynchronize(after);
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        blocks.push_back({dev, bytes, p, ev});
    }
};
ScratchPool g_scratch_pool;

// Scratch buffers of one build, taken from the cache; returned when the build's stream is
// idle.
struct Scratch {
    hipStream_t st;
    std::vector<std::pair<void*, size_t>> blocks;
    explicit Scratch(hipStream_t s) : st(s) {}
    template <class T>
    T* get(size_t n) {
        size_t got = 0;
        T* p = (T*)scratch_take(std::max<size_t>(n, 1) * sizeof(T), got, st);
        if (p) blocks.push_back({p, got});
        return p;
    }
    ~Scratch() {
        (void)hipStreamSynchronize(st);   // no kernel of this build still uses the blocks
        for (auto& b : blocks) scratch_give(b.first, b.second);
    }
};

}  // namespace

void* scratch_take(size_t bytes, size_t& got, hipStream_t st) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    return g_scratch_pool.take(dev, bytes, got, st);
}
void scratch_give(void* p, size_t bytes) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    g_scratch_pool.give(dev, bytes, p, nullptr, false);
}
void scratch_give_after(void* p, size_t bytes, hipStream_t st) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    g_scratch_pool.give(dev, bytes, p, st, true);
}
uint64_t scratch_peak(bool reset) {
    std::lock_guard<std::mutex> lk(g_scratch_pool.mu);
    const uint64_t peak = g_scratch_pool.peak;
    if (reset) g_scratch_pool.peak = g_scratch_pool.live_bytes;
    return peak;
}

namespace {

void free_products(rtx_bvh* b) {
    void* ps[] = {b->verts, b->nodes48, b->tris3, b->prims, b->gnodes, b->wnodes, b->leaves, b->gtris, b->codes,
                  b->sorted,
                  b->bfs48, b->work, b->rs_keys[0], b->rs_keys[1], b->rs_vals[0], b->rs_vals[1], b->rs_hist,
                  b->rs_part, b->host_rays, b->host_hits, b->host_nrm};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (b->scratch_ev) (void)hipEventDestroy(b->scratch_ev);
    b->scratch_ev = nullptr;
    void* pins[] = {b->pin_rays, b->pin_hits, b->pin_nrm};
    for (void* p : pins)
        if (p) (void)hipHostFree(p);
    b->pin_rays = b->pin_hits = nullptr;
    b->pin_nrm = nullptr;
    b->pin_rays_d = b->pin_hits_d = b->pin_nrm_d = nullptr;
    b->pin_cap = 0;
    if (b->host_stream) (void)hipStreamDestroy(b->host_stream);
    b->host_stream = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
static int build_impl(const float* verts_d, uint32_t nv, const uint32_t* tris_d, uint32_t nt, hipStream_t st,
                      rtx_bvh* B) {
    if (nt == 0 || nv == 0) return fail(RTX_E_ARG, "no triangles or vertices on rtx_build()");
    if (nt > (1u << 27)) return fail(RTX_E_LIMIT, "more than 2^27 triangles on rtx_build()");
    Scratch S(st);
    const uint32_t N = nt;
    const uint32_t cap = std::max<uint32_t>((uint32_t)((float)N * 1.5f), N);
    const uint32_t ncap = 2 * cap + 1;

    uint32_t* err = S.get<uint32_t>(1);
    float4* tbox = S.get<float4>(2 * (size_t)N);
    float4* cent = S.get<float4>(N);
    float* box6 = S.get<float>(8);
    uint64_t* keysB = S.get<uint64_t>(N);
    uint64_t* keysC = S.get<uint64_t>(N);
    uint32_t* valsA = S.get<uint32_t>(N);
    uint32_t* valsB = S.get<uint32_t>(N);
    unsigned long long* orand = S.get<unsigned long long>(2);
    uint32_t* sorted = S.get<uint32_t>(cap);
    uint32_t* tindex = S.get<uint32_t>(cap);
    BNode* nodes = S.get<BNode>(ncap);
    Dec* dec = S.get<Dec>(cap + 1);
    uint32_t* child = S.get<uint32_t>(cap + 1);
    uint32_t* toff = S.get<uint32_t>(cap + 1);
    uint32_t* soff = S.get<uint32_t>(cap + 1);
    uint32_t* acc = S.get<uint32_t>(cap + 1);
    Counters* ctr = S.get<Counters>(1);
    uint32_t* bpart = S.get<uint32_t>(3 * ((size_t)ncap / 1024 + 2));   // chunk sums of k_budget_*
    uint32_t* wlist = S.get<uint32_t>(cap + 1);
    Dfs D;
    uint32_t* dbuf = S.get<uint32_t>((size_t)ncap * 8);
    B->codes = dalloc<uint64_t>(N);
    B->sorted = dalloc<uint32_t>(N);
    B->verts = dalloc<float>((size_t)nv * 6);
    if (!err || !tbox || !cent || !box6 || !keysB || !keysC || !valsA || !valsB || !orand || !sorted ||
        !tindex || !nodes || !dec || !child || !toff || !soff || !acc || !ctr || !bpart || !dbuf || !B->codes || !B->sorted ||
        !B->verts)
        return fail(RTX_E_HIP, "device allocation failed on rtx_build()");
    D.size = dbuf; D.tcnt = dbuf + ncap; D.icnt = dbuf + 2 * (size_t)ncap; D.lcnt = dbuf + 3 * (size_t)ncap;
    D.dfs = dbuf + 4 * (size_t)ncap; D.toff = dbuf + 5 * (size_t)ncap; D.ioff = dbuf + 6 * (size_t)ncap;
    D.loff = dbuf + 7 * (size_t)ncap;

    RTX_HIP(hipMemcpyAsync(B->verts, verts_d, sizeof(float) * 6 * nv, hipMemcpyDeviceToDevice, st));
    RTX_HIP(hipMemsetAsync(err, 0, 4, st));
    k_check_indices<<<nblk(3ull * N, 256), 256, 0, st>>>(tris_d, 3 * N, nv, err);
    uint32_t herr = 0;
    RTX_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
    RTX_HIP(hipStreamSynchronize(st));
    if (herr) return fail(RTX_E_ARG, "triangle has out-of-bounds vertex index on rtx_build()");

    const float* V = B->verts;
    k_tri_prep<<<nblk(N, 256), 256, 0, st>>>(V, tris_d, N, tbox, cent);
    // scene box: ordered two-level reduction
    const uint32_t chunk = 16;
    const uint32_t nparts = nblk(N, chunk * kRedThreads);
    float4* part = S.get<float4>(2 * (size_t)nparts);
    if (!part) return fail(RTX_E_HIP, "device allocation failed on rtx_build()");
    k_scene_box_partial<<<nparts, kRedThreads, 0, st>>>(tbox, N, chunk, part);
    k_scene_box_final<<<1, 64, 0, st>>>(part, nparts, box6);
    unsigned long long orand_init[2] = {0ull, ~0ull};
    RTX_HIP(hipMemcpyAsync(orand, orand_init, sizeof(orand_init), hipMemcpyHostToDevice, st));
    k_morton<<<std::min(nblk(N, 256), kMortonBlocks), 256, 0, st>>>(cent, N, box6, B->codes, valsA, orand);
    unsigned long long hor[2];
    RTX_HIP(hipMemcpyAsync(hor, orand, sizeof(hor), hipMemcpyDeviceToHost, st));
    RTX_HIP(hipMemcpyAsync(B->scene, box6, 6 * sizeof(float), hipMemcpyDeviceToHost, st));
    RTX_HIP(hipStreamSynchronize(st));

    // radix sort (rtx_sort.hip): skip passes whose digit is the same for every key
    uint32_t* hist = S.get<uint32_t>(sort_hist_entries(N));
    uint32_t* spart = S.get<uint32_t>(sort_part_entries(N));
    if (!hist || !spart) return fail(RTX_E_HIP, "device allocation failed on rtx_build()");
    const uint64_t vary = hor[0] ^ hor[1];
    uint32_t pass_mask = 0;
    for (int pass = 0; pass < 8; ++pass)
        if ((vary >> (8 * pass)) & 255u) pass_mask |= 1u << pass;
    RTX_HIP(hipMemcpyAsync(keysB, B->codes, sizeof(uint64_t) * N, hipMemcpyDeviceToDevice, st));
    SortBuffers SB{{keysB, keysC}, {valsA, valsB}, hist, spart, 0};
    if (radix_sort_pairs(SB, N, pass_mask, st)) return fail(RTX_E_HIP, "Morton code sort failed on rtx_build()");
    const uint64_t* scodes = SB.keys[SB.cur];
    const uint32_t* vin = SB.vals[SB.cur];
    RTX_HIP(hipMemcpyAsync(sorted, vin, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, st));
    RTX_HIP(hipMemcpyAsync(B->sorted, vin, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, st));
    k_iota<<<nblk(N, 256), 256, 0, st>>>(tindex, N);
    k_init_root<<<1, 64, 0, st>>>(nodes, box6, N, tindex, ctr);

    Ctx C{V, tris_d, tbox, cent, scodes, sorted, tindex, nodes, dec, child, toff, soff, acc, cap, ncap, ctr, wlist};
    // Levels are queued one ahead of the host: level L+1 is launched (grids sized for at
    // most two children per node of level L) before the host reads level L's counters, so
    // the GPU does not idle between levels while the host waits on a copy.
    std::vector<std::pair<uint32_t, uint32_t>> levels;
    PinnedCounters pc;
    if (!pc.ok()) return fail(RTX_E_HIP, "pinned host allocation failed on rtx_build()");
    auto launch_level = [&](int slot, uint32_t bound) -> int {
        k_level_start<<<1, 64, 0, st>>>(ctr);
        k_level_decide<<<nblk(bound, 64), 64, 0, st>>>(C);
        k_sah_wave<<<bound, 64, 0, st>>>(C);
        if (bound > 1024) {   // more than one 1024-node chunk: try the multi-block budget pass
            const uint32_t nb = nblk(bound, 1024);
            k_budget_sums<<<nb, 1024, 0, st>>>(C, bpart);
            k_budget_scan<<<1, 64, 0, st>>>(C, bpart);
            k_budget_apply<<<nb, 1024, 0, st>>>(C, bpart);
        }
        k_level_budget<<<1, 1024, 0, st>>>(C);
        k_level_apply<<<bound, 64, 0, st>>>(C);
        RTX_HIP(hipMemcpyAsync(pc.h + slot, ctr, sizeof(Counters), hipMemcpyDeviceToHost, st));
        RTX_HIP(hipEventRecord(pc.ev[slot], st));
        return RTX_OK;
    };
    uint32_t ls = 0, cnt = 1, total = 1;
    Counters hlast{};
    int rc = launch_level(0, 1);
    for (int L = 0; rc == RTX_OK; ++L) {
        levels.push_back({ls, cnt});
        rc = launch_level((L + 1) & 1, 2 * cnt);
        if (rc != RTX_OK) break;
        if (!hip_ok(hipEventSynchronize(pc.ev[L & 1]), "hipEventSynchronize")) { rc = RTX_E_HIP; break; }
        const Counters hc = pc.h[L & 1];
        if (hc.err || hc.total > ncap) { rc = fail(RTX_E_LIMIT, "node capacity exceeded on rtx_build()"); break; }
        hlast = hc;
        total = hc.total;
        ls += cnt;
        cnt = total - ls;
        if (cnt == 0) break;   // the level queued last has no nodes: its blocks all exit
    }
    if (rc != RTX_OK) {
        (void)hipStreamSynchronize(st);   // no kernel may still use the scratch freed on return
        return rc;
    }
    for (size_t l = levels.size(); l-- > 0;)
        k_refit<<<nblk(levels[l].second, 256), 256, 0, st>>>(nodes, levels[l].first, levels[l].second);
    for (size_t l = levels.size(); l-- > 0;)
        k_subtree<<<nblk(levels[l].second, 256), 256, 0, st>>>(nodes, levels[l].first, levels[l].second, D);
    for (size_t l = 0; l < levels.size(); ++l)
        k_dfs<<<nblk(levels[l].second, 256), 256, 0, st>>>(nodes, levels[l].first, levels[l].second, D);
    uint32_t root_counts[4];  // tcnt, icnt, lcnt
    RTX_HIP(hipMemcpyAsync(&root_counts[0], D.tcnt, 4, hipMemcpyDeviceToHost, st));
    RTX_HIP(hipMemcpyAsync(&root_counts[1], D.icnt, 4, hipMemcpyDeviceToHost, st));
    RTX_HIP(hipMemcpyAsync(&root_counts[2], D.lcnt, 4, hipMemcpyDeviceToHost, st));
    RTX_HIP(hipStreamSynchronize(st));
    B->nn = total;
    B->ntri = root_counts[0];
    B->ninternal = root_counts[1];
    B->nleaves = root_counts[2];
    B->nodes48 = dalloc<Node48>(total);
    B->bfs48 = dalloc<Node48>(total);
    B->tris3 = dalloc<uint32_t>(3 * (size_t)B->ntri);
    B->prims = dalloc<uint32_t>(B->ntri);
    B->gnodes = dalloc<GNode>(std::max<uint32_t>(B->ninternal, 1));
    B->leaves = dalloc<Leaf>(B->nleaves);
    B->gtris = dalloc<GTri>(B->ntri);
    B->work = dalloc<uint32_t>(64);
    if (!B->nodes48 || !B->bfs48 || !B->tris3 || !B->prims || !B->gnodes || !B->leaves || !B->gtris || !B->work)
        return fail(RTX_E_HIP, "device allocation failed on rtx_build()");
    Out O{B->nodes48, B->bfs48, B->tris3, B->prims, B->gnodes, B->leaves, B->gtris};
    k_emit<<<nblk(total, 128), 128, 0, st>>>(C, total, D, O);
    B->max_depth = (uint32_t)levels.size() - 1;
    rc = build_wide(B, st);
    if (rc != RTX_OK) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    RTX_HIP(hipMemsetAsync(B->work + 1, 0, sizeof(uint32_t), st));
    k_boxes_nonfinite<<<nblk(total, 256), 256, 0, st>>>(B->nodes48, total, B->work + 1);
    BNode root;
    uint32_t nonfinite = 1;
    RTX_HIP(hipMemcpyAsync(&root, nodes, sizeof(root), hipMemcpyDeviceToHost, st));
    RTX_HIP(hipMemcpyAsync(&nonfinite, B->work + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RTX_HIP(hipStreamSynchronize(st));
    B->boxes_finite = nonfinite == 0;
    B->root_box[0] = root.lo.x; B->root_box[1] = root.lo.y; B->root_box[2] = root.lo.z;
    B->root_box[3] = root.hi.x; B->root_box[4] = root.hi.y; B->root_box[5] = root.hi.z;
    B->root_ref = (root.flags & F_LEAF) ? kLeafBit : 0u;
    B->nv = nv;
    B->nt = nt;
    B->levels = (uint32_t)levels.size();
    B->max_depth = B->levels - 1;
    B->tindex_size = hlast.tsize;
    B->sorted_size = hlast.ssize;
    return RTX_OK;
}

int upload_reference_tree(rtx_bvh* B, const Node48* nodes, uint32_t nn, const uint32_t* tris3, uint32_t ntri,
                          const float* verts6, uint32_t nv) {
    if (nn == 0) return fail(RTX_E_ARG, "empty tree");
    std::vector<uint32_t> iidx(nn), lidx(nn);
    uint32_t ni = 0, nl = 0;
    for (uint32_t i = 0; i < nn; ++i) {
        if (nodes[i].is_leaf) {
            lidx[i] = nl++;
            if ((uint64_t)nodes[i].a + nodes[i].b > ntri) return fail(RTX_E_ARG, "leaf triangle range out of bounds");
        } else {
            iidx[i] = ni++;
            if (nodes[i].a >= nn || nodes[i].b >= nn) return fail(RTX_E_ARG, "child index out of bounds");
        }
    }
    for (uint64_t i = 0; i < 3ull * ntri; ++i)
        if (tris3[i] >= nv) return fail(RTX_E_ARG, "triangle vertex index out of bounds");
    B->nn = nn; B->ntri = ntri; B->nv = nv; B->nt = ntri;
    if (nl >= (1u << 30)) return fail(RTX_E_LIMIT, "more than 2^30 leaves on rtx_load()");
    B->ninternal = ni; B->nleaves = nl;
    B->nodes48 = dalloc<Node48>(nn);
    B->tris3 = dalloc<uint32_t>(3 * (size_t)std::max<uint32_t>(ntri, 1));
    B->prims = dalloc<uint32_t>(std::max<uint32_t>(ntri, 1));
    B->verts = dalloc<float>(6 * (size_t)nv);
    B->gnodes = dalloc<GNode>(std::max<uint32_t>(ni, 1));
    B->leaves = dalloc<Leaf>(std::max<uint32_t>(nl, 1));
    B->gtris = dalloc<GTri>(std::max<uint32_t>(ntri, 1));
    B->work = dalloc<uint32_t>(64);
    uint32_t* d_i = dalloc<uint32_t>(nn);
    uint32_t* d_l = dalloc<uint32_t>(nn);
    if (!B->nodes48 || !B->tris3 || !B->prims || !B->verts || !B->gnodes || !B->leaves || !B->gtris || !B->work ||
        !d_i || !d_l)
        return fail(RTX_E_HIP, "device allocation failed on rtx_load()");
    std::vector<uint32_t> prims(ntri);
    for (uint32_t i = 0; i < ntri; ++i) prims[i] = i;  // a .bvh file keeps no original ids
    bool ok = hip_ok(hipMemcpy(B->nodes48, nodes, sizeof(Node48) * nn, hipMemcpyHostToDevice), "copy") &&
              hip_ok(hipMemcpy(B->tris3, tris3, 12ull * ntri, hipMemcpyHostToDevice), "copy") &&
              hip_ok(hipMemcpy(B->prims, prims.data(), 4ull * ntri, hipMemcpyHostToDevice), "copy") &&
              hip_ok(hipMemcpy(B->verts, verts6, 24ull * nv, hipMemcpyHostToDevice), "copy") &&
              hip_ok(hipMemcpy(d_i, iidx.data(), 4ull * nn, hipMemcpyHostToDevice), "copy") &&
              hip_ok(hipMemcpy(d_l, lidx.data(), 4ull * nn, hipMemcpyHostToDevice), "copy");
    if (ok) {
        k_from_ref<<<nblk(nn, 256), 256>>>(B->nodes48, nn, d_i, d_l, B->gnodes, B->leaves);
        if (ntri) k_gtris_from_ref<<<nblk(ntri, 256), 256>>>(B->tris3, B->prims, ntri, B->verts, B->gtris);
        ok = hip_ok(hipDeviceSynchronize(), "k_from_ref");
    }
    (void)hipFree(d_i);
    (void)hipFree(d_l);
    if (!ok) return RTX_E_HIP;
    for (int k = 0; k < 3; ++k) { B->root_box[k] = nodes[0].bmin[k]; B->root_box[3 + k] = nodes[0].bmax[k]; }
    B->root_ref = nodes[0].is_leaf ? kLeafBit : 0u;
    B->boxes_finite = true;
    for (uint32_t i = 0; i < nn; ++i)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(nodes[i].bmin[k]) || !std::isfinite(nodes[i].bmax[k])) B->boxes_finite = false;
    // Depth by a host walk; every node must be reached exactly once (a tree, no cycles).
    std::vector<uint8_t> seen(nn, 0);
    std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 0u}};
    uint32_t depth = 0, reached = 0;
    while (!st.empty()) {
        auto [i, d] = st.back();
        st.pop_back();
        if (seen[i]) return fail(RTX_E_ARG, "node reached twice: not a tree");
        seen[i] = 1;
        ++reached;
        depth = std::max(depth, d);
        if (!nodes[i].is_leaf) { st.push_back({nodes[i].a, d + 1}); st.push_back({nodes[i].b, d + 1}); }
    }
    B->max_depth = depth;
    return build_wide(B, nullptr);
}

// Widening pass shared by builds and loaded trees (see k_wide_level). The level scratch
// comes from the scratch cache and goes back once the pass has run.
int build_wide(rtx_bvh* B, hipStream_t st) {
    const uint32_t n = B->ninternal;
    if (n == 0) return RTX_OK;   // the root is a leaf
    B->wnodes = dalloc<WNode>(n);
    size_t got = 0;
    uint8_t* wl = static_cast<uint8_t*>(scratch_take(n, got, st));
    if (!B->wnodes || !wl) {
        if (wl) scratch_give(wl, got);
        return fail(RTX_E_HIP, "device allocation failed for the wide traversal nodes");
    }
    k_wide_init<<<nblk(n, 256), 256, 0, st>>>(wl, n);
    // a wide level holds at least one binary level, so max_depth + 1 launches reach every node
    for (uint32_t L = 0; L <= B->max_depth && L < 255; ++L)
        k_wide_level<<<nblk(n, 256), 256, 0, st>>>(B->gnodes, n, L, wl, B->wnodes);
    const bool ok = hip_ok(hipGetLastError(), "k_wide_level") && hip_ok(hipStreamSynchronize(st), "wide pass");
    scratch_give(wl, got);
    return ok ? RTX_OK : RTX_E_HIP;
}

}  // namespace rtx

using namespace rtx;

extern "C" int rtx_build_device(const float* verts6_d, uint32_t nv, const uint32_t* tris_d, uint32_t nt, void* stream,
                                rtx_bvh** out) {
    if (!out || !verts6_d || !tris_d) return fail(RTX_E_ARG, "null argument on rtx_build_device()");
    *out = nullptr;
    rtx_bvh* b = new rtx_bvh;
    (void)hipGetDevice(&b->device);
    int rc = build_impl(verts6_d, nv, tris_d, nt, (hipStream_t)stream, b);
    if (rc != RTX_OK) {
        free_products(b);
        delete b;
        return rc;
    }
    *out = b;
    return RTX_OK;
}

extern "C" int rtx_build(const float* verts6, uint32_t nv, const uint32_t* tris, uint32_t nt, rtx_bvh** out) {
    if (!out || !verts6 || !tris) return fail(RTX_E_ARG, "null argument on rtx_build()");
    if (nt == 0 || nv == 0) return fail(RTX_E_ARG, "no triangles or vertices on rtx_build()");
    float* dv = dalloc<float>(6 * (size_t)nv);
    uint32_t* dt = dalloc<uint32_t>(3 * (size_t)nt);
    int rc = RTX_OK;
    if (!dv || !dt) rc = fail(RTX_E_HIP, "device allocation failed on rtx_build()");
    if (rc == RTX_OK && !(hip_ok(hipMemcpy(dv, verts6, 24ull * nv, hipMemcpyHostToDevice), "copy") &&
                          hip_ok(hipMemcpy(dt, tris, 12ull * nt, hipMemcpyHostToDevice), "copy")))
        rc = RTX_E_HIP;
    if (rc == RTX_OK) rc = rtx_build_device(dv, nv, dt, nt, nullptr, out);
    if (dv) (void)hipFree(dv);
    if (dt) (void)hipFree(dt);
    return rc;
}

extern "C" void rtx_bvh_free(rtx_bvh* b) {
    if (!b) return;
    free_products(b);
    delete b;
}

extern "C" int rtx_debug_scratch_peak(uint64_t* peak_bytes, int reset) {
    if (!peak_bytes) return fail(RTX_E_ARG, "null argument on rtx_debug_scratch_peak()");
    *peak_bytes = scratch_peak(reset != 0);
    return RTX_OK;
}
