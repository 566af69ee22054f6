// CPU oracle: serial restatement of the reference ray tracer's hot path.
//
// TEST INFRASTRUCTURE ONLY -- the checker the HIP product is compared against.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Float semantics follow the reference exactly (compile with -O2 -ffp-contract=off):
//  * vec3 min/max are SSE minps/maxps: min(a,b) = a<b ? a : b  (src/math/vec3.hpp:71-72)
//  * std::min(a,b) = b<a ? b : a, std::max(a,b) = a<b ? b : a  (libstdc++ stl_algobase.h)
//  * dot(a,b) = (ax*bx + ay*by) + (az*bz + aw*bw) with w lanes +0  (vec3.hpp:41-49)
//  * normalize(v) = v * (1/sqrtf(dot(v,v)))                       (vec3.hpp:65-69)
//  * float->uint is x86-64 cvttss2si with a 64-bit destination; float->int saturates to INT_MIN.
//
// Parity pinning of each part is described in the header comment of each section.
#include "rtx_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Constants of src/bvh/bvh_defines.hpp:7-21.
constexpr int      kLbvhThreshold = 64;
constexpr int      kBinsMin = 8;
constexpr int      kBinsMax = 32;
constexpr float    kCTrav = 1.0f;
constexpr float    kCIsect = 1.2f;
constexpr float    kCentroidEps = 1e-6f;
constexpr float    kSpatialEps = 1e-6f;
constexpr uint32_t kMaxLeaf = 16;
constexpr uint32_t kMortonMax = (1u << 21) - 1;

struct V3 { float x, y, z; };

inline float& at(V3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline float at(const V3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 scl(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + 0.0f); }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 normalize(V3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return scl(inv, v); }

inline float sse_min(float a, float b) { return a < b ? a : b; }
inline float sse_max(float a, float b) { return a > b ? a : b; }
inline float std_min(float a, float b) { return (b < a) ? b : a; }
inline float std_max(float a, float b) { return (a < b) ? b : a; }

// x86-64 float->int32 (cvttss2si): out-of-range and NaN give INT_MIN.
inline int f2i(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int)x;
}
// x86-64 float->uint32 as g++ emits it: 64-bit cvttss2si, keep the low 32 bits.
inline uint32_t f2u(float x) {
    if (!(x >= -9223372036854775808.0f && x < 9223372036854775808.0f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)x;
}
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }

// ---------------------------------------------------------------------------
// AABB (src/bvh/AABB.cpp:5-52, 86-171). Empty box = (+inf, -inf).
// ---------------------------------------------------------------------------
struct Box {
    V3 lo{kInf, kInf, kInf};
    V3 hi{-kInf, -kInf, -kInf};
    void extend(V3 p) {
        lo = {sse_min(lo.x, p.x), sse_min(lo.y, p.y), sse_min(lo.z, p.z)};
        hi = {sse_max(hi.x, p.x), sse_max(hi.y, p.y), sse_max(hi.z, p.z)};
    }
    void extend(const Box& b) {
        lo = {sse_min(lo.x, b.lo.x), sse_min(lo.y, b.lo.y), sse_min(lo.z, b.lo.z)};
        hi = {sse_max(hi.x, b.hi.x), sse_max(hi.y, b.hi.y), sse_max(hi.z, b.hi.z)};
    }
    float half_area() const {
        V3 d = sub(hi, lo);
        return (d.x * d.y + d.y * d.z) + d.z * d.x;
    }
    bool valid() const { return lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z; }
};

// Sutherland-Hodgman clip of a triangle against the 6 planes of `r`, then its AABB
// (AABB.cpp:86-171). Clip order: axis x,y,z; per axis the min plane then the max plane.
Box clip_triangle(V3 v0, V3 v1, V3 v2, const Box& r) {
    V3 poly[12], tmp[12];
    int n = 3;
    poly[0] = v0; poly[1] = v1; poly[2] = v2;
    for (int axis = 0; axis < 3; ++axis) {
        for (int side = 0; side < 2; ++side) {
            float plane = side == 0 ? at(r.lo, axis) : at(r.hi, axis);
            int m = 0;
            for (int i = 0; i < n; ++i) {
                int j = (i + 1) % n;
                float pi = at(poly[i], axis), pj = at(poly[j], axis);
                bool ci = side == 0 ? (pi >= plane) : (pi <= plane);
                bool cj = side == 0 ? (pj >= plane) : (pj <= plane);
                if (ci) tmp[m++] = poly[i];
                if (ci != cj) {
                    float t = (plane - pi) / (pj - pi);
                    tmp[m++] = add(poly[i], scl(t, sub(poly[j], poly[i])));
                }
            }
            for (int i = 0; i < m; ++i) poly[i] = tmp[i];
            n = m;
            if (n == 0) return Box{};
        }
    }
    Box b;
    for (int i = 0; i < n; ++i) b.extend(poly[i]);
    return b;
}

// libstdc++'s std::partition for bidirectional iterators (stl_algo.h:1487-1510): a
// two-ended scan that swaps the first false element with the last true one.
template <class Pred>
uint32_t partition_range(uint32_t* arr, uint32_t first, uint32_t last, Pred pred) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (pred(arr[first])) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!pred(arr[last])) --last; else break;
        }
        std::swap(arr[first], arr[last]);
        ++first;
    }
}

// ---------------------------------------------------------------------------
// Morton codes (bvh_morton.cpp:19-105): 21 bits per axis, x in the top lane.
// Pinned by closed-form known answers and a bit-loop interleave in the tests.
// ---------------------------------------------------------------------------
uint64_t interleave3(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t c = 0;
    for (int i = 0; i < 21; ++i) {
        c |= (uint64_t)((x >> i) & 1u) << (3 * i + 2);
        c |= (uint64_t)((y >> i) & 1u) << (3 * i + 1);
        c |= (uint64_t)((z >> i) & 1u) << (3 * i + 0);
    }
    return c;
}

struct Mesh {
    const float* v6;
    uint32_t nv;
    const uint32_t* tri;
    uint32_t nt;
    V3 pos(uint32_t vi) const { return {v6[6 * vi], v6[6 * vi + 1], v6[6 * vi + 2]}; }
    V3 nrm(uint32_t vi) const { return {v6[6 * vi + 3], v6[6 * vi + 4], v6[6 * vi + 5]}; }
    void verts(uint32_t t, V3& a, V3& b, V3& c) const {
        a = pos(tri[3 * t]); b = pos(tri[3 * t + 1]); c = pos(tri[3 * t + 2]);
    }
    V3 centroid(uint32_t t) const {
        V3 a, b, c;
        verts(t, a, b, c);
        return {((a.x + b.x) + c.x) / 3.0f, ((a.y + b.y) + c.y) / 3.0f, ((a.z + b.z) + c.z) / 3.0f};
    }
};

Box scene_box(const Mesh& m) {
    Box b;
    for (uint32_t i = 0; i < m.nt; ++i) {
        V3 a, c, d;
        m.verts(i, a, c, d);
        b.extend(a); b.extend(c); b.extend(d);
    }
    return b;
}

void morton_codes(const Mesh& m, const Box& sb, uint64_t* codes) {
    V3 mn = sb.lo;
    V3 size = sub(sb.hi, sb.lo);
    if (size.x == 0.f) size.x = 1.f;
    if (size.y == 0.f) size.y = 1.f;
    if (size.z == 0.f) size.z = 1.f;
    auto q = [](float c, float lo, float sz) {
        uint32_t v = f2u(((c - lo) * 2097151.0f) / sz);
        return v > kMortonMax ? kMortonMax : v;
    };
    for (uint32_t i = 0; i < m.nt; ++i) {
        V3 c = m.centroid(i);
        codes[i] = interleave3(q(c.x, mn.x, size.x), q(c.y, mn.y, size.y), q(c.z, mn.z, size.z));
    }
}

// Stable LSD radix sort, 8 passes of 8 bits (bvh_morton.cpp:121-156).
void radix_sort(const uint64_t* codes, uint32_t n, uint32_t* idx) {
    std::vector<uint32_t> tmp(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    for (int pass = 0; pass < 8; ++pass) {
        int shift = pass * 8;
        uint32_t cnt[256] = {0};
        for (uint32_t i = 0; i < n; ++i) cnt[(codes[idx[i]] >> shift) & 255u]++;
        uint32_t run = 0;
        for (int b = 0; b < 256; ++b) { uint32_t c = cnt[b]; cnt[b] = run; run += c; }
        for (uint32_t i = 0; i < n; ++i) tmp[cnt[(codes[idx[i]] >> shift) & 255u]++] = idx[i];
        std::memcpy(idx, tmp.data(), sizeof(uint32_t) * n);
    }
}

// ---------------------------------------------------------------------------
// Builder: FIFO top-down build (bvh_bfs.cpp:50-134) with binned SAH + spatial
// splits (bvh_sah.cpp:24-460), then DFS renumbering and triangle flattening
// (bvh_bake.cpp:75-149). Pinned node-by-node against the compiled reference.
// ---------------------------------------------------------------------------
struct Node {
    Box box;
    bool leaf = true;
    uint32_t a = 0, b = 0;   // childID[0..1] or {beginTriIndex, triSize}
};

struct Job { uint32_t node; bool sah; };

struct Builder {
    Mesh m;
    std::vector<uint64_t> code;
    std::vector<uint32_t> sorted;     // s_sortedTris
    std::vector<uint32_t> tindex;     // s_triIndex
    uint32_t tindex_size = 0, sorted_size = 0;
    std::vector<V3> cent;
    std::vector<Node> nodes;
    uint32_t total = 0;
    Box sbox;
    std::deque<Job> queue;
    uint32_t spatial_accepted = 0, budget_rejected = 0;

    uint32_t tri_of(uint32_t pos) const { return sorted[tindex[pos]]; }

    Box range_box(uint32_t begin, uint32_t end) const {
        Box b;
        for (uint32_t i = begin; i < end; ++i) {
            V3 p0, p1, p2;
            m.verts(tri_of(i), p0, p1, p2);
            b.extend(p0); b.extend(p1); b.extend(p2);
        }
        return b;
    }

    void make_children(uint32_t node, uint32_t begin, uint32_t split, uint32_t end,
                       const Box& lb, const Box& rb, bool lsah, bool rsah) {
        uint32_t l = total++, r = total++;
        nodes[node].leaf = false;
        nodes[node].a = l;
        nodes[node].b = r;
        nodes[l] = Node{lb, true, begin, split - begin};
        nodes[r] = Node{rb, true, split, end - split};
        queue.push_back({l, lsah});
        queue.push_back({r, rsah});
    }

    bool sah_split(uint32_t ni);
    void run();
};

bool Builder::sah_split(uint32_t ni) {
    const uint32_t begin = nodes[ni].a;
    const uint32_t n = nodes[ni].b;
    const uint32_t end = begin + n;
    const int nb = clampi((int)n / 4, kBinsMin, kBinsMax);
    const Box pbox = nodes[ni].box;
    const float invA = 1.0f / pbox.half_area();

    float best = kCIsect * (float)n;
    int bAxis = -1, bBin = -1;
    float bCMin = 0.f, bScale = 0.f;

    struct Bin { Box box; uint32_t cnt = 0; };
    Bin bins[kBinsMax];
    Box lbox[kBinsMax];
    uint32_t lcnt[kBinsMax];

    // Object split over centroid bins (bvh_sah.cpp:58-126).
    for (int axis = 0; axis < 3; ++axis) {
        float cmin = std::numeric_limits<float>::max();
        float cmax = -std::numeric_limits<float>::max();
        for (uint32_t i = begin; i < end; ++i) {
            float c = at(cent[tri_of(i)], axis);
            cmin = std_min(cmin, c);
            cmax = std_max(cmax, c);
        }
        if (cmax - cmin < kCentroidEps) continue;
        float scale = (float)nb / (cmax - cmin);
        for (int b = 0; b < nb; ++b) bins[b] = Bin{};
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t t = tri_of(i);
            int b = std::min(f2i((at(cent[t], axis) - cmin) * scale), nb - 1);
            bins[b].cnt++;
            V3 p0, p1, p2;
            m.verts(t, p0, p1, p2);
            bins[b].box.extend(p0); bins[b].box.extend(p1); bins[b].box.extend(p2);
        }
        Box la; uint32_t lc = 0;
        for (int b = 0; b < nb - 1; ++b) {
            la.extend(bins[b].box); lc += bins[b].cnt;
            lbox[b] = la; lcnt[b] = lc;
        }
        Box ra; uint32_t rc = 0;
        for (int b = nb - 1; b >= 1; --b) {
            ra.extend(bins[b].box); rc += bins[b].cnt;
            uint32_t lc2 = lcnt[b - 1];
            if (lc2 == 0 || rc == 0) continue;
            float cost = kCTrav + kCIsect * (lbox[b - 1].half_area() * (float)lc2 + ra.half_area() * (float)rc) * invA;
            if (cost < best) { best = cost; bAxis = axis; bBin = b; bCMin = cmin; bScale = scale; }
        }
    }
    const int oAxis = bAxis, oBin = bBin;
    const float oCMin = bCMin, oScale = bScale;

    // Spatial split over node-extent bins with per-bin clipping (bvh_sah.cpp:157-243).
    struct SBin { Box box; uint32_t in = 0, out = 0; };
    SBin sb[kBinsMax];
    bool spatial = false;
    float sNodeMin = 0.f, sScale = 0.f;
    for (int axis = 0; axis < 3; ++axis) {
        float nmin = at(pbox.lo, axis), nmax = at(pbox.hi, axis);
        if (nmax - nmin < kSpatialEps) continue;
        float scale = (float)nb / (nmax - nmin);
        for (int b = 0; b < nb; ++b) sb[b] = SBin{};
        for (uint32_t i = begin; i < end; ++i) {
            V3 p0, p1, p2;
            m.verts(tri_of(i), p0, p1, p2);
            Box tb; tb.extend(p0); tb.extend(p1); tb.extend(p2);
            int b0 = clampi(f2i((at(tb.lo, axis) - nmin) * scale), 0, nb - 1);
            int b1 = clampi(f2i((at(tb.hi, axis) - nmin) * scale), 0, nb - 1);
            sb[b0].in++;
            sb[b1].out++;
            Box reg = pbox;
            for (int b = b0; b <= b1; ++b) {
                at(reg.lo, axis) = nmin + (float)b / scale;
                at(reg.hi, axis) = nmin + (float)(b + 1) / scale;
                Box cb = clip_triangle(p0, p1, p2, reg);
                if (cb.valid()) sb[b].box.extend(cb);
            }
        }
        Box la; uint32_t lc = 0;
        for (int b = 0; b < nb - 1; ++b) {
            la.extend(sb[b].box); lc += sb[b].in;
            lbox[b] = la; lcnt[b] = lc;
        }
        Box ra; uint32_t rc = 0;
        for (int b = nb - 1; b >= 1; --b) {
            ra.extend(sb[b].box); rc += sb[b].out;
            uint32_t lc2 = lcnt[b - 1];
            if (lc2 == 0 || rc == 0) continue;
            float cost = kCTrav + kCIsect * (lbox[b - 1].half_area() * (float)lc2 + ra.half_area() * (float)rc) * invA;
            if (cost < best) { best = cost; bAxis = axis; bBin = b; spatial = true; sNodeMin = nmin; sScale = scale; }
        }
    }

    if (bAxis == -1) { nodes[ni].leaf = true; return false; }

    uint32_t* ti = tindex.data();
    auto object_split = [&](int axis, int bin, float cmin, float scale) -> bool {
        if (axis == -1) { nodes[ni].leaf = true; return false; }
        uint32_t mid = partition_range(ti, begin, end, [&](uint32_t si) {
            float c = at(cent[sorted[si]], axis);
            return std::min(f2i((c - cmin) * scale), nb - 1) < bin;
        });
        if (mid == begin || mid == end) { nodes[ni].leaf = true; return false; }
        make_children(ni, begin, mid, end, range_box(begin, mid), range_box(mid, end), true, true);
        return true;
    };

    if (!spatial) return object_split(bAxis, bBin, bCMin, bScale);

    // Spatial split: 3-way partition [left-only | straddling | right-only] (bvh_sah.cpp:296-333).
    const float plane = sNodeMin + (float)bBin / sScale;
    const int ax = bAxis;
    auto tri_box = [&](uint32_t si) {
        V3 p0, p1, p2;
        m.verts(sorted[si], p0, p1, p2);
        Box tb; tb.extend(p0); tb.extend(p1); tb.extend(p2);
        return tb;
    };
    uint32_t leftEnd = partition_range(ti, begin, end, [&](uint32_t si) { return at(tri_box(si).hi, ax) <= plane; });
    uint32_t strEnd = partition_range(ti, leftEnd, end, [&](uint32_t si) { return at(tri_box(si).lo, ax) < plane; });
    uint32_t nl = leftEnd - begin, ns = strEnd - leftEnd, nr = end - strEnd;

    if (nl + ns == 0 || ns + nr == 0) return object_split(oAxis, oBin, oCMin, oScale);
    if (tindex_size + nr + ns > (uint32_t)tindex.size() || sorted_size + ns > (uint32_t)sorted.size()) {
        budget_rejected++;
        return object_split(oAxis, oBin, oCMin, oScale);
    }
    spatial_accepted++;
    uint32_t rbegin = tindex_size;
    for (uint32_t k = 0; k < nr; ++k) tindex[tindex_size++] = tindex[strEnd + k];
    for (uint32_t k = 0; k < ns; ++k) {
        uint32_t old = tindex[leftEnd + k];
        sorted[sorted_size] = sorted[old];
        tindex[tindex_size++] = sorted_size++;
    }
    uint32_t rsize = nr + ns, lsize = nl + ns;
    Box lclip = pbox, rclip = pbox;
    at(lclip.hi, ax) = plane;
    at(rclip.lo, ax) = plane;
    auto clipped_range = [&](uint32_t b0, uint32_t cnt, const Box& cb) {
        Box out;
        for (uint32_t k = b0; k < b0 + cnt; ++k) {
            V3 p0, p1, p2;
            m.verts(tri_of(k), p0, p1, p2);
            Box c = clip_triangle(p0, p1, p2, cb);
            if (c.valid()) out.extend(c);
        }
        return out;
    };
    Box lb = clipped_range(begin, lsize, lclip);
    Box rb = clipped_range(rbegin, rsize, rclip);
    uint32_t l = total++, r = total++;
    nodes[ni].leaf = false; nodes[ni].a = l; nodes[ni].b = r;
    nodes[l] = Node{lb, true, begin, lsize};
    nodes[r] = Node{rb, true, rbegin, rsize};
    queue.push_back({l, true});
    queue.push_back({r, true});
    return true;
}

void Builder::run() {
    const uint32_t N = m.nt;
    const uint32_t cap = (uint32_t)((float)N * 1.5f);
    code.resize(N);
    sbox = scene_box(m);
    morton_codes(m, sbox, code.data());
    sorted.assign(cap, 0);
    tindex.assign(cap, 0);
    {
        std::vector<uint32_t> idx(N);
        radix_sort(code.data(), N, idx.data());
        for (uint32_t i = 0; i < N; ++i) { sorted[i] = idx[i]; tindex[i] = i; }
    }
    tindex_size = N;
    sorted_size = N;
    cent.resize(N);
    for (uint32_t i = 0; i < N; ++i) cent[i] = m.centroid(i);

    nodes.assign((size_t)cap * 2 - 1 > 0 ? (size_t)cap * 2 - 1 : 1, Node{});
    total = 1;
    nodes[0] = Node{sbox, true, 0, N};
    queue.push_back({0, N <= (uint32_t)kLbvhThreshold});
    while (!queue.empty()) {
        Job j = queue.front();
        queue.pop_front();
        Node& nd = nodes[j.node];
        if (nd.b <= kMaxLeaf) { nd.leaf = true; continue; }
        if (j.sah) { sah_split(j.node); continue; }
        uint32_t begin = nd.a, cnt = nd.b, end = begin + cnt;
        uint64_t x = code[sorted[begin]] ^ code[sorted[end - 1]];
        if (x == 0) { sah_split(j.node); continue; }
        uint64_t mask = (uint64_t)1 << (63 - __builtin_clzll(x));
        uint32_t split = kInvalid;
        for (uint32_t i = 0; i + 1 < cnt; ++i) {
            if ((code[sorted[begin + i]] & mask) != (code[sorted[begin + i + 1]] & mask)) { split = begin + i + 1; break; }
        }
        Box lb = range_box(begin, split), rb = range_box(split, end);
        make_children(j.node, begin, split, end, lb, rb,
                      split - begin <= (uint32_t)kLbvhThreshold, end - split <= (uint32_t)kLbvhThreshold);
    }
}

}  // namespace

struct orc_tree {
    std::vector<orc_node48> nodes;      // DFS order
    std::vector<orc_node48> bfs_nodes;  // builder order
    std::vector<uint32_t> tris;         // 3 per reordered triangle
    std::vector<uint32_t> prims;
    std::vector<uint32_t> sorted;
    uint32_t tindex_size = 0, sorted_size = 0, depth = 0, spatial = 0, rejected = 0;
};

static orc_node48 to48(const Node& n) {
    orc_node48 o;
    std::memset(&o, 0, sizeof(o));
    o.bmin[0] = n.box.lo.x; o.bmin[1] = n.box.lo.y; o.bmin[2] = n.box.lo.z;
    o.bmax[0] = n.box.hi.x; o.bmax[1] = n.box.hi.y; o.bmax[2] = n.box.hi.z;
    o.is_leaf = n.leaf ? 1 : 0;
    o.a = n.a; o.b = n.b;
    return o;
}

extern "C" {

int orc_morton(const float* verts6, uint32_t nv, const uint32_t* tris, uint32_t nt,
               uint64_t* codes_out, float* box6) {
    if (nt == 0) return -1;
    Mesh m{verts6, nv, tris, nt};
    Box b = scene_box(m);
    morton_codes(m, b, codes_out);
    if (box6) { box6[0] = b.lo.x; box6[1] = b.lo.y; box6[2] = b.lo.z; box6[3] = b.hi.x; box6[4] = b.hi.y; box6[5] = b.hi.z; }
    return 0;
}

int orc_sort(const uint64_t* codes, uint32_t nt, uint32_t* sorted_out) {
    radix_sort(codes, nt, sorted_out);
    return 0;
}

int orc_build(const float* verts6, uint32_t nv, const uint32_t* tris, uint32_t nt, orc_tree** out) {
    if (nt == 0 || nv == 0) return -1;
    for (uint32_t i = 0; i < 3 * nt; ++i) if (tris[i] >= nv) return -2;
    Builder B;
    B.m = Mesh{verts6, nv, tris, nt};
    B.run();
    orc_tree* t = new orc_tree;
    t->sorted.assign(B.sorted.begin(), B.sorted.begin() + nt);
    t->tindex_size = B.tindex_size;
    t->sorted_size = B.sorted_size;
    t->spatial = B.spatial_accepted;
    t->rejected = B.budget_rejected;
    t->bfs_nodes.resize(B.total);
    for (uint32_t i = 0; i < B.total; ++i) t->bfs_nodes[i] = to48(B.nodes[i]);
    // DFS pre-order renumbering, left child first (bvh_bake.cpp:75-118).
    std::vector<uint32_t> remap(B.total, kInvalid), order;
    order.reserve(B.total);
    std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 0u}};
    while (!st.empty()) {
        auto [o, d] = st.back();
        st.pop_back();
        remap[o] = (uint32_t)order.size();
        order.push_back(o);
        t->depth = std::max(t->depth, d);
        if (!B.nodes[o].leaf) { st.push_back({B.nodes[o].b, d + 1}); st.push_back({B.nodes[o].a, d + 1}); }
    }
    t->nodes.resize(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
        Node n = B.nodes[order[i]];
        if (!n.leaf) { n.a = remap[n.a]; n.b = remap[n.b]; }
        t->nodes[i] = to48(n);
    }
    // Leaves in DFS order own consecutive triangle runs (bvh_bake.cpp:120-149).
    for (auto& n : t->nodes) {
        if (!n.is_leaf) continue;
        uint32_t old = n.a;
        n.a = (uint32_t)t->prims.size();
        for (uint32_t k = 0; k < n.b; ++k) {
            uint32_t tri = B.sorted[B.tindex[old + k]];
            t->prims.push_back(tri);
            t->tris.push_back(tris[3 * tri]); t->tris.push_back(tris[3 * tri + 1]); t->tris.push_back(tris[3 * tri + 2]);
        }
    }
    *out = t;
    return 0;
}

void orc_tree_free(orc_tree* t) { delete t; }
void orc_tree_counts(const orc_tree* t, uint32_t* c) {
    c[0] = (uint32_t)t->nodes.size(); c[1] = (uint32_t)t->prims.size(); c[2] = t->tindex_size;
    c[3] = t->sorted_size; c[4] = t->depth; c[5] = t->spatial; c[6] = t->rejected;
}
void orc_tree_nodes(const orc_tree* t, orc_node48* out) { std::memcpy(out, t->nodes.data(), t->nodes.size() * sizeof(orc_node48)); }
void orc_tree_bfs_nodes(const orc_tree* t, orc_node48* out) { std::memcpy(out, t->bfs_nodes.data(), t->bfs_nodes.size() * sizeof(orc_node48)); }
void orc_tree_tris(const orc_tree* t, uint32_t* out) { std::memcpy(out, t->tris.data(), t->tris.size() * sizeof(uint32_t)); }
void orc_tree_prims(const orc_tree* t, uint32_t* out) { std::memcpy(out, t->prims.data(), t->prims.size() * sizeof(uint32_t)); }
void orc_tree_sorted(const orc_tree* t, uint32_t* out) { std::memcpy(out, t->sorted.data(), t->sorted.size() * sizeof(uint32_t)); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Traversal (bvh_traverse.cpp:22-178). Pinned against the compiled reference
// trace() and against brute force over all triangles.
// ---------------------------------------------------------------------------
namespace {

struct Tracer {
    const orc_node48* nodes;
    const uint32_t* tris;
    Mesh m;
};

struct RayF { V3 o, d; float tmin, tmax; };

inline float ray_box(const RayF& r, const orc_node48& n, float tmax) {
    float ix = 1.f / r.d.x, iy = 1.f / r.d.y, iz = 1.f / r.d.z;
    float t0x = (n.bmin[0] - r.o.x) * ix, t1x = (n.bmax[0] - r.o.x) * ix;
    float t0y = (n.bmin[1] - r.o.y) * iy, t1y = (n.bmax[1] - r.o.y) * iy;
    float t0z = (n.bmin[2] - r.o.z) * iz, t1z = (n.bmax[2] - r.o.z) * iz;
    float tn = std_max(std_max(std_min(t0x, t1x), std_min(t0y, t1y)), std_max(std_min(t0z, t1z), r.tmin));
    float tf = std_min(std_min(std_max(t0x, t1x), std_max(t0y, t1y)), std_min(std_max(t0z, t1z), tmax));
    return tn <= tf ? tn : kInf;
}

struct TH { float t, u, v; };

inline TH ray_tri(const RayF& r, V3 v0, V3 v1, V3 v2) {
    const TH miss{kInf, 0.f, 0.f};
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 h = cross(r.d, e2);
    float a = dot(e1, h);
    if (a >= 0) return miss;
    float f = 1.f / a;
    V3 s = sub(r.o, v0);
    float u = f * dot(s, h);
    if (u < 0.f || u > 1.f) return miss;
    V3 q = cross(s, e1);
    float v = f * dot(r.d, q);
    if (v < 0.f || u + v > 1.f) return miss;
    float t = f * dot(e2, q);
    if (t < r.tmin || t > r.tmax) return miss;
    return {t, u, v};
}

struct HitF { bool hit = false; float t = 0, u = 0, v = 0; uint32_t tri = 0; V3 n{0, 0, 0}; };

HitF traverse(const Tracer& T, const RayF& r, bool any, uint32_t* cnt) {
    HitF best;
    float bestT = r.tmax;
    uint32_t stack[256];
    int top = 0;
    stack[top++] = 0;
    uint32_t ni_visits = 0, ntests = 0;
    while (top > 0) {
        uint32_t ni = stack[--top];
        const orc_node48& nd = T.nodes[ni];
        float tb = ray_box(r, nd, bestT);
        if (tb >= bestT) continue;
        if (nd.is_leaf) {
            for (uint32_t k = 0; k < nd.b; ++k) {
                uint32_t ti = nd.a + k;
                uint32_t i0 = T.tris[3 * ti], i1 = T.tris[3 * ti + 1], i2 = T.tris[3 * ti + 2];
                ++ntests;
                TH th = ray_tri(r, T.m.pos(i0), T.m.pos(i1), T.m.pos(i2));
                if (th.t < bestT) {
                    bestT = th.t;
                    best.hit = true; best.t = th.t; best.u = th.u; best.v = th.v; best.tri = ti;
                    float w = 1.f - th.u - th.v;
                    best.n = normalize(add(add(scl(w, T.m.nrm(i0)), scl(th.u, T.m.nrm(i1))), scl(th.v, T.m.nrm(i2))));
                    if (any) { top = 0; break; }
                }
            }
        } else {
            ++ni_visits;
            uint32_t l = nd.a, rr = nd.b;
            float tl = ray_box(r, T.nodes[l], bestT), tr = ray_box(r, T.nodes[rr], bestT);
            if (tl <= tr) {
                if (tr < bestT) stack[top++] = rr;
                if (tl < bestT) stack[top++] = l;
            } else {
                if (tl < bestT) stack[top++] = l;
                if (tr < bestT) stack[top++] = rr;
            }
        }
    }
    if (cnt) { cnt[0] = ni_visits; cnt[1] = ntests; }
    return best;
}

template <class F>
void parallel_for(size_t n, int nthreads, F&& f) {
    if (nthreads <= 1 || n < 1024) { f(0, n); return; }
    std::vector<std::thread> th;
    size_t chunk = (n + nthreads - 1) / nthreads;
    for (int i = 0; i < nthreads; ++i) {
        size_t b = i * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&, b, e] { f(b, e); });
    }
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" int orc_trace(const orc_node48* nodes, uint32_t nn, const uint32_t* tris, uint32_t ntri,
                         const float* verts6, uint32_t nv, const orc_ray* rays, size_t n,
                         orc_hit* hits, float* normals, uint32_t* counts, int any_hit, int nthreads) {
    if (nn == 0) return -1;
    (void)ntri;
    Tracer T{nodes, tris, Mesh{verts6, nv, nullptr, 0}};
    parallel_for(n, nthreads, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const orc_ray& q = rays[i];
            RayF r{{q.ox, q.oy, q.oz}, {q.dx, q.dy, q.dz}, q.tmin, q.tmax};
            HitF h = traverse(T, r, any_hit != 0, counts ? counts + 2 * i : nullptr);
            hits[i] = orc_hit{h.t, h.u, h.v, h.hit ? h.tri : 0xFFFFFFFFu};
            if (normals) { normals[3 * i] = h.n.x; normals[3 * i + 1] = h.n.y; normals[3 * i + 2] = h.n.z; }
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// Sample table and the project's sin/cos: both are specifications shared by the
// oracle and the GPU renderer; pinned against numpy within 3e-7 in the tests.
// ---------------------------------------------------------------------------
static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

extern "C" float orc_sample(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    uint32_t h = mix32(seed ^ 0x9e3779b9u);
    h = mix32(h ^ pixel);
    h = mix32(h ^ sample);
    h = mix32(h ^ dim);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

extern "C" void orc_sincos(float x, float* s, float* c) {
    // Cody-Waite reduction by pi/2 with a 3-part constant, Taylor polynomials on [-pi/4, pi/4].
    float kf = floorf(x * 0.636619747f + 0.5f);
    float r = ((x - kf * 1.57080078125f) - kf * -4.453584551811218e-06f) - kf * -8.705516307827565e-10f;
    float r2 = r * r;
    float sp = r + (r * r2) * (-1.66666672e-1f + r2 * (8.33333377e-3f + r2 * (-1.98412701e-4f + r2 * 2.75573188e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666679e-2f + r2 * (-1.38888892e-3f + r2 * (2.48015876e-5f + r2 * -2.75573188e-7f))));
    int q = ((int)kf) & 3;
    float ss = (q == 0) ? sp : (q == 1) ? cp : (q == 2) ? -sp : -cp;
    float cc = (q == 0) ? cp : (q == 1) ? -sp : (q == 2) ? -cp : sp;
    *s = ss; *c = cc;
}

// ---------------------------------------------------------------------------
// Path-trace loop (main.cpp:69-83, 123-141, 145-204, 339-365) driven by the
// sample table. Extensions beyond the reference (documented in the design):
// per-triangle albedo/emission, max_bounce/rr_bounce parameters, constant env.
// ---------------------------------------------------------------------------
namespace {

template <class R>
struct Vec3T { R x, y, z; };

struct SceneRef {
    Tracer T;
    const uint32_t* prims;
    const float* albedo;
    const float* emission;
};

inline RayF primary_ray(const orc_camera* c, int px, int py, int W, int H, float jx, float jy) {
    float aspect = (float)W / (float)H;
    float u = (((float)px + jx) / (float)W * 2.f - 1.f) * aspect;
    float v = ((float)py + jy) / (float)H * 2.f - 1.f;
    V3 fwd{c->fwd[0], c->fwd[1], c->fwd[2]}, rt{c->right[0], c->right[1], c->right[2]}, up{c->up[0], c->up[1], c->up[2]};
    V3 d = normalize(add(add(fwd, scl(u * c->tan_half_fov_y, rt)), scl(v * c->tan_half_fov_y, up)));
    return RayF{{c->pos[0], c->pos[1], c->pos[2]}, d, 1e-3f, kInf};
}

// ORC_MUTANT (never set in liboracle.so): planted mistakes compiled into separate
// test-only libraries, so the tests can show that their closed-form checks catch each one.
//   1 uniform instead of cosine-weighted hemisphere   2 normal in the x slot of toWorld
//   3 roulette without throughput /= p                4 roulette probability = channel mean
//   5 tangent not normalized
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

// sampleCosineHemisphere (main.cpp:123-131)
inline V3 cos_hemisphere(float u1, float u2) {
    float r = sqrtf(u1);
    float theta = 6.28318548f * u2;
    float s, c;
    orc_sincos(theta, &s, &c);
#if ORC_MUTANT == 1
    r = sqrtf(std_max(0.f, 1.f - u1 * u1));
    return {r * c, u1, r * s};
#endif
    return {r * c, sqrtf(std_max(0.f, 1.f - u1)), r * s};
}

// toWorld (main.cpp:133-141)
inline V3 to_world(V3 l, V3 n) {
    V3 up = fabsf(n.x) < 0.9f ? V3{1.f, 0.f, 0.f} : V3{0.f, 1.f, 0.f};
#if ORC_MUTANT == 5
    V3 t = cross(n, up);
#else
    V3 t = normalize(cross(n, up));
#endif
    V3 b = cross(n, t);
#if ORC_MUTANT == 2
    return add(add(scl(l.x, n), scl(l.y, t)), scl(l.z, b));
#endif
    return add(add(scl(l.x, t), scl(l.y, n)), scl(l.z, b));
}

// atan2/acos of the sample-table renderer: octant reduction and a degree-7 polynomial in
// a^2 for atan(a)/a on [0,1] (the same specification the GPU renderer implements).
inline float atan2_spec(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    const float a = mx == 0.f ? 0.f : mn / mx;
    const float s = a * a;
    const float c[8] = {0.9999998211860657f, -0.33331647515296936f, 0.19965273141860962f, -0.13995960354804993f,
                        0.09853174537420273f, -0.058659493923187256f, 0.023694388568401337f, -0.004545152187347412f};
    float p = c[7];
    for (int k = 6; k >= 0; --k) p = p * s + c[k];
    float r = p * a;
    if (ay > ax) r = 1.57079637f - r;
    if (x < 0.f) r = 3.14159274f - r;
    if (y < 0.f) r = -r;
    return r;
}

extern "C" float orc_atan2(float y, float x) { return atan2_spec(y, x); }

// IBL::sample (main.cpp:107-117) with the atan2/acos above.
inline V3 ibl_sample(const orc_render_params* p, V3 d) {
    float phi = atan2_spec(d.z, d.x);
    float cy = d.y < -1.f ? -1.f : (1.f < d.y ? 1.f : d.y);
    float theta = atan2_spec(sqrtf((1.f - cy) * (1.f + cy)), cy);
    float u = (phi + 3.14159274f) / 6.28318548f;
    float v = theta / 3.14159274f;
    int px = std::min(f2i(u * (float)p->ibl_w), p->ibl_w - 1);
    int py = std::min(f2i(v * (float)p->ibl_h), p->ibl_h - 1);
    px = std::max(px, 0);
    py = std::max(py, 0);
    const float* t = p->ibl + 3 * ((size_t)py * p->ibl_w + px);
    return {t[0], t[1], t[2]};
}

template <class R>
inline Vec3T<R> env_color(const orc_render_params* p, V3 d) {
    if (p->env_mode == 1) return {(R)p->env_color[0], (R)p->env_color[1], (R)p->env_color[2]};
    if (p->env_mode == 3) {
        // constant sky plus a sun disk (a distant light seen by the rays that reach it)
        const float c = dot(d, V3{p->sun_dir[0], p->sun_dir[1], p->sun_dir[2]});
        float r = p->env_color[0], g = p->env_color[1], b = p->env_color[2];
        if (c >= p->sun_cos) { r = r + p->sun_rad[0]; g = g + p->sun_rad[1]; b = b + p->sun_rad[2]; }
        return {(R)r, (R)g, (R)b};
    }
    if (p->env_mode == 2) { V3 e = ibl_sample(p, d); return {(R)e.x, (R)e.y, (R)e.z}; }
    float t = d.y * 0.5f + 0.5f;
    float omt = 1.f - t;
    return {(R)(1.0f * omt + 0.5f * t), (R)(1.0f * omt + 0.7f * t), (R)(1.0f * omt + 1.0f * t)};
}

// Texel of a hit for texture albedo: uv interpolated like the shading normal, then the
// nearest texel (x86 truncation, clamped to the texture).
inline uint32_t texel_at(const orc_render_params* p, const uint32_t* t3, float u, float v) {
    const float w = 1.f - u - v;
    const float* a = p->uv + 2 * (size_t)t3[0];
    const float* b = p->uv + 2 * (size_t)t3[1];
    const float* c = p->uv + 2 * (size_t)t3[2];
    const float U = (w * a[0] + u * b[0]) + v * c[0];
    const float V = (w * a[1] + u * b[1]) + v * c[1];
    int px = std::min(f2i(U * (float)p->tex_w), p->tex_w - 1);
    int py = std::min(f2i(V * (float)p->tex_h), p->tex_h - 1);
    px = std::max(px, 0);
    py = std::max(py, 0);
    return (uint32_t)py * (uint32_t)p->tex_w + (uint32_t)px;
}

// One path. R = float reproduces the renderer; R = double is the finite-difference
// variant (same float32 geometry, float64 throughput and radiance).
template <class R>
Vec3T<R> path(const SceneRef& S, const orc_render_params* p, RayF ray, uint32_t pixel, uint32_t s) {
    Vec3T<R> tp{1, 1, 1}, L{0, 0, 0};
    for (int bounce = 0; bounce <= p->max_bounce; ++bounce) {
        HitF h = traverse(S.T, ray, false, nullptr);
        if (!h.hit) {
            Vec3T<R> e = env_color<R>(p, ray.d);
            L = {L.x + tp.x * e.x, L.y + tp.y * e.y, L.z + tp.z * e.z};
            break;
        }
        uint32_t prim = S.prims[h.tri];
        if (S.emission) {
            const float* e = S.emission + 3 * prim;
            L = {L.x + tp.x * (R)e[0], L.y + tp.y * (R)e[1], L.z + tp.z * (R)e[2]};
        }
        R ax = (R)0.8f, ay = (R)0.8f, az = (R)0.8f;
        if (p->tex) {
            const uint32_t* t3 = S.T.tris + 3 * (size_t)h.tri;
            const float* a = p->tex + 3 * (size_t)texel_at(p, t3, h.u, h.v);
            ax = a[0]; ay = a[1]; az = a[2];
        } else if (S.albedo) {
            ax = S.albedo[3 * prim]; ay = S.albedo[3 * prim + 1]; az = S.albedo[3 * prim + 2];
        }
        float u1 = orc_sample(p->seed, pixel, s, 2 + 3 * bounce);
        float u2 = orc_sample(p->seed, pixel, s, 3 + 3 * bounce);
        V3 wd = to_world(cos_hemisphere(u1, u2), h.n);
        tp = {tp.x * ax, tp.y * ay, tp.z * az};
        // Russian roulette (main.cpp:188-193)
        if (bounce >= p->rr_bounce) {
            R pr = tp.x;
            if (pr < tp.y) pr = tp.y;
            if (pr < tp.z) pr = tp.z;
#if ORC_MUTANT == 4
            pr = (tp.x + tp.y + tp.z) / (R)3;
#endif
            if ((R)orc_sample(p->seed, pixel, s, 4 + 3 * bounce) > pr) break;
#if ORC_MUTANT != 3
            tp = {tp.x / pr, tp.y / pr, tp.z / pr};
#endif
        }
        ray.o = add(ray.o, scl(h.t, ray.d));
        ray.d = wd;
        ray.tmin = 1e-4f;
        ray.tmax = kInf;
    }
    return L;
}

}  // namespace

extern "C" int orc_render(const orc_node48* nodes, uint32_t nn, const uint32_t* tris, const uint32_t* prims,
                          uint32_t ntri, const float* verts6, uint32_t nv, const float* albedo,
                          const float* emission, const orc_camera* cam, const orc_render_params* p,
                          float* image, int nthreads) {
    if (nn == 0 || p->spp <= 0) return -1;
    (void)ntri;
    SceneRef S{Tracer{nodes, tris, Mesh{verts6, nv, nullptr, 0}}, prims, albedo, emission};
    const int W = p->width, H = p->height;
    parallel_for((size_t)W * H, nthreads, [&](size_t b, size_t e) {
        for (size_t pix = b; pix < e; ++pix) {
            int x = (int)(pix % W), y = (int)(pix / W);
            float ax = 0.f, ay = 0.f, az = 0.f;
            for (int s = 0; s < p->spp; ++s) {
                float jx = orc_sample(p->seed, (uint32_t)pix, s, 0);
                float jy = orc_sample(p->seed, (uint32_t)pix, s, 1);
                Vec3T<float> L = path<float>(S, p, primary_ray(cam, x, y, W, H, jx, jy), (uint32_t)pix, s);
                ax += L.x; ay += L.y; az += L.z;
            }
            float fs = (float)p->spp;
            image[3 * pix] = ax / fs; image[3 * pix + 1] = ay / fs; image[3 * pix + 2] = az / fs;
        }
    });
    return 0;
}

extern "C" double orc_render_loss_f64(const orc_node48* nodes, uint32_t nn, const uint32_t* tris,
                                      const uint32_t* prims, uint32_t ntri, const float* verts6, uint32_t nv,
                                      const float* albedo, const float* emission, const orc_camera* cam,
                                      const orc_render_params* p, const float* target, int nthreads) {
    if (nn == 0 || p->spp <= 0) return -1.0;
    (void)ntri;
    SceneRef S{Tracer{nodes, tris, Mesh{verts6, nv, nullptr, 0}}, prims, albedo, emission};
    const int W = p->width, H = p->height;
    int nt = nthreads < 1 ? 1 : nthreads;
    std::vector<double> part(nt, 0.0);
    size_t npix = (size_t)W * H;
    size_t chunk = (npix + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            double acc = 0.0;
            for (size_t pix = t * chunk; pix < std::min(npix, (t + 1) * chunk); ++pix) {
                int x = (int)(pix % W), y = (int)(pix / W);
                double sx = 0, sy = 0, sz = 0;
                for (int s = 0; s < p->spp; ++s) {
                    float jx = orc_sample(p->seed, (uint32_t)pix, s, 0);
                    float jy = orc_sample(p->seed, (uint32_t)pix, s, 1);
                    Vec3T<double> L = path<double>(S, p, primary_ray(cam, x, y, W, H, jx, jy), (uint32_t)pix, s);
                    sx += L.x; sy += L.y; sz += L.z;
                }
                double dx = sx / p->spp - target[3 * pix], dy = sy / p->spp - target[3 * pix + 1],
                       dz = sz / p->spp - target[3 * pix + 2];
                acc += dx * dx + dy * dy + dz * dz;
            }
            part[t] = acc;
        });
    }
    for (auto& x : th) x.join();
    double sum = 0.0;
    for (double v : part) sum += v;
    return sum;
}
