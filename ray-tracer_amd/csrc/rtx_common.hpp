// Shared device/host definitions for librtx: data layouts in HBM and the exact float
// semantics of the reference (minsuk0519/ray-tracer) that parity depends on.
//
// All sources are compiled with -ffp-contract=off (no FMA contraction) and keep HIP's
// correctly rounded fp32 divide and sqrt.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RTX_HD __host__ __device__ __forceinline__

namespace rtx {

constexpr float kInf = __builtin_huge_valf();
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;

// Builder constants (src/bvh/bvh_defines.hpp:7-21).
constexpr int      kLbvhThreshold = 64;
constexpr int      kBinsMin = 8;
constexpr int      kBinsMax = 32;
constexpr float    kCTrav = 1.0f;
constexpr float    kCIsect = 1.2f;
constexpr float    kCentroidEps = 1e-6f;
constexpr float    kSpatialEps = 1e-6f;
constexpr uint32_t kMaxLeaf = 16;
constexpr uint32_t kMortonMax = (1u << 21) - 1;

struct f3 { float x, y, z; };

RTX_HD float comp(const f3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
RTX_HD void set_comp(f3& v, int a, float s) { if (a == 0) v.x = s; else if (a == 1) v.y = s; else v.z = s; }
RTX_HD f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RTX_HD f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RTX_HD f3 scl(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
// vec3 dot sums lanes as (x+y)+(z+w) with a +0 w lane (src/math/vec3.hpp:41-49).
RTX_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + 0.0f); }
RTX_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// 1.0f / a, bit for bit. On the GPU, v_rcp_f32 followed by one FMA Newton step
// y + y * (1 - a * y) is the correctly rounded reciprocal for every 2^-126 <= |a| < 2^126
// (all 2^32 inputs were compared with the IEEE division on gfx950): 3 instructions instead
// of the 11 of the scaled division sequence. Zero, subnormal, huge, infinite and NaN a take
// the division.
RTX_HD float recip(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float m = __builtin_fabsf(a);
    if (m >= 0x1p-126f && m < 0x1p126f) {
        const float y = __builtin_amdgcn_rcpf(a);
        return __builtin_fmaf(__builtin_fmaf(-a, y, 1.0f), y, y);
    }
#endif
    return 1.0f / a;
}

RTX_HD f3 normalize(f3 v) { float inv = recip(sqrtf(dot(v, v))); return scl(inv, v); }

// SSE minps/maxps (vec3 min/max, vec3.hpp:71-72) and libstdc++ std::min/std::max.
RTX_HD float sse_min(float a, float b) { return a < b ? a : b; }
RTX_HD float sse_max(float a, float b) { return a > b ? a : b; }
RTX_HD float std_min(float a, float b) { return (b < a) ? b : a; }
RTX_HD float std_max(float a, float b) { return (a < b) ? b : a; }

// x86-64 conversions as the reference binary performs them: float->int is cvttss2si
// (INT_MIN when out of range or NaN); float->uint32 is the low half of a 64-bit cvttss2si.
RTX_HD int f2i_x86(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return (int)0x80000000;
    return (int)x;
}
RTX_HD uint32_t f2u_x86(float x) {
    if (!(x >= -9223372036854775808.0f && x < 9223372036854775808.0f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)x;
}
RTX_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }

// AABB with the reference's empty box (+inf,-inf) and SSE min/max extension
// (src/bvh/AABB.cpp:5-52). Extension keeps, per component, the last of equal extrema,
// so an ordered tree reduction combine(earlier, later) equals the serial fold.
struct Box {
    f3 lo, hi;
    RTX_HD static Box empty() { return Box{{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}}; }
    RTX_HD void extend(f3 p) {
        lo = {sse_min(lo.x, p.x), sse_min(lo.y, p.y), sse_min(lo.z, p.z)};
        hi = {sse_max(hi.x, p.x), sse_max(hi.y, p.y), sse_max(hi.z, p.z)};
    }
    RTX_HD void extend(const Box& b) {
        lo = {sse_min(lo.x, b.lo.x), sse_min(lo.y, b.lo.y), sse_min(lo.z, b.lo.z)};
        hi = {sse_max(hi.x, b.hi.x), sse_max(hi.y, b.hi.y), sse_max(hi.z, b.hi.z)};
    }
    RTX_HD float half_area() const {
        f3 d = sub(hi, lo);
        return (d.x * d.y + d.y * d.z) + d.z * d.x;
    }
    RTX_HD bool valid() const { return lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z; }
};

// ---------------------------------------------------------------------------
// HBM layouts
// ---------------------------------------------------------------------------

// Reference BVHNode (src/bvh/bvh.hpp:15-30), 48 bytes, used for export and .bvh files.
struct Node48 {
    float bmin[4];
    float bmax[4];
    uint8_t is_leaf;
    uint8_t pad0[3];
    uint32_t a, b;
    uint32_t pad1;
};
static_assert(sizeof(Node48) == 48, "Node48 layout");

// Traversal node: one 64-byte line per internal node holding both children's boxes.
// A child reference is an internal node index, or for a leaf (kLeafBit set) either
//  - kInlineLeaf | count << 25 | begin, when count < 32 and begin < 2^25: the leaf's
//    triangle range itself, so reaching the leaf needs no further load; or
//  - the leaf's index into the leaf table (bit 30 clear, index < 2^30).
struct alignas(64) GNode {
    float4 l0;   // lmin.x lmin.y lmin.z lmax.x
    float4 l1;   // lmax.y lmax.z rmin.x rmin.y
    float4 r1;   // rmin.z rmax.x rmax.y rmax.z
    uint32_t lref, rref, pad0, pad1;
};
static_assert(sizeof(GNode) == 64, "GNode layout");

// Wide traversal node (trace entry points, tuning trace_tree 1): up to four children of a
// binary node collapsed into one 128-byte line. Child boxes are the binary tree's own boxes
// (exact floats), stored per coordinate so one float4 load brings a coordinate of all four;
// refs use the binary child-reference encoding (internal node index, or a leaf reference),
// kInvalid marks an empty slot (slots fill from 0, in the binary tree's left-to-right order).
struct alignas(128) WNode {
    float4 lx, ly, lz;   // box minima of children 0..3
    float4 hx, hy, hz;   // box maxima
    uint4 ref;
    uint4 pad;
};
static_assert(sizeof(WNode) == 128, "WNode layout");

// Traversal triangle: 36 bytes, v0 and the two edges precomputed (bit-identical to
// computing v1-v0, v2-v0 at trace time): v0.xyz e1.xyz e2.xyz. Packed at a 36-byte
// stride (4-byte aligned): the cooperative leaf phase loads a leaf's triangles with
// neighbouring lanes, so a round of 64 reads 18 cache lines instead of 24 at 48 bytes.
struct GTri {
    float v[9];
};
static_assert(sizeof(GTri) == 36 && alignof(GTri) == 4, "GTri layout");

// 16-byte vectors at 4-byte alignment: gfx950 loads and stores dwordx4 at any dword
// address, which the packed triangles need.
typedef float gtri_f4 __attribute__((ext_vector_type(4), aligned(4)));
typedef __attribute__((address_space(1))) gtri_f4 gtri_f4_global;
typedef __attribute__((address_space(1))) float float_global;

__device__ __forceinline__ void store_gtri(GTri* p, f3 v0, f3 e1, f3 e2) {
    gtri_f4_global* q = (gtri_f4_global*)p->v;
    q[0] = gtri_f4{v0.x, v0.y, v0.z, e1.x};
    q[1] = gtri_f4{e1.y, e1.z, e2.x, e2.y};
    ((float_global*)p->v)[8] = e2.z;
}

struct Leaf { uint32_t begin, count; };

constexpr uint32_t kInlineLeaf = 0x40000000u;
// The one inline encoding equal to kInvalid (31 triangles from 2^25 - 1), which traversal
// keeps for a finished ray, takes the leaf-table form instead.
RTX_HD uint32_t leaf_ref(uint32_t leaf_index, uint32_t begin, uint32_t count) {
    return (count < 32u && begin < (1u << 25) && !(count == 31u && begin == (1u << 25) - 1u))
               ? (kLeafBit | kInlineLeaf | (count << 25) | begin)
               : (kLeafBit | leaf_index);
}

struct Ray32 { float ox, oy, oz, tmin, dx, dy, dz, tmax; };
struct Hit16 { float t, u, v; uint32_t tri; };

// ---------------------------------------------------------------------------
// Slab and triangle tests (src/bvh/bvh_traverse.cpp:22-87)
// ---------------------------------------------------------------------------
RTX_HD float slab(float ox, float oy, float oz, float ix, float iy, float iz, float tmin,
                  float x0, float y0, float z0, float x1, float y1, float z1, float tmax) {
    float t0x = (x0 - ox) * ix, t1x = (x1 - ox) * ix;
    float t0y = (y0 - oy) * iy, t1y = (y1 - oy) * iy;
    float t0z = (z0 - oz) * iz, t1z = (z1 - oz) * iz;
    float tn = std_max(std_max(std_min(t0x, t1x), std_min(t0y, t1y)), std_max(std_min(t0z, t1z), tmin));
    float tf = std_min(std_min(std_max(t0x, t1x), std_max(t0y, t1y)), std_min(std_max(t0z, t1z), tmax));
    return tn <= tf ? tn : kInf;
}

// slab() with single-instruction min/max. The two agree whenever no t is NaN: every min
// and max then returns the same value (a -0 against +0 tie may flip the zero's sign, which
// no comparison on the result can see). NaN needs 0 * inf, so the caller takes this form
// only for rays with finite, nonzero inverse directions, a finite origin, non-NaN
// tmin/tmax, and a tree whose boxes are all finite.
RTX_HD float slab_fast(float ox, float oy, float oz, float ix, float iy, float iz, float tmin,
                       float x0, float y0, float z0, float x1, float y1, float z1, float tmax) {
    float t0x = (x0 - ox) * ix, t1x = (x1 - ox) * ix;
    float t0y = (y0 - oy) * iy, t1y = (y1 - oy) * iy;
    float t0z = (z0 - oz) * iz, t1z = (z1 - oz) * iz;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    return tn <= tf ? tn : kInf;
}

// Moller-Trumbore with back-face culling and no epsilons. Returns t (kInf on a miss).
RTX_HD float tri_test(f3 o, f3 d, float tmin, float tmax, f3 v0, f3 e1, f3 e2, float& u_out, float& v_out) {
    f3 h = cross(d, e2);
    float a = dot(e1, h);
    if (a >= 0) return kInf;
    float f = recip(a);
    f3 s = sub(o, v0);
    float u = f * dot(s, h);
    if (u < 0.f || u > 1.f) return kInf;
    f3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.f || u + v > 1.f) return kInf;
    float t = f * dot(e2, q);
    if (t < tmin || t > tmax) return kInf;
    u_out = u;
    v_out = v;
    return t;
}

// ---------------------------------------------------------------------------
// Sample table and sin/cos shared by the GPU renderer and its CPU checker.
// ---------------------------------------------------------------------------
RTX_HD uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
RTX_HD float sample01(uint32_t seed, uint32_t pixel, uint32_t s, uint32_t dim) {
    uint32_t h = mix32(seed ^ 0x9e3779b9u);
    h = mix32(h ^ pixel);
    h = mix32(h ^ s);
    h = mix32(h ^ dim);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
// Cody-Waite reduction by pi/2 (3-part constant) and Taylor polynomials on [-pi/4, pi/4].
RTX_HD void sincos_rtx(float x, float& s, float& c) {
    float kf = floorf(x * 0.636619747f + 0.5f);
    float r = ((x - kf * 1.57080078125f) - kf * -4.453584551811218e-06f) - kf * -8.705516307827565e-10f;
    float r2 = r * r;
    float sp = r + (r * r2) * (-1.66666672e-1f + r2 * (8.33333377e-3f + r2 * (-1.98412701e-4f + r2 * 2.75573188e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666679e-2f + r2 * (-1.38888892e-3f + r2 * (2.48015876e-5f + r2 * -2.75573188e-7f))));
    int q = ((int)kf) & 3;
    s = (q == 0) ? sp : (q == 1) ? cp : (q == 2) ? -sp : -cp;
    c = (q == 0) ? cp : (q == 1) ? -sp : (q == 2) ? -cp : sp;
}

// atan2 and acos shared by the GPU renderer and its CPU checker (so environment-map texel
// choices agree bit for bit): octant reduction to a = min/max in [0,1], then a degree-7
// polynomial in a^2 fitted to atan(a)/a (max error 2.5e-7 rad).
RTX_HD float atan2_rtx(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    const float a = mx == 0.f ? 0.f : mn / mx;
    const float s = a * a;
    float p = -0.004545152187347412f;
    p = p * s + 0.023694388568401337f;
    p = p * s + -0.058659493923187256f;
    p = p * s + 0.09853174537420273f;
    p = p * s + -0.13995960354804993f;
    p = p * s + 0.19965273141860962f;
    p = p * s + -0.33331647515296936f;
    p = p * s + 0.9999998211860657f;
    float r = p * a;
    if (ay > ax) r = 1.57079637f - r;
    if (x < 0.f) r = 3.14159274f - r;
    if (y < 0.f) r = -r;
    return r;
}
RTX_HD float acos_rtx(float y) { return atan2_rtx(sqrtf((1.f - y) * (1.f + y)), y); }

// Equirectangular environment lookup of IBL::sample (src/main.cpp:107-117): nearest texel,
// row 0 = straight up. rgb holds w*h*3 floats, top row first.
RTX_HD f3 ibl_lookup(const float* rgb, int w, int h, f3 d) {
    const float phi = atan2_rtx(d.z, d.x);
    const float cy = d.y < -1.f ? -1.f : (1.f < d.y ? 1.f : d.y);
    const float theta = acos_rtx(cy);
    const float u = (phi + 3.14159274f) / 6.28318548f;
    const float v = theta / 3.14159274f;
    int px = f2i_x86(u * (float)w), py = f2i_x86(v * (float)h);
    px = px < w - 1 ? px : w - 1;
    py = py < h - 1 ? py : h - 1;
    px = px < 0 ? 0 : px;
    py = py < 0 ? 0 : py;
    const float* p = rgb + 3 * ((size_t)py * w + px);
    return f3{p[0], p[1], p[2]};
}

}  // namespace rtx
