// Register-only fast path of the builder's triangle-to-box clip (src/bvh/AABB.cpp:86-171)
// for the common case of the spatial-split passes: only the two planes of one axis cut the
// triangle.
#pragma once
#include "rtx_common.hpp"

namespace rtx {

// One Sutherland-Hodgman intersection as the reference computes it (AABB.cpp:118-119).
RTX_HD f3 clip_lerp(const f3& p, const f3& q, float plane, int a) {
    const float pa = comp(p, a), qa = comp(q, a);
    const float t = (plane - pa) / (qa - pa);
    return f3{p.x + t * (q.x - p.x), p.y + t * (q.y - p.y), p.z + t * (q.z - p.z)};
}

// The reference clips against the planes min.x, max.x, min.y, max.y, min.z, max.z in that
// order. A plane that every current vertex is inside of emits the same vertices in the
// same order, so it can be skipped. This path applies when that holds for the four planes
// of the axes other than `a`:
//  - planes of axes before `a` are checked on the triangle itself;
//  - planes of axes after `a` are checked on the polygon left by the two planes of `a`.
// The two planes of `a` run Sutherland-Hodgman in slot form: the polygon after min[a] is
// the fixed sequence (v0, I01, v1, I12, v2, I20) with a valid flag per slot (the reference
// appends exactly these, in this order), and after max[a] each of those slots becomes (the
// slot, its intersection with the next valid slot). No index depends on the data, so the
// polygon stays in registers. The box is folded over the valid slots in order, as the
// reference folds its polygon.
// Returns false when a plane of another axis would cut (the caller then runs the full
// clip); otherwise out holds the reference's result (the empty box when nothing is left).
RTX_HD bool clip_tri_slab(const f3& v0, const f3& v1, const f3& v2, const Box& r, int a, Box& out) {
    for (int b = 0; b < a; ++b) {
        const float lo = comp(r.lo, b), hi = comp(r.hi, b);
        const float x0 = comp(v0, b), x1 = comp(v1, b), x2 = comp(v2, b);
        if (!(x0 >= lo && x1 >= lo && x2 >= lo && x0 <= hi && x1 <= hi && x2 <= hi)) return false;
    }
    const float lo = comp(r.lo, a), hi = comp(r.hi, a);
    const f3 v[3] = {v0, v1, v2};
    f3 S[6];
    bool F[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = i == 2 ? 0 : i + 1;
        const bool ci = comp(v[i], a) >= lo, cj = comp(v[j], a) >= lo;
        S[2 * i] = v[i];
        F[2 * i] = ci;
        F[2 * i + 1] = ci != cj;
        S[2 * i + 1] = (ci != cj) ? clip_lerp(v[i], v[j], lo, a) : v[i];
    }
    if (!(F[0] || F[1] || F[2] || F[3] || F[4] || F[5])) {
        out = Box::empty();
        return true;
    }
    f3 T[12];
    bool G[12];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        f3 q = S[k];   // the next valid slot after k, cyclically (k itself if it is the only one)
#pragma unroll
        for (int s = 5; s >= 1; --s) {
            const int idx = (k + s) % 6;
            if (F[idx]) q = S[idx];
        }
        const bool ci = comp(S[k], a) <= hi, cj = comp(q, a) <= hi;
        T[2 * k] = S[k];
        G[2 * k] = F[k] && ci;
        G[2 * k + 1] = F[k] && ci != cj;
        T[2 * k + 1] = (F[k] && ci != cj) ? clip_lerp(S[k], q, hi, a) : S[k];
    }
    bool any = false;
#pragma unroll
    for (int k = 0; k < 12; ++k) any = any || G[k];
    if (!any) {
        out = Box::empty();
        return true;
    }
    for (int b = a + 1; b < 3; ++b) {
        const float blo = comp(r.lo, b), bhi = comp(r.hi, b);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const float x = comp(T[k], b);
            if (G[k] && !(x >= blo && x <= bhi)) return false;
        }
    }
    Box bx = Box::empty();
#pragma unroll
    for (int k = 0; k < 12; ++k)
        if (G[k]) bx.extend(T[k]);
    out = bx;
    return true;
}

}  // namespace rtx
