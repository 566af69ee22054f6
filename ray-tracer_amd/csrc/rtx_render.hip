// Wavefront path tracer, multi-scene batching and the path-replay adjoint.
//
// Replaces the raster loop + pathTrace of the reference (src/main.cpp:145-204, 339-365).
// Every random number comes from the counter-based sample table (seed, pixel, sample,
// dim): dims 0/1 jitter, then per bounce b: 2+3b, 3+3b (direction), 4+3b (roulette).
//
// Paths of all scenes are numbered globally: scene-major, then sample, then pixel in 8 x 8
// tile order (tile_pixel). They
// are processed in batches of `batch_paths`; per batch:
//   k_camera -> { k_wtrace -> k_shade } x (max_bounce + 1) -> k_accum
// k_shade appends surviving paths to the next queue (an LDS scan per block, one atomic per
// block). The adjoint's replay runs the same wavefront with k_shade<kShadeReplay>, which
// also adds each bounce's gradient contributions (see k_shade).
// k_accum adds each pixel's samples into the image in sample order, so the sum is the
// reference's serial `accum += pathTrace(...)`; k_finalize divides by spp.
// No stage needs the host: queue lengths stay in device memory, so a whole render can be
// captured into one hipGraph (counters are cleared by a kernel, never by a memset).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "rtx_trace_impl.hpp"

namespace rtx {
namespace {

struct SceneDesc {
    TraceArgs T;
    const uint32_t* prims;
    const float* albedo;     // 3 per original triangle, or null (0.8)
    const float* emission;   // 3 per original triangle, or null
    float pos[3], fwd[3], right[3], up[3], tanh_;
    int W, H, spp, max_bounce, rr_bounce, env_mode;
    uint32_t seed;
    float env[3];
    float* image;
    uint64_t path_begin;
    uint32_t npix;
    uint32_t nprim;      // original triangle count (gradient array length / 3)
    const float* ibl;
    int ibl_w, ibl_h;
    const float* uv;     // texture albedo: 2 floats per input vertex
    const float* tex;    // tex_w * tex_h * 3; null = per-triangle albedo
    int tex_w, tex_h;
    float sun_dir[3], sun_cos, sun_rad[3];
};

struct PathState {
    float4* ray0;      // ox oy oz tmin
    float4* ray1;      // dx dy dz tmax
    float4* hit;       // t u v tri
    float4* tp;
    float4* L;
    uint32_t* sid;
    uint32_t* q[2];
    uint32_t* cnt;     // [b]: input queue length of bounce b; [ncnt/2 + b]: persistent-trace counter of bounce b
    uint32_t B;
    uint64_t g0;       // path number of the batch's first path (its paths are g0, g0 + 1, ...)
};

constexpr int kShadeForward = 0, kShadeReplay = 1;
// Where the replay adds its gradients: private copies of the arrays in the block's LDS,
// flushed once per block; or global atomics after merging equal entries within the wave.
constexpr int kGradNone = 0, kGradLds = 1, kGradGlobal = 2;

// Inputs and outputs of the replay's gradient scatter (scene 0 only). recL[rbase + i] holds
// path i's primal radiance (its path record, see k_store_L); dimg = dLoss/dimage (3 floats per pixel); ga, ge: gradient
// arrays (3 floats per albedo entry, per original triangle), either may be null.
struct GradArgs {
    const float4* recL;
    const float* dimg;
    float* ga;
    float* ge;
    uint32_t nprim;
    float4* hit_out;   // forward shading of bounce 0: path i's camera-ray hit to hit_out[rbase + i]
    uint64_t rbase;    // record of the batch's first path
};


__device__ __forceinline__ uint32_t find_scene(const uint64_t* off, int n, uint64_t g) {
    int lo = 0, hi = n;   // off[lo] <= g < off[hi]
    while (hi - lo > 1) {
        int m = (lo + hi) >> 1;
        if (off[m] <= g) lo = m; else hi = m;
    }
    return (uint32_t)lo;
}

// primaryRay (main.cpp:69-83)
__device__ __forceinline__ void primary_ray(const SceneDesc& S, int px, int py, float jx, float jy, f3& o, f3& d) {
    float aspect = (float)S.W / (float)S.H;
    float u = (((float)px + jx) / (float)S.W * 2.f - 1.f) * aspect;
    float v = ((float)py + jy) / (float)S.H * 2.f - 1.f;
    f3 fwd{S.fwd[0], S.fwd[1], S.fwd[2]}, rt{S.right[0], S.right[1], S.right[2]}, up{S.up[0], S.up[1], S.up[2]};
    d = normalize(add(add(fwd, scl(u * S.tanh_, rt)), scl(v * S.tanh_, up)));
    o = f3{S.pos[0], S.pos[1], S.pos[2]};
}

// sampleCosineHemisphere + toWorld (main.cpp:123-141)
__device__ __forceinline__ f3 bounce_dir(float u1, float u2, f3 n) {
    float r = sqrtf(u1);
    float theta = 6.28318548f * u2;
    float s, c;
    sincos_rtx(theta, s, c);
    f3 l{r * c, sqrtf(std_max(0.f, 1.f - u1)), r * s};
    f3 up = fabsf(n.x) < 0.9f ? f3{1.f, 0.f, 0.f} : f3{0.f, 1.f, 0.f};
    f3 t = normalize(cross(n, up));
    f3 b = cross(n, t);
    return add(add(scl(l.x, t), scl(l.y, n)), scl(l.z, b));
}

// Miss shading: the sky gradient of main.cpp:169-172, a constant, the IBL map, or a
// constant sky plus a sun disk (a distant light seen by the rays that reach its cone).
__device__ __forceinline__ f3 env_color(const SceneDesc& S, f3 d) {
    if (S.env_mode == 1) return f3{S.env[0], S.env[1], S.env[2]};
    if (S.env_mode == 2) return ibl_lookup(S.ibl, S.ibl_w, S.ibl_h, d);
    if (S.env_mode == 3) {
        const float c = dot(d, f3{S.sun_dir[0], S.sun_dir[1], S.sun_dir[2]});
        if (c >= S.sun_cos) return f3{S.env[0] + S.sun_rad[0], S.env[1] + S.sun_rad[1], S.env[2] + S.sun_rad[2]};
        return f3{S.env[0], S.env[1], S.env[2]};
    }
    float t = d.y * 0.5f + 0.5f;
    float omt = 1.f - t;
    return f3{1.0f * omt + 0.5f * t, 1.0f * omt + 0.7f * t, 1.0f * omt + 1.0f * t};
}

// Texel of a hit for texture albedo: uv interpolated like the shading normal, then the
// nearest texel (x86 truncation, clamped to the texture).
__device__ __forceinline__ uint32_t texel_at(const SceneDesc& S, uint32_t tri, float u, float v) {
    const uint32_t* t3 = S.T.tris3 + 3 * (size_t)tri;
    const float w = 1.f - u - v;
    const uint32_t i0 = ld_g(t3), i1 = ld_g(t3 + 1), i2 = ld_g(t3 + 2);
    const float2 a{ld_g(S.uv + 2 * i0), ld_g(S.uv + 2 * i0 + 1)};
    const float2 b{ld_g(S.uv + 2 * i1), ld_g(S.uv + 2 * i1 + 1)};
    const float2 c{ld_g(S.uv + 2 * i2), ld_g(S.uv + 2 * i2 + 1)};
    const float U = (w * a.x + u * b.x) + v * c.x;
    const float V = (w * a.y + u * b.y) + v * c.y;
    int px = f2i_x86(U * (float)S.tex_w), py = f2i_x86(V * (float)S.tex_h);
    px = px < S.tex_w - 1 ? px : S.tex_w - 1;
    py = py < S.tex_h - 1 ? py : S.tex_h - 1;
    px = px < 0 ? 0 : px;
    py = py < 0 ? 0 : py;
    return (uint32_t)py * (uint32_t)S.tex_w + (uint32_t)px;
}

__global__ void k_zero_u32(uint32_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Start of a batch: the first queue holds `a` paths; every other queue length and every
// bounce's trace counter starts at 0, so no bounce needs a clearing launch of its own.
__global__ void k_set_counts(uint32_t* p, uint32_t n, uint32_t a) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = i == 0 ? a : 0u;
}

// Pixel of slot q of a sample's W x H slots: 8 x 8 tiles, tile by tile along each band of 8
// rows, row-major inside a tile (the last band and each band's last tile may be narrower).
// Consecutive paths, and so the 64 of a wave, then cover a compact patch of the image, and
// their bounce rays start from neighbouring surface points. A bijection of [0, W H), so each
// pixel's samples keep their order; results do not depend on it.
__device__ __forceinline__ uint32_t tile_pixel(uint32_t q, uint32_t W, uint32_t H) {
    const uint32_t band = q / (8u * W), y0 = 8u * band, hb = min(8u, H - y0);
    const uint32_t r = q - band * 8u * W, full = W / 8u;
    uint32_t row, col;
    if (r < full * 8u * hb) {
        const uint32_t t = r / (8u * hb), k = r - t * 8u * hb;
        row = k / 8u;
        col = 8u * t + (k & 7u);
    } else {
        const uint32_t wl = W - 8u * full, k = r - full * 8u * hb;
        row = k / wl;
        col = 8u * full + k % wl;
    }
    return (y0 + row) * W + col;
}

__global__ void k_camera(const SceneDesc* __restrict__ scenes, const uint64_t* __restrict__ off, int ns, uint64_t g0,
                         uint32_t cnt, PathState P) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint64_t g = g0 + i;
        const uint32_t s = find_scene(off, ns, g);
        const SceneDesc& S = scenes[s];
        const uint64_t q = g - S.path_begin;
        const uint32_t sample = (uint32_t)(q / S.npix),
                       pixel = tile_pixel((uint32_t)(q % S.npix), (uint32_t)S.W, (uint32_t)S.H);
        const int x = (int)(pixel % (uint32_t)S.W), y = (int)(pixel / (uint32_t)S.W);
        f3 o, d;
        primary_ray(S, x, y, sample01(S.seed, pixel, sample, 0), sample01(S.seed, pixel, sample, 1), o, d);
        P.ray1[i] = make_float4(d.x, d.y, d.z, kInf);
        P.sid[i] = s;
    }
}

// Pixel and sample of batch path i, which belongs to scene S: as k_camera derives them from
// the path's number, so they are not stored.
__device__ __forceinline__ void path_pixel_sample(const PathState& P, const SceneDesc& S, uint32_t i, uint32_t& pixel,
                                                  uint32_t& sample) {
    const uint64_t q = P.g0 + i - S.path_begin;
    sample = (uint32_t)(q / S.npix);
    pixel = tile_pixel((uint32_t)(q % S.npix), (uint32_t)S.W, (uint32_t)S.H);
}

// What bounce 0 does not read from the path state, because it is the same for every camera
// ray: the origin and tmin (the scene's camera position and 1e-3, as primary_ray gives
// them), the throughput (1, 1, 1), the radiance (0, 0, 0) and the queue (the identity: the
// batch's paths in order). k_camera writes only the direction and the scene.
__device__ __forceinline__ float4 ray_origin(const PathState& P, const SceneDesc& S, uint32_t i, int bounce) {
    return bounce == 0 ? make_float4(S.pos[0], S.pos[1], S.pos[2], 1e-3f) : P.ray0[i];
}
__device__ __forceinline__ uint32_t queued_path(const uint32_t* Q, uint32_t k, int bounce) {
    return bounce == 0 ? k : Q[k];
}

// Path records of the primal render (one scene), the replay's only input per path, in two
// halves of the render's N paths: recL[g] = (radiance L, 0) and recH[g] = the hit of the
// path's camera ray (t, u, v, triangle bits), g the path's number (sample * npix + slot, the
// slot's pixel by tile_pixel), recH = recL + N. 32 bytes per path, whatever max_bounce is;
// with the camera rays' hits stored the replay does not trace them again. recH is written by
// the forward shading of bounce 0 (GradArgs::hit_out).
// Each path's radiance L into recL[base + i].
__global__ void k_store_L(PathState P, uint32_t cnt, float4* __restrict__ recL, uint64_t base) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const float4 L = P.L[i];
        recL[base + i] = make_float4(L.x, L.y, L.z, 0.f);
    }
}
// The replay's camera rays take their hits from the records instead of a trace: the same
// rays (the same sample table) give the same hits. Queue 0 is the identity.
__global__ void k_load_hit(PathState P, uint32_t cnt, const float4* __restrict__ recH, uint64_t base) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x)
        P.hit[i] = recH[base + i];
}

// The wavefront trace loads leaf triangles two at a time and is held to 64 VGPRs so that
// 8 waves fit per SIMD (measured faster than 4-triangle batches at 5 waves, both for the
// 64-scene batched render and for the 4118-triangle differentiable render). The
// cooperative-leaf loop is held to 7 waves (72 VGPRs, a few spilled): measured against 5
// (86 VGPRs), 6 and 8 waves, it was fastest on all three render workloads (C5 215 -> 184
// ms), and at 8 waves the spills cost more than the wave gained.
constexpr int kRenderTriBatch = 2;

// Where a lane finds its scene's tree (SC):
//  - kSceneUni: every queued path belongs to scenes[0] (a launch sequence for one scene).
//    The tree pointers are wave-uniform kernel data that stay in scalar registers.
//  - kSceneLds: paths of up to kLdsScenes scenes share the queue. The block copies each
//    scene's node, leaf and triangle pointers into an LDS table; a lane keeps its scene
//    index and reads the pointers from LDS before a node or leaf access.
//  - kSceneGlobal: more scenes: the lane reads them through its scene descriptor in global
//    memory (a dependent load before every node and leaf access).
constexpr int kSceneUni = 0, kSceneLds = 1, kSceneGlobal = 2;
constexpr uint32_t kLdsScenes = 64;
struct SceneTrees {
    const GNode* nodes;
    const Leaf* leaves;
    const GTri* tris;
};

// Chunked form: each wave traces 64 consecutive queue entries to the end.
template <int DEPTH, int LOOP, int SC>
__global__ void __launch_bounds__(kTraceThreads, LOOP == kLoopCoop ? 7 : 8)
    k_wtrace(const SceneDesc* __restrict__ scenes, uint32_t nscenes, PathState P, int qi, int bounce, uint32_t wait) {
    __shared__ uint2 lst[DEPTH][kTraceThreads];
    __shared__ CoopLds coop[LOOP == kLoopCoop ? kTraceThreads / 64 : 1];
    __shared__ SceneTrees tab[SC == kSceneLds ? kLdsScenes : 1];
    if (SC == kSceneLds) {
        for (uint32_t j = threadIdx.x; j < nscenes; j += kTraceThreads)
            tab[j] = SceneTrees{scenes[j].T.nodes, scenes[j].T.leaves, scenes[j].T.tris};
        __syncthreads();
    }
    const uint32_t n = P.cnt[bounce];
    const uint32_t* Q = P.q[qi];
    // wave-uniform trip count (the while-while loop votes across the wave)
    for (uint32_t wb = xcd_block(blockIdx.x) * kTraceThreads + (threadIdx.x & ~63u); wb < n;
         wb += gridDim.x * kTraceThreads) {
        const uint32_t k = wb + __lane_id();
        const bool in = k < n;
        uint32_t i = 0;
        float4 r0 = make_float4(0.f, 0.f, 0.f, 1.f), r1 = make_float4(0.f, 0.f, 1.f, 0.f);
        const SceneDesc* S = scenes;
        uint32_t sid = 0;
        if (in) {
            i = queued_path(Q, k, bounce);
            r1 = P.ray1[i];
            if (SC != kSceneUni) {
                sid = P.sid[i];
                S = scenes + sid;
            }
            r0 = ray_origin(P, *S, i, bounce);
        }
        TraceArgs A = SC == kSceneUni ? scenes->T : S->T;
        if (SC == kSceneLds) {
            A.nodes = tab[sid].nodes;
            A.leaves = tab[sid].leaves;
            A.tris = tab[sid].tris;
        }
        const TraceOut h = trace_ray<false, false, DEPTH, LOOP, kRenderTriBatch, SC != kSceneUni>(
            A, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, r0.w, r1.w, lst, threadIdx.x, wait,
            coop + (LOOP == kLoopCoop ? threadIdx.x / 64 : 0));
        if (in) P.hit[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.tri));
    }
}

// Persistent variant (render_loop 2): a grid of waves that refill idle lanes (at least
// kRenderRefill at a time) from the queue, so a wave does not wait for the slowest path of
// a 64-path chunk. Each wave takes kRenderChunk queue positions per atomic on `work` and
// serves its refills from them: with one atomic per refill the single counter serialized
// the launch (C3 162 ms against 89 ms for the chunked kernel; 128-path chunks: 96 ms, and
// C4 10.4 -> 8.7 ms, C5 184 -> 152 ms). The control flow follows k_trace_persist: every
// ballot, the leader's atomic, the shuffle and the exit test are wave-uniform, and no lane
// leaves the loop body early.
constexpr int kRenderRefill = 32;   // 24: C5 123.7 vs 122.4 ms; 16: 126.1 ms
constexpr int kRenderChunk = 128;   // queue positions per atomic on the path counter

template <int DEPTH, int LOOP, int SC>
__global__ void __launch_bounds__(kTraceThreads, LOOP == kLoopCoop ? 7 : 8)
    k_wtrace_persist(const SceneDesc* __restrict__ scenes, uint32_t nscenes, PathState P, int qi, int bounce,
                     uint32_t wait, uint32_t* work) {
    __shared__ uint2 lst[DEPTH][kTraceThreads];
    __shared__ CoopLds coop[LOOP == kLoopCoop ? kTraceThreads / 64 : 1];
    __shared__ SceneTrees tab[SC == kSceneLds ? kLdsScenes : 1];
    if (SC == kSceneLds) {
        for (uint32_t i = threadIdx.x; i < nscenes; i += kTraceThreads)
            tab[i] = SceneTrees{scenes[i].T.nodes, scenes[i].T.leaves, scenes[i].T.tris};
        __syncthreads();
    }
    const uint32_t n = P.cnt[bounce];
    const uint32_t* Q = P.q[qi];
    const uint32_t tid = threadIdx.x;
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    Walker<DEPTH, kRenderTriBatch> w;
    uint2 spill[kSpillDepth];
    w.spill = spill;
    w.kill();
    w.best = kInvalid;
    uint32_t id = kInvalid;            // path being traced
    const SceneDesc* S = scenes;       // its scene
    uint32_t sid = 0;                  // its scene index
    const TraceArgs TU = scenes->T;    // kSceneUni: the one scene's tree, loaded once into scalar registers
    // the lane's tree for node and leaf accesses (the other fields are only read at start)
    auto args = [&]() -> TraceArgs {
        if (SC == kSceneUni) return TU;
        if (SC == kSceneGlobal) return S->T;
        TraceArgs A = TU;
        A.nodes = tab[sid].nodes;
        A.leaves = tab[sid].leaves;
        A.tris = tab[sid].tris;
        return A;
    };
    uint32_t rbase = 0, rend = 0;      // wave-uniform reservoir of queue positions [rbase, rend)
    bool gdone = false;                // `work` has passed n
    bool more = true;                  // the same on every lane: set only from wave-wide values
    for (;;) {
        const uint64_t idle = __ballot(!w.alive());
        const uint32_t nidle = (uint32_t)__popcll(idle);
        const bool refill = more && nidle >= (uint32_t)kRenderRefill;   // wave-uniform
        const bool done = !more && nidle == 64u;                       // wave-uniform
        // finished paths are written when their lanes take new ones or the wave ends, so
        // finish()'s test runs for many lanes at once
        if ((refill || done) && !w.alive() && id != kInvalid) {
            w.finish(args());
            P.hit[id] = make_float4(w.best != kInvalid ? w.bestT : 0.f, w.bu, w.bv, __uint_as_float(w.best));
            id = kInvalid;
        }
        if (done) break;
        if (refill) {
            if (rbase == rend) {   // one atomic takes kRenderChunk queue positions
                const int leader = __ffsll((unsigned long long)idle) - 1;
                uint32_t b = 0;
                if (__lane_id() == leader) b = atomicAdd(work, (uint32_t)kRenderChunk);
                b = __shfl(b, leader);
                rbase = min(b, n);
                rend = min(b + (uint32_t)kRenderChunk, n);
                gdone = b + (uint32_t)kRenderChunk >= n;
            }
            const uint32_t base = rbase;
            rbase = min(rbase + nidle, rend);
            more = !gdone || rbase < rend;
            if (!w.alive()) {
                const uint32_t k = base + (uint32_t)__popcll(idle & lt);
                if (k < rend) {
                    id = queued_path(Q, k, bounce);
                    if (SC != kSceneUni) {
                        sid = P.sid[id];
                        S = scenes + sid;
                    }
                    const float4 r0 = ray_origin(P, *S, id, bounce), r1 = P.ray1[id];
                    w.start(SC == kSceneUni ? TU : S->T, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, r0.w, r1.w);
                }
            }
        }
        // a started path that missed the root is written at the next refill or at the end;
        // once the queue is empty the leaf-phase wait scales with the lanes still tracing
        const uint32_t wt = more ? wait : (wait * (uint32_t)__popcll(__ballot(w.alive()))) >> 6;
        w.template node_phase<false>(args(), lst, tid, wt);
        if (LOOP == kLoopCoop) {   // all lanes active: the loop is wave-uniform
            const bool at = w.alive() && (w.ref & kLeafBit);
            w.template leaf_coop<false, false, SC != kSceneUni>(args(), coop + tid / 64);
            if (at) w.pop(lst, tid);
        } else if (w.alive() && (w.ref & kLeafBit)) {
            w.template leaf_step<false, false>(args());
            w.pop(lst, tid);
        }
    }
}

// Atomic accumulation for the adjoint: before adding, the wave merges lanes that target
// the same triangle (one merge round per distinct triangle, up to kMergeRounds); lanes
// left after that add their own contributions.
constexpr int kMergeRounds = 8;

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ void merged_add(float* __restrict__ buf, bool has, uint32_t prim, f3 g) {
    uint64_t pending = __ballot(has);
    const uint64_t me = 1ull << __lane_id();
    for (int round = 0; pending && round < kMergeRounds; ++round) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t p = __shfl(prim, leader);
        const bool in = has && (pending & me) && prim == p;
        const uint64_t m = __ballot(in);
        const float sx = wave_sum(in ? g.x : 0.f), sy = wave_sum(in ? g.y : 0.f), sz = wave_sum(in ? g.z : 0.f);
        if (__lane_id() == leader) {
            atomicAdd(&buf[3 * p], sx);
            atomicAdd(&buf[3 * p + 1], sy);
            atomicAdd(&buf[3 * p + 2], sz);
        }
        pending &= ~m;
    }
    if (has && (pending & me)) {
        atomicAdd(&buf[3 * prim], g.x);
        atomicAdd(&buf[3 * prim + 1], g.y);
        atomicAdd(&buf[3 * prim + 2], g.z);
    }
}

// Sum over the 64 lanes of a wave (every lane active): DPP row shifts and row broadcasts
// into lane 63 (the order of wave_incl_scan), read back as a uniform value.
__device__ __forceinline__ float wave_total(float x) {
#define RTX_DPP_ADD(ctrl, rows) x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), ctrl, rows, 0xf, false))
    RTX_DPP_ADD(0x111, 0xf);   // row_shr:1
    RTX_DPP_ADD(0x112, 0xf);   // row_shr:2
    RTX_DPP_ADD(0x114, 0xf);   // row_shr:4
    RTX_DPP_ADD(0x118, 0xf);   // row_shr:8
    RTX_DPP_ADD(0x142, 0xa);   // row_bcast:15 into rows 1, 3
    RTX_DPP_ADD(0x143, 0xc);   // row_bcast:31 into rows 2, 3
#undef RTX_DPP_ADD
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// The replay's gradient adds into the block's LDS copies (ga at lg, ge at lg + eoff; ua / ue:
// the array is kept). Lanes of a wave often add to the same entry (neighbouring paths hit the
// same wall), and LDS atomics on one address serialize, so per round the lanes sharing the
// first pending lane's entry are summed by wave_total and added once; a round whose entry has
// a single lane is that lane's own add. After kLdsMergeRounds rounds the lanes left add
// their own contributions. Every lane of the wave must be active.
constexpr int kLdsMergeRounds = 4;
__device__ __forceinline__ void lds_grad_add(float* lg, uint32_t eoff, bool ua, bool ue, bool has, uint32_t idx,
                                             f3 gA, f3 gE) {
    auto add = [&](uint32_t e, f3 a, f3 g) {
        if (ua) { atomicAdd(&lg[3 * e], a.x); atomicAdd(&lg[3 * e + 1], a.y); atomicAdd(&lg[3 * e + 2], a.z); }
        if (ue) {
            atomicAdd(&lg[eoff + 3 * e], g.x); atomicAdd(&lg[eoff + 3 * e + 1], g.y); atomicAdd(&lg[eoff + 3 * e + 2], g.z);
        }
    };
    uint64_t pending = __ballot(has);
    const uint64_t me = 1ull << __lane_id();
    for (int round = 0; pending && round < kLdsMergeRounds; ++round) {
        const int leader = __ffsll((unsigned long long)pending) - 1;   // uniform
        const uint32_t p = (uint32_t)__builtin_amdgcn_readlane((int)idx, leader);
        const bool in = (pending & me) && idx == p;
        const uint64_t m = __ballot(in);
        pending &= ~m;
        if (__popcll(m) == 1) {
            if (in) add(p, gA, gE);
            continue;
        }
        f3 sa{0.f, 0.f, 0.f}, se{0.f, 0.f, 0.f};
        if (ua) sa = f3{wave_total(in ? gA.x : 0.f), wave_total(in ? gA.y : 0.f), wave_total(in ? gA.z : 0.f)};
        if (ue) se = f3{wave_total(in ? gE.x : 0.f), wave_total(in ? gE.y : 0.f), wave_total(in ? gE.z : 0.f)};
        if (__lane_id() == leader) add(p, sa, se);
    }
    if (pending & me) add(idx, gA, gE);
}

// Dynamic LDS the replay may take for its gradient arrays (a gfx950 workgroup may allocate
// 160 KiB; k_shade's own LDS takes under 256 bytes), and its launch shape when they fit: 256
// blocks of 1024 threads (C4 forward+adjoint, with a separate gradient kernel: 128 blocks of
// 512 threads one launch per array 19.6 ms; both arrays in one launch 16.5 ms; 256 blocks of
// 1024 threads 16.2 ms).
constexpr size_t kLdsGradBytes = 160 * 1024 - 256;
constexpr int kGradThreads = 1024;
constexpr uint32_t kGradLdsBlocks = 256;

// One bounce of pathTrace (main.cpp:158-200) for every queued path. UNI: every path belongs
// to scenes[0], whose descriptor is then read once as uniform data.
//
// MODE kShadeReplay is the path-replay adjoint's bounce: the replay is the forward wavefront
// again (the same sample table gives the same paths, bit for bit), and at the surface hit k
// of path i this bounce adds, weighted by dLoss/dI of the path's pixel / spp,
//   dL/da_k = (L - L_upto_k) / a_k   to ga[albedo entry],   dL/dLe_k = tp_k   to ge[triangle],
// with L the path's primal radiance (from G.recL), L_upto_k the radiance gathered up to and
// including hit k's own emission and tp_k the throughput arriving at k: with
// L = sum_m (prod_{j<m} a_j) X_m everything past hit k is proportional to a_k. Nothing is
// kept per path between bounces beyond the forward path state, so memory does not grow with
// max_bounce. At a zero albedo channel the path carried nothing past the vertex: that vertex
// adds 0 to the channel (not 0/0). GM: kGradLds (both arrays in LDS, launched as at most
// kGradLdsBlocks blocks of kGradThreads) or kGradGlobal.
template <int MODE, bool UNI, int GM = kGradNone>
__global__ void k_shade(const SceneDesc* __restrict__ scenes, PathState P, int qi, int bounce, GradArgs G) {
    constexpr bool ADJ = MODE == kShadeReplay;
    extern __shared__ float lg[];   // kGradLds: ga's copy (if set), then ge's (if set)
    const uint32_t n3 = 3 * G.nprim, eoff = G.ga ? n3 : 0u, nlds = GM == kGradLds ? eoff + (G.ge ? n3 : 0u) : 0u;
    if (GM == kGradLds) {
        for (uint32_t j = threadIdx.x; j < nlds; j += blockDim.x) lg[j] = 0.f;
        __syncthreads();
    }
    const uint32_t n = P.cnt[bounce];
    const uint32_t* Q = P.q[qi];
    uint32_t* Qo = P.q[qi ^ 1];
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    __shared__ uint32_t wcount[16];   // surviving paths per wave of the block (<= 1024 threads)
    __shared__ uint32_t bbase;
    const uint32_t wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // trip count uniform per block: the queue append below is one atomic per block, and the
    // replay's wave merge votes across the wave
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t bb = blockIdx.x * blockDim.x; bb < n; bb += stride) {
        const uint32_t k = bb + threadIdx.x;
        bool cont = false, has = false;
        uint32_t i = 0, gidx = 0;
        f3 gA{0.f, 0.f, 0.f}, gE{0.f, 0.f, 0.f};
        if (k < n) {
            i = queued_path(Q, k, bounce);
            const SceneDesc& S = scenes[UNI ? 0u : P.sid[i]];
            const float4 h = P.hit[i];
            if (!ADJ && G.hit_out) G.hit_out[G.rbase + i] = h;   // bounce 0: queue 0 is the identity
            const uint32_t tri = __float_as_uint(h.w);
            float4 tp = make_float4(1.f, 1.f, 1.f, 0.f), L = make_float4(0.f, 0.f, 0.f, 0.f);
            if (bounce > 0) {
                tp = P.tp[i];
                L = P.L[i];
            }
            const float4 r1 = P.ray1[i];
            const f3 d{r1.x, r1.y, r1.z};
            if (tri == kInvalid) {
                f3 e = env_color(S, d);
                L = make_float4(L.x + tp.x * e.x, L.y + tp.y * e.y, L.z + tp.z * e.z, 0.f);
            } else {
                const uint32_t prim = ld_g(S.prims + tri);
                // albedo entry: the hit's texel (texture albedo) or its original triangle
                const float* alb = S.tex ? S.tex : S.albedo;
                const uint32_t aidx = S.tex ? texel_at(S, tri, h.y, h.z) : prim;
                const float4 tpin = tp;
                if (S.emission) {
                    const float* e = S.emission + 3 * prim;
                    L = make_float4(L.x + tp.x * ld_g(e), L.y + tp.y * ld_g(e + 1), L.z + tp.z * ld_g(e + 2), 0.f);
                }
                float ax = 0.8f, ay = 0.8f, az = 0.8f;
                const bool more = bounce < S.max_bounce;
                if (alb && (more || ADJ)) {
                    ax = ld_g(alb + 3 * aidx); ay = ld_g(alb + 3 * aidx + 1); az = ld_g(alb + 3 * aidx + 2);
                }
                uint32_t px, sm;
                path_pixel_sample(P, S, i, px, sm);
                if (ADJ) {
                    has = true;
                    gidx = aidx;
                    const float fs = (float)S.spp;
                    const f3 delta{ld_g(G.dimg + 3 * px) / fs, ld_g(G.dimg + 3 * px + 1) / fs,
                                   ld_g(G.dimg + 3 * px + 2) / fs};
                    if (G.ga) {
                        const float4 lf = ld_g(G.recL + G.rbase + i);
                        gA = f3{ax != 0.f ? delta.x * (lf.x - L.x) / ax : 0.f,
                                ay != 0.f ? delta.y * (lf.y - L.y) / ay : 0.f,
                                az != 0.f ? delta.z * (lf.z - L.z) / az : 0.f};
                    }
                    gE = f3{delta.x * tpin.x, delta.y * tpin.y, delta.z * tpin.z};
                }
                if (more) {
                    const float u1 = sample01(S.seed, px, sm, 2 + 3 * bounce);
                    const float u2 = sample01(S.seed, px, sm, 3 + 3 * bounce);
                    const f3 nrm = hit_normal(S.T, tri, h.y, h.z);
                    const f3 wd = bounce_dir(u1, u2, nrm);
                    tp = make_float4(tp.x * ax, tp.y * ay, tp.z * az, 0.f);
                    cont = true;
                    if (bounce >= S.rr_bounce) {
                        float pr = tp.x;
                        if (pr < tp.y) pr = tp.y;
                        if (pr < tp.z) pr = tp.z;
                        if (sample01(S.seed, px, sm, 4 + 3 * bounce) > pr) cont = false;
                        else tp = make_float4(tp.x / pr, tp.y / pr, tp.z / pr, 0.f);
                    }
                    if (cont) {
                        const float4 r0 = ray_origin(P, S, i, bounce);
                        const f3 o = add(f3{r0.x, r0.y, r0.z}, scl(h.x, d));
                        P.ray0[i] = make_float4(o.x, o.y, o.z, 1e-4f);
                        P.ray1[i] = make_float4(wd.x, wd.y, wd.z, kInf);
                        P.tp[i] = tp;
                    }
                }
            }
            P.L[i] = L;
        }
        if (GM == kGradLds) {
            lds_grad_add(lg, eoff, G.ga != nullptr, G.ge != nullptr, has, gidx, gA, gE);
        } else if (GM == kGradGlobal) {
            if (G.ga) merged_add(G.ga, has, gidx, gA);
            if (G.ge) merged_add(G.ge, has, gidx, gE);
        }
        const uint64_t m = __ballot(cont);
        if (__lane_id() == 0) wcount[wid] = (uint32_t)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (uint32_t w = 0; w < nw; ++w) { const uint32_t c = wcount[w]; wcount[w] = run; run += c; }
            bbase = run ? atomicAdd(&P.cnt[bounce + 1], run) : 0u;
        }
        __syncthreads();
        if (cont) Qo[bbase + wcount[wid] + __popcll(m & lt)] = i;
        __syncthreads();
    }
    if (GM == kGradLds) {
        for (uint32_t j = threadIdx.x; j < nlds; j += blockDim.x) {
            const float x = lg[j];
            if (x != 0.f) atomicAdd(j < eoff ? &G.ga[j] : &G.ge[j - eoff], x);
        }
    }
}

// In-order accumulation: the first path of each pixel in the batch adds that pixel's
// samples in increasing sample order.
__global__ void k_accum(const SceneDesc* __restrict__ scenes, const uint64_t* __restrict__ off, int ns, uint64_t g0,
                        uint32_t cnt, PathState P) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint32_t s = P.sid[i];
        const SceneDesc& S = scenes[s];
        uint32_t pixel, sample;
        path_pixel_sample(P, S, i, pixel, sample);
        if (sample != 0 && i >= S.npix) continue;   // an earlier sample of this pixel is in the batch
        float* img = S.image + 3 * (size_t)pixel;
        float ax = img[0], ay = img[1], az = img[2];
        for (uint32_t j = i, sm = sample; j < cnt && sm < (uint32_t)S.spp; j += S.npix, ++sm) {
            const float4 L = P.L[j];
            ax += L.x; ay += L.y; az += L.z;
        }
        img[0] = ax; img[1] = ay; img[2] = az;
    }
    (void)off; (void)ns; (void)g0;
}

// Device copies of the scene table and path offsets, written by kernels that carry them as
// launch arguments: ordered on the render's stream like its other work, with no host copy
// to wait for and no host buffer to keep alive.
constexpr int kPutDescs = 8;     // 8 descriptors per launch (kernel arguments stay under 4 KB)
constexpr int kPutOffs = 64;
struct DescChunk {
    SceneDesc d[kPutDescs];
    uint32_t n;
};
struct OffChunk {
    uint64_t v[kPutOffs];
    uint32_t n;
};
__global__ void k_put_descs(SceneDesc* __restrict__ dst, DescChunk c) {
    for (uint32_t i = threadIdx.x; i < c.n; i += blockDim.x) dst[i] = c.d[i];
}
__global__ void k_put_offs(uint64_t* __restrict__ dst, OffChunk c) {
    for (uint32_t i = threadIdx.x; i < c.n; i += blockDim.x) dst[i] = c.v[i];
}
static_assert(sizeof(DescChunk) < 4096 && sizeof(OffChunk) < 4096, "kernel arguments");

__global__ void k_finalize(const SceneDesc* __restrict__ scenes, int s) {
    const SceneDesc& S = scenes[s];
    const float fs = (float)S.spp;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < S.npix * 3u; p += gridDim.x * blockDim.x)
        S.image[p] = S.image[p] / fs;
}

SceneDesc make_desc(const rtx_bvh* b, const rtx_camera* c, const rtx_render_params* p, const float* albedo,
                    const float* emission, float* image) {
    SceneDesc S;
    S.T = make_trace_args(b);
    S.prims = b->prims;
    S.albedo = albedo;
    S.emission = emission;
    for (int k = 0; k < 3; ++k) {
        S.pos[k] = c->pos[k]; S.fwd[k] = c->fwd[k]; S.right[k] = c->right[k]; S.up[k] = c->up[k];
        S.env[k] = p->env_color[k];
    }
    S.tanh_ = c->tan_half_fov_y;
    S.W = p->width; S.H = p->height; S.spp = p->spp; S.max_bounce = p->max_bounce; S.rr_bounce = p->rr_bounce;
    S.env_mode = p->env_mode;
    S.seed = p->seed;
    S.image = image;
    S.path_begin = 0;
    S.npix = (uint32_t)(p->width * p->height);
    S.nprim = b->nt;
    S.ibl = p->ibl_d;
    S.ibl_w = p->ibl_width;
    S.ibl_h = p->ibl_height;
    S.uv = p->uv_d;
    S.tex = p->tex_d;
    S.tex_w = p->tex_width;
    S.tex_h = p->tex_height;
    if (S.tex) S.nprim = (uint32_t)(p->tex_width * p->tex_height);   // gradient entries: texels
    for (int k = 0; k < 3; ++k) { S.sun_dir[k] = p->sun_dir[k]; S.sun_rad[k] = p->sun_radiance[k]; }
    S.sun_cos = p->sun_cos;
    return S;
}

int check_scene(const rtx_bvh* b, const rtx_camera* c, const rtx_render_params* p, const void* image,
                const char* fn) {
    if (!b || !c || !p || !image) return fail(RTX_E_ARG, std::string("null argument on ") + fn);
    if (!b->gnodes) return fail(RTX_E_STATE, std::string("BVH not loaded on ") + fn);
    if (p->width <= 0 || p->height <= 0 || p->spp <= 0 || p->max_bounce < 0 || p->env_mode < 0 || p->env_mode > 3)
        return fail(RTX_E_ARG, std::string("bad render parameters on ") + fn);
    if (p->tex_d && (!p->uv_d || p->tex_width <= 0 || p->tex_height <= 0 ||
                     (uint64_t)p->tex_width * p->tex_height > 0x3FFFFFFFull))
        return fail(RTX_E_ARG, std::string("texture albedo needs uv_d and a valid texture size on ") + fn);
    if (p->env_mode == 2 && (!p->ibl_d || p->ibl_width <= 0 || p->ibl_height <= 0))
        return fail(RTX_E_ARG, std::string("env_mode 2 needs an IBL map on ") + fn);
    if ((uint64_t)p->width * p->height > 0xFFFFFFFFull / 3)
        return fail(RTX_E_LIMIT, std::string("image too large on ") + fn);
    if (b->max_depth > kMaxTraceDepth)
        return fail(RTX_E_LIMIT, std::string("tree deeper than the traversal stack on ") + fn);
    return RTX_OK;
}

}  // namespace

// A prepared render of one or more scenes: device scene table, path state, and (for the
// batched mode) the captured graph.
struct Batch {
    std::vector<SceneDesc> descs;
    std::vector<uint64_t> off;
    SceneDesc* d_scenes = nullptr;
    uint64_t* d_off = nullptr;
    PathState P{};
    std::vector<std::pair<void*, size_t>> allocs;   // blocks of the scratch cache
    hipEvent_t done = nullptr;                      // recorded after every launch
    uint32_t B = 0;
    int mode = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int max_b = 0;
    uint32_t ncnt = 0;   // entries of P.cnt: queue lengths, then trace counters, one per bounce

    hipStream_t sst = nullptr;    // stream of setup's uploads
    hipStream_t last = nullptr;   // stream of the latest launch

    ~Batch() {
        // A graph batch may still run: wait for it before its graph and blocks go. Otherwise
        // the blocks go back to the cache ordered after the latest launch's stream: whoever
        // takes one next waits for that work on the GPU, not here (launches of one batch are
        // ordered by its caller: they share the path state).
        if (exec && done) (void)hipEventSynchronize(done);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        const hipStream_t after = last ? last : sst;   // setup's uploads may be the last work
        const bool pending = !exec && (done || sst);
        if (done) (void)hipEventDestroy(done);
        for (auto& a : allocs) {
            if (pending) scratch_give_after(a.first, a.second, after);
            else scratch_give(a.first, a.second);
        }
    }
    template <class T>
    T* alloc(size_t n) {
        size_t got = 0;
        T* p = (T*)scratch_take(std::max<size_t>(n, 1) * sizeof(T), got, sst);
        if (p) allocs.push_back({p, got});
        return p;
    }
    int note_launch(hipStream_t s) {
        if (!done) RTX_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        RTX_HIP(hipEventRecord(done, s));
        last = s;
        return RTX_OK;
    }
    // adjoint (scene 0 only): dLoss/dimage in, gradients out. adj 0: a forward render (which
    // also stores each path's record in rec when that is set); 1: a primal pass and the
    // replay, per batch; 2: the replay alone, with the path records read from rec. A record
    // array holds the radiance half (one float4 per path) and then the hit half (k_store_L).
    int adj = 0;
    const float* dimg = nullptr;
    float* ga = nullptr;
    float* ge = nullptr;
    float4* rec = nullptr;     // the render's path records (2 * paths float4), by global path number
    float4* recb = nullptr;    // adj 1: the batch's path records (2 * B float4)
    // Takes the path state from the scratch cache and uploads the scene table, ordered on st.
    int setup(uint32_t batch_paths, hipStream_t st, bool adjoint = false);
    void enqueue_range(int s_first, int s_last, hipStream_t st);   // scenes [s_first, s_last)
    int launch(hipStream_t st);
};

int Batch::setup(uint32_t batch_paths, hipStream_t st, bool adjoint) {
    sst = st;
    const int ns = (int)descs.size();
    off.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        descs[s].path_begin = off[s];
        off[s + 1] = off[s] + (uint64_t)descs[s].npix * descs[s].spp;
        max_b = std::max(max_b, descs[s].max_bounce);
    }
    uint64_t total = off[ns];
    B = (uint32_t)std::min<uint64_t>(std::max<uint32_t>(batch_paths, 1024u), total);
    d_scenes = alloc<SceneDesc>(ns);
    d_off = alloc<uint64_t>(ns + 1);
    P.ray0 = alloc<float4>(B); P.ray1 = alloc<float4>(B); P.hit = alloc<float4>(B);
    P.tp = alloc<float4>(B); P.L = alloc<float4>(B);
    P.sid = alloc<uint32_t>(B);
    P.q[0] = alloc<uint32_t>(B); P.q[1] = alloc<uint32_t>(B);
    ncnt = 2 * (uint32_t)(max_b + 2);
    P.cnt = alloc<uint32_t>(ncnt);
    P.B = B;
    if (adjoint) {   // flat in max_bounce: only the batch's path records are added
        // the replay's LDS form may take more than the default 64 KiB of dynamic LDS
        static const bool lds_ok =
            hip_ok(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_shade<kShadeReplay, true, kGradLds>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsGradBytes),
                   "hipFuncSetAttribute") &&
            hip_ok(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_shade<kShadeReplay, false, kGradLds>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsGradBytes),
                   "hipFuncSetAttribute");
        if (!lds_ok) return RTX_E_HIP;
        recb = alloc<float4>(2ull * B);
        if (!recb) return fail(RTX_E_HIP, "device allocation failed on render setup");
    }
    if (!d_scenes || !d_off || !P.ray0 || !P.ray1 || !P.hit || !P.tp || !P.L || !P.sid ||
        !P.q[0] || !P.q[1] || !P.cnt)
        return fail(RTX_E_HIP, "device allocation failed on render setup");
    for (int s0 = 0; s0 < ns; s0 += kPutDescs) {
        DescChunk c{};
        c.n = (uint32_t)std::min(kPutDescs, ns - s0);
        for (uint32_t k = 0; k < c.n; ++k) c.d[k] = descs[s0 + k];
        k_put_descs<<<1, 64, 0, st>>>(d_scenes + s0, c);
    }
    for (int s0 = 0; s0 <= ns; s0 += kPutOffs) {
        OffChunk c{};
        c.n = (uint32_t)std::min(kPutOffs, ns + 1 - s0);
        for (uint32_t k = 0; k < c.n; ++k) c.v[k] = off[s0 + k];
        k_put_offs<<<1, 64, 0, st>>>(d_off + s0, c);
    }
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}

// Kernel sequence for the paths of scenes [s_first, s_last): batches of B paths.
void Batch::enqueue_range(int s_first, int s_last, hipStream_t st) {
    const uint32_t grid = g_tuning.grid_blocks, sh = g_tuning.shade_threads, wait = g_tuning.render_wait_lanes;
    const uint32_t pgrid = g_tuning.trace_grid_blocks;   // blocks of the persistent trace kernels
    int mb = 0;
    for (int s = s_first; s < s_last; ++s) {
        const SceneDesc& S = descs[s];
        uint64_t n3 = 3ull * S.npix;
        if (!adj)
            k_zero_u32<<<std::min<uint64_t>((n3 + 255) / 256, 4096), 256, 0, st>>>(reinterpret_cast<uint32_t*>(S.image),
                                                                                 n3);
        mb = std::max(mb, S.max_bounce);
    }
    const int ns = (int)descs.size();
    const bool one = s_last - s_first == 1;
    const SceneDesc* sd = one ? d_scenes + s_first : d_scenes;
    for (uint64_t g0 = off[s_first]; g0 < off[s_last]; g0 += B) {
        const uint32_t cnt = (uint32_t)std::min<uint64_t>(B, off[s_last] - g0);
        P.g0 = g0;   // launches take P by value: this batch's kernels see this batch's g0
        const uint32_t cgrid = std::min<uint32_t>((cnt + 255) / 256, 65535);
        // One wavefront over the batch: camera rays, then max_bounce + 1 trace/shade rounds.
        // The replay (kShadeReplay) scatters each bounce's gradients as it shades it, reads the
        // paths' radiance from recL + rbase and takes its camera rays' hits from recH + rbase;
        // the forward wavefront writes its camera rays' hits to recH + rbase when recH is set.
        auto wavefront = [&](int mode, const float4* recL, float4* recH, uint64_t rbase) {
            k_set_counts<<<1, 64, 0, st>>>(P.cnt, ncnt, cnt);
            k_camera<<<cgrid, 256, 0, st>>>(d_scenes, d_off, ns, g0, cnt, P);
            // bounce b traces queue q[qi] of cnt[b] paths; shade appends the survivors to
            // q[qi^1], counting them in cnt[b+1]; the persistent trace kernel deals from
            // cnt[ncnt/2 + b]
            int qi = 0;
            for (int b = 0; b <= mb; ++b) {
                const uint32_t loop = g_tuning.render_loop == 0 ? (uint32_t)kLoopSingle
                                      : g_tuning.render_leaf_coop ? (uint32_t)kLoopCoop : (uint32_t)kLoopWhileWhile;
                // The persistent kernel hands each wave kRenderChunk paths per atomic: a batch
                // with fewer paths than its waves can take in one round would leave most of
                // the GPU idle, so such batches (the single scenes of a plain C3 render, say)
                // take the chunked kernel, 64 paths per wave.
                const bool persist = g_tuning.render_loop == 2 &&
                                     (uint64_t)cnt >= (uint64_t)pgrid * (kTraceThreads / 64) * kRenderChunk;
                if (mode == kShadeReplay && b == 0) {
                    k_load_hit<<<cgrid, 256, 0, st>>>(P, cnt, recH, rbase);
                } else if (persist) {
                    uint32_t* work = P.cnt + ncnt / 2 + b;
                    with_trace_variant(g_tuning.render_lds_depth, loop, [&](auto D, auto L) {
                        constexpr int DD = decltype(D)::value, LL = decltype(L)::value;
                        if (one)
                            k_wtrace_persist<DD, LL, kSceneUni>
                                <<<pgrid, kTraceThreads, 0, st>>>(d_scenes + s_first, 1u, P, qi, b, wait, work);
                        else if ((uint32_t)ns <= kLdsScenes)
                            k_wtrace_persist<DD, LL, kSceneLds>
                                <<<pgrid, kTraceThreads, 0, st>>>(d_scenes, (uint32_t)ns, P, qi, b, wait, work);
                        else
                            k_wtrace_persist<DD, LL, kSceneGlobal>
                                <<<pgrid, kTraceThreads, 0, st>>>(d_scenes, (uint32_t)ns, P, qi, b, wait, work);
                    });
                } else {
                    with_trace_variant(g_tuning.render_lds_depth, loop, [&](auto D, auto L) {
                        constexpr int DD = decltype(D)::value, LL = decltype(L)::value;
                        if (one)
                            k_wtrace<DD, LL, kSceneUni>
                                <<<grid, kTraceThreads, 0, st>>>(d_scenes + s_first, 1u, P, qi, b, wait);
                        else if ((uint32_t)ns <= kLdsScenes)
                            k_wtrace<DD, LL, kSceneLds>
                                <<<grid, kTraceThreads, 0, st>>>(d_scenes, (uint32_t)ns, P, qi, b, wait);
                        else
                            k_wtrace<DD, LL, kSceneGlobal>
                                <<<grid, kTraceThreads, 0, st>>>(d_scenes, (uint32_t)ns, P, qi, b, wait);
                    });
                }
                if (mode == kShadeReplay) {
                    // the gradient scatter rides on the replay's shading: both arrays in LDS
                    // when they fit, global atomics otherwise
                    const GradArgs G{recL, dimg, ga, ge, descs[0].nprim, nullptr, rbase};
                    const size_t lds = ((ga ? 1u : 0u) + (ge ? 1u : 0u)) * 3ull * descs[0].nprim * sizeof(float);
                    if (lds <= kLdsGradBytes) {
                        const uint32_t gb = std::min<uint32_t>(kGradLdsBlocks, (cnt + kGradThreads - 1) / kGradThreads);
                        if (one) k_shade<kShadeReplay, true, kGradLds><<<gb, kGradThreads, lds, st>>>(sd, P, qi, b, G);
                        else k_shade<kShadeReplay, false, kGradLds><<<gb, kGradThreads, lds, st>>>(sd, P, qi, b, G);
                    } else {
                        if (one) k_shade<kShadeReplay, true, kGradGlobal><<<grid, sh, 0, st>>>(sd, P, qi, b, G);
                        else k_shade<kShadeReplay, false, kGradGlobal><<<grid, sh, 0, st>>>(sd, P, qi, b, G);
                    }
                } else {
                    const GradArgs G{nullptr, nullptr, nullptr, nullptr, 0, b == 0 ? recH : nullptr, rbase};
                    if (one) k_shade<kShadeForward, true><<<grid, sh, 0, st>>>(sd, P, qi, b, G);
                    else k_shade<kShadeForward, false><<<grid, sh, 0, st>>>(sd, P, qi, b, G);
                }
                qi ^= 1;
            }
        };
        // the caller's records (one scene): off[s_last] = the scene's paths
        float4* recH = rec ? rec + off[s_last] : nullptr;
        if (adj == 0) {
            wavefront(kShadeForward, nullptr, recH, g0);
            if (rec) k_store_L<<<cgrid, 256, 0, st>>>(P, cnt, rec, g0);
            k_accum<<<cgrid, 256, 0, st>>>(d_scenes, d_off, ns, g0, cnt, P);
        } else if (adj == 1) {
            wavefront(kShadeForward, nullptr, recb + B, 0);
            k_store_L<<<cgrid, 256, 0, st>>>(P, cnt, recb, 0);
            wavefront(kShadeReplay, recb, recb + B, 0);
        } else {
            wavefront(kShadeReplay, rec, recH, g0);
        }
    }
    if (!adj)
        for (int s = s_first; s < s_last; ++s)
            k_finalize<<<std::min<uint32_t>((3 * descs[s].npix + 255) / 256, 4096), 256, 0, st>>>(d_scenes, s);
}

int Batch::launch(hipStream_t st) {
    if (mode == 1) {
        RTX_HIP(hipGraphLaunch(exec, st));
        return note_launch(st);
    }
    if (mode == 2) {
        enqueue_range(0, (int)descs.size(), st);
    } else {
        for (int s = 0; s < (int)descs.size(); ++s) enqueue_range(s, s + 1, st);
    }
    RTX_HIP(hipGetLastError());
    return note_launch(st);
}

}  // namespace rtx

using namespace rtx;

struct rtx_batch {
    rtx::Batch b;
};

extern "C" int rtx_set_tuning(const rtx_tuning* t) {
    if (!t || t->batch_paths < 1024 || t->grid_blocks == 0 || t->shade_threads == 0 || t->shade_threads % 64 ||
        t->shade_threads > 1024 || !(t->trace_lds_depth == 8 || t->trace_lds_depth == 12 || t->trace_lds_depth == 16) ||
        t->trace_loop > 2 || t->ray_sort > 1 || t->wait_lanes > 64 || t->trace_grid_blocks == 0 ||
        !(t->render_lds_depth == 8 || t->render_lds_depth == 12 || t->render_lds_depth == 16) || t->render_loop > 2 ||
        t->render_wait_lanes > 64 || t->leaf_coop > 1 || t->render_leaf_coop > 1 || t->refill_lanes == 0 ||
        t->refill_lanes > 64 || t->trace_tree > 1)
        return fail(RTX_E_ARG, "bad tuning on rtx_set_tuning()");
    g_tuning = *t;
    return RTX_OK;
}

extern "C" int rtx_get_tuning(rtx_tuning* t) {
    if (!t) return fail(RTX_E_ARG, "null argument on rtx_get_tuning()");
    *t = g_tuning;
    return RTX_OK;
}

extern "C" int rtx_batch_create(const rtx_bvh* const* scenes, int n, const rtx_camera* cams,
                                const rtx_render_params* params, const float* const* albedo_d,
                                const float* const* emission_d, float* const* images_d, int mode, rtx_batch** out) {
    if (!scenes || n <= 0 || !cams || !params || !images_d || !out || mode < 0 || mode > 2)
        return fail(RTX_E_ARG, "bad argument on rtx_batch_create()");
    *out = nullptr;
    rtx_batch* bt = new rtx_batch;
    Batch& B = bt->b;
    B.mode = mode;
    for (int s = 0; s < n; ++s) {
        int rc = check_scene(scenes[s], &cams[s], &params[s], images_d[s], "rtx_batch_create()");
        if (rc) { delete bt; return rc; }
        B.descs.push_back(make_desc(scenes[s], &cams[s], &params[s], albedo_d ? albedo_d[s] : nullptr,
                                    emission_d ? emission_d[s] : nullptr, images_d[s]));
    }
    // A batch may be launched on any stream: its setup runs on a stream of its own and is
    // complete when creation returns.
    hipStream_t cs;
    if (!hip_ok(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream")) { delete bt; return RTX_E_HIP; }
    int rc = B.setup(g_tuning.batch_paths, cs);
    bool ok = hip_ok(hipStreamSynchronize(cs), "batch setup");
    B.sst = nullptr;
    if (rc || !ok) {
        (void)hipStreamDestroy(cs);
        delete bt;
        return rc ? rc : RTX_E_HIP;
    }
    if (mode == 1) {
        ok = hip_ok(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
        if (ok) {
            B.enqueue_range(0, n, cs);
            ok = hip_ok(hipStreamEndCapture(cs, &B.graph), "end capture");
        }
        if (ok) ok = hip_ok(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0), "instantiate");
    }
    (void)hipStreamDestroy(cs);
    if (!ok) { delete bt; return RTX_E_HIP; }
    *out = bt;
    return RTX_OK;
}

extern "C" int rtx_batch_launch(rtx_batch* bt, void* stream) {
    if (!bt) return fail(RTX_E_ARG, "null argument on rtx_batch_launch()");
    return bt->b.launch((hipStream_t)stream);
}

extern "C" void rtx_batch_free(rtx_batch* bt) { delete bt; }

extern "C" int rtx_render_batched(const rtx_bvh* const* scenes, int n, const rtx_camera* cams,
                                  const rtx_render_params* params, const float* const* albedo_d,
                                  const float* const* emission_d, float* const* images_d, int mode, void* stream) {
    rtx_batch* bt = nullptr;
    int rc = rtx_batch_create(scenes, n, cams, params, albedo_d, emission_d, images_d, mode, &bt);
    if (rc) return rc;
    rc = rtx_batch_launch(bt, stream);
    if (rc == RTX_OK && !hip_ok(hipStreamSynchronize((hipStream_t)stream), "sync")) rc = RTX_E_HIP;
    rtx_batch_free(bt);
    return rc;
}

namespace {
// The single-scene entry points take their path state from the scratch cache and give it
// back as soon as their work is queued, so their launches cannot be captured into a graph
// (a replay would use blocks that belong to later calls): a capturing stream is refused.
// rtx_batch_create's mode 1 is the graph form.
int refuse_capture(hipStream_t st, const char* fn) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    RTX_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return fail(RTX_E_STATE, std::string("stream capture of ") + fn + " (use rtx_batch_create mode 1 for a graph)");
    return RTX_OK;
}

// One scene's render, queued on `stream` and not waited for; the path state goes back to the
// scratch cache ordered after it.
int render_one(const rtx_bvh* b, const rtx_camera* cam, const rtx_render_params* p, const float* albedo_d,
               const float* emission_d, float* image_d, float* path_records_d, void* stream) {
    Batch B;
    B.descs.push_back(make_desc(b, cam, p, albedo_d, emission_d, image_d));
    B.rec = reinterpret_cast<float4*>(path_records_d);
    hipStream_t st = (hipStream_t)stream;
    int rc = refuse_capture(st, "a render");
    if (rc) return rc;
    rc = B.setup(g_tuning.batch_paths, st);
    if (rc) return rc;
    B.enqueue_range(0, 1, st);
    rc = B.note_launch(st);
    if (rc) return rc;
    return hip_ok(hipGetLastError(), "render launch") ? RTX_OK : RTX_E_HIP;
}
}  // namespace

extern "C" int rtx_render(const rtx_bvh* b, const rtx_camera* cam, const rtx_render_params* p, const float* albedo_d,
                          const float* emission_d, float* image_d, void* stream) {
    int rc = check_scene(b, cam, p, image_d, "rtx_render()");
    if (rc) return rc;
    return render_one(b, cam, p, albedo_d, emission_d, image_d, nullptr, stream);
}

extern "C" int rtx_render_primal(const rtx_bvh* b, const rtx_camera* cam, const rtx_render_params* p,
                                 const float* albedo_d, const float* emission_d, float* image_d,
                                 float* path_records_d, void* stream) {
    int rc = check_scene(b, cam, p, image_d, "rtx_render_primal()");
    if (rc) return rc;
    if (!path_records_d) return fail(RTX_E_ARG, "null path_records_d on rtx_render_primal()");
    if (reinterpret_cast<uintptr_t>(path_records_d) % 16)
        return fail(RTX_E_ARG, "path_records_d is not 16-byte aligned on rtx_render_primal()");
    return render_one(b, cam, p, albedo_d, emission_d, image_d, path_records_d, stream);
}

namespace {
int adjoint_impl(const rtx_bvh* b, const rtx_camera* cam, const rtx_render_params* p, const float* albedo_d,
                 const float* emission_d, const float* path_records_d, const float* d_image_d, float* grad_albedo_d,
                 float* grad_emission_d, void* stream, const char* fn) {
    int rc = check_scene(b, cam, p, d_image_d, fn);
    if (rc) return rc;
    if (reinterpret_cast<uintptr_t>(path_records_d) % 16)
        return fail(RTX_E_ARG, std::string("path_records_d is not 16-byte aligned on ") + fn);
    if (!albedo_d && !p->tex_d) return fail(RTX_E_ARG, std::string("albedo is required on ") + fn);
    if (p->tex_d && grad_emission_d)
        return fail(RTX_E_ARG, std::string("emission gradients are per triangle: not with texture albedo on ") + fn);
    if (p->rr_bounce <= p->max_bounce)
        return fail(RTX_E_ARG, "the adjoint needs rr_bounce > max_bounce (no Russian roulette)");
    if (p->max_bounce > 64) return fail(RTX_E_LIMIT, std::string("max_bounce above 64 on ") + fn);
    Batch B;
    B.descs.push_back(make_desc(b, cam, p, albedo_d, emission_d, nullptr));
    B.adj = path_records_d ? 2 : 1;
    B.rec = reinterpret_cast<float4*>(const_cast<float*>(path_records_d));
    B.dimg = d_image_d;
    B.ga = grad_albedo_d;
    B.ge = emission_d ? grad_emission_d : nullptr;
    // the state per path does not depend on max_bounce: the path state and one path record;
    // queued on the stream and not waited for, like render_one
    hipStream_t st = (hipStream_t)stream;
    rc = refuse_capture(st, fn);
    if (rc) return rc;
    rc = B.setup(g_tuning.batch_paths, st, true);
    if (rc) return rc;
    B.enqueue_range(0, 1, st);
    rc = B.note_launch(st);
    if (rc) return rc;
    return hip_ok(hipGetLastError(), "adjoint launch") ? RTX_OK : RTX_E_HIP;
}
}  // namespace

extern "C" int rtx_render_adjoint(const rtx_bvh* b, const rtx_camera* cam, const rtx_render_params* p,
                                  const float* albedo_d, const float* emission_d, const float* d_image_d,
                                  float* grad_albedo_d, float* grad_emission_d, void* stream) {
    return adjoint_impl(b, cam, p, albedo_d, emission_d, nullptr, d_image_d, grad_albedo_d, grad_emission_d, stream,
                        "rtx_render_adjoint()");
}

extern "C" int rtx_render_adjoint_replay(const rtx_bvh* b, const rtx_camera* cam, const rtx_render_params* p,
                                         const float* albedo_d, const float* emission_d, const float* path_records_d,
                                         const float* d_image_d, float* grad_albedo_d, float* grad_emission_d,
                                         void* stream) {
    if (!path_records_d) return fail(RTX_E_ARG, "null path_records_d on rtx_render_adjoint_replay()");
    return adjoint_impl(b, cam, p, albedo_d, emission_d, path_records_d, d_image_d, grad_albedo_d, grad_emission_d,
                        stream, "rtx_render_adjoint_replay()");
}
