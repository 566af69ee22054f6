/* librtx: MI355X-native BVH build, traversal and path tracing for the
 * minsuk0519/ray-tracer hot path.  C ABI: plain pointers and sizes.
 *
 * Conventions
 *  - Return 0 on success, a negative RTX_E_* code on failure; rtx_last_error()
 *    gives the message (thread-local).  Mirrors the reference's bool + "Error : ..."
 *    convention (e.g. src/bvh/bvh_traverse.cpp:187-241).
 *  - Vertices use the reference Vertex layout: 6 floats (x,y,z,nx,ny,nz) per vertex
 *    (src/bvh/bvh.hpp:32-41).  Triangles are 3 uint32 vertex indices (bvh.hpp:10-13).
 *  - "_d" pointers are device (HBM) pointers; stream is a hipStream_t passed as void*
 *    (NULL = default stream).  The library owns rtx_bvh handles; the caller owns ray,
 *    hit and image buffers.  Calls on one handle are not re-entrant across threads.
 */
#ifndef RTX_H
#define RTX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RTX_OK = 0,
    RTX_E_ARG = -1,      /* invalid argument (null pointer, zero count, bad index) */
    RTX_E_HIP = -2,      /* HIP runtime error */
    RTX_E_IO = -3,       /* file missing, truncated or unwritable */
    RTX_E_VERSION = -4,  /* .bvh version mismatch */
    RTX_E_LIMIT = -5,    /* size beyond a documented limit */
    RTX_E_STATE = -6     /* object not ready (e.g. trace before load) */
};

typedef struct rtx_bvh rtx_bvh;

/* 32-byte ray: replaces bvh::Ray (src/bvh/bvh_traverse.hpp:10-16; 48 B there). */
typedef struct rtx_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} rtx_ray;

/* 16-byte closest hit: replaces bvh::Hit (bvh_traverse.hpp:18-26).  tri is the
 * reference Hit.triIndex (index into the reordered triangle array), 0xFFFFFFFF on a
 * miss (the reference's hit == false).  Normals are an optional separate output. */
typedef struct rtx_hit {
    float t, u, v;
    uint32_t tri;
} rtx_hit;

typedef struct rtx_bvh_info {
    uint32_t node_count;        /* reference BVHNode count (DFS order) */
    uint32_t tri_count;         /* reordered triangles, spatial-split duplicates included */
    uint32_t vert_count;
    uint32_t input_tri_count;
    uint32_t internal_count;
    uint32_t leaf_count;
    uint32_t max_depth;
    uint32_t tri_index_size;    /* s_triIndexSize at the end of the build */
    uint32_t sorted_tris_size;  /* s_sortedTrisSize at the end of the build */
    uint32_t levels;
    float scene_min[3];
    float scene_max[3];
} rtx_bvh_info;

const char* rtx_last_error(void);
int rtx_device_count(void);

/* ---- build -------------------------------------------------------------------
 * In-memory bake: replaces bvh::bakeBVH's pipeline (src/bvh/bvh_bake.cpp:180-194:
 * sortTrisByMorton, buildCentroidCache, initRoot, bfsLoop, reorderNodes,
 * reorderTriangles) on the GPU.  verts6/tris are host pointers. */
int rtx_build(const float* verts6, uint32_t nv, const uint32_t* tris, uint32_t nt, rtx_bvh** out);
/* Same, from device-resident inputs (verts6_d, tris_d stay owned by the caller). */
int rtx_build_device(const float* verts6_d, uint32_t nv, const uint32_t* tris_d, uint32_t nt,
                     void* stream, rtx_bvh** out);
void rtx_bvh_free(rtx_bvh* bvh);
int rtx_bvh_get_info(const rtx_bvh* bvh, rtx_bvh_info* info);
/* Copy the tree to host in the reference's layout (BVHNode 48 B each, DFS order;
 * Triangle 3 x uint32; original triangle id per reordered triangle).  Any pointer may
 * be NULL.  Replaces the in-memory state writeBakedData serialises (bvh_IO.cpp:253-264). */
int rtx_bvh_export(const rtx_bvh* bvh, void* nodes48, uint32_t* tris3, uint32_t* prims);
/* Morton codes per input triangle and the sorted permutation (s_sortedTris[0..N)),
 * from sortTrisByMorton (src/bvh/bvh_morton.cpp:40-159).  Either pointer may be NULL. */
int rtx_bvh_morton(const rtx_bvh* bvh, uint64_t* codes, uint32_t* sorted);
/* Node array before DFS renumbering (the s_nodes order of bfsLoop, bvh_bfs.cpp:72-134). */
int rtx_bvh_export_bfs(const rtx_bvh* bvh, void* nodes48, uint32_t* count);
/* Host run of the builder's register-only clip (clipTriangleToAABB, src/bvh/AABB.cpp:86-171,
 * when only the two planes of `axis` cut the triangle), for the parity tests: returns 1 and
 * the box in out6 (min xyz, max xyz) when that path applies, 0 when the builder would run
 * the full clip instead.  No GPU work. */
int rtx_debug_clip_slab(const float* v9, const float* region6, int axis, float* out6);
/* The traversal's reciprocal (v_rcp_f32 and one FMA Newton step where 2^-126 <= |x| < 2^126,
 * the IEEE division elsewhere) against the IEEE quotient 1.0f / x, on the GPU, for the float
 * bit patterns [first, first + count) (first + count <= 2^32): the number of inputs whose
 * result bits differ, NaN results counted as equal, and the lowest such pattern
 * (0xFFFFFFFF when none).  Pins the Moller-Trumbore f = 1/a of bvh_traverse.cpp:62 and
 * the 1/dir of rayAABB (:26-28) to the reference's division.  Synchronous. */
int rtx_debug_recip_check(uint64_t first, uint64_t count, uint64_t* mismatches, uint32_t* first_bad);
/* Peak bytes of device scratch (build scratch, render path state) held by library calls at
 * once since the previous call with reset != 0, which starts a new window at the bytes held
 * now.  For the tests that check the adjoint's memory does not grow with max_bounce.  No
 * reference counterpart. */
int rtx_debug_scratch_peak(uint64_t* peak_bytes, int reset);

/* ---- trace ------------------------------------------------------------------
 * Closest hit for n rays: replaces bvh::trace (src/bvh/bvh_traverse.cpp:237-246,
 * traverse :91-178).  rays_d/hits_d device pointers; normals_d optional (3 floats per
 * ray, the interpolated shading normal of bvh_traverse.cpp:135-139).  On every trace
 * entry point n = 0 does nothing and the buffers may then be NULL.  Calls of at most
 * 16,384 rays (closest and any hit) walk each ray's tree breadth-first with one wave per
 * ray, which keeps the reference's hit, including which triangle wins an exact-t tie. */
int rtx_trace_closest(const rtx_bvh* bvh, const rtx_ray* rays_d, size_t n, rtx_hit* hits_d,
                      float* normals_d, void* stream);
/* rtx_trace_closest without allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same kernels and results as rtx_trace_closest; the
 * persistent kernel's ray counter and the ray sort's buffers live in a fixed file-scope
 * device array (about 34 MB, held by each GPU the library runs on until the process ends)
 * instead of on the handle.  n <= 2097152 (RTX_E_LIMIT above); the BVH's device must be
 * current.  Every call on one GPU shares that array: calls that overlap in time on one
 * GPU (other streams or threads) race and must be ordered by the caller. */
int rtx_trace_closest_noalloc(const rtx_bvh* bvh, const rtx_ray* rays_d, size_t n, rtx_hit* hits_d,
                              float* normals_d, void* stream);
/* Any hit (occlusion): occluded_d[i] = 1 iff trace(rays[i]).hit.  No reference
 * counterpart (the reference only has closest hit); pinned by that identity. */
int rtx_trace_any(const rtx_bvh* bvh, const rtx_ray* rays_d, size_t n, uint8_t* occluded_d, void* stream);
/* Instrumented closest hit: counts_d[2*i] = internal-node visits, [2*i+1] = triangle
 * tests of ray i (the per-ray work the roofline's algorithmic bytes are priced on). */
int rtx_trace_counts(const rtx_bvh* bvh, const rtx_ray* rays_d, size_t n, uint32_t* counts_d, void* stream);
/* Host-pointer convenience wrapper around rtx_trace_closest (copies in and out). */
int rtx_trace_closest_host(const rtx_bvh* bvh, const rtx_ray* rays, size_t n, rtx_hit* hits, float* normals);

/* ---- files and scenes ----------------------------------------------------------
 * .bvh file: version-1 header + raw BVHNode/Triangle/Vertex arrays, bit-compatible
 * with writeBakedData (src/bvh/bvh_IO.cpp:236-275) and loadBVH (bvh_traverse.cpp:182-235). */
int rtx_save(const rtx_bvh* bvh, const char* path);
int rtx_load(const char* path, rtx_bvh** out);
/* Parse a .scn file's geometry lines (sphere / box / mesh) into host arrays, as
 * readScene + initGeos do (src/bvh/bvh_IO.cpp:72-234, bvh_geo.cpp:22-192).  `mesh` reads
 * ASCII PLY in place of the assimp import of loadMesh (bvh_IO.cpp:24-70); other mesh
 * formats fail with RTX_E_IO.  Free the arrays with rtx_free_host. */
int rtx_scene_load(const char* scn_path, float** verts6, uint32_t* nv, uint32_t** tris, uint32_t* nt);
void rtx_free_host(void* p);
/* Radiance .hdr files (the stbi_write_hdr output and stbi_loadf input of main.cpp:95-104,
 * 378).  Write: rgb is width*height*3 floats; flip_y != 0 writes row height-1 first, as
 * main.cpp:367-376 does for its bottom-up image.  Read: flat and run-length encoded
 * scanlines; *rgb is allocated (free with rtx_free_host), top row first. */
int rtx_hdr_write(const char* path, int width, int height, const float* rgb, int flip_y);
int rtx_hdr_read(const char* path, int* width, int* height, float** rgb);
/* Drop-in for bvh::bakeBVH(scenePath) (bvh_bake.cpp:180-194): scene -> GPU build ->
 * <scene>.bvh next to it. */
int rtx_bake_scene(const char* scn_path);
/* Reference geometry generators addSphere / addBox (src/bvh/bvh_geo.cpp:22-192).
 * Pass NULL outputs to query counts.  quat is (x, y, z, w). */
int rtx_geo_sphere(const float center[3], float radius, int rings, int sectors,
                   float* verts6, uint32_t* nv, uint32_t* tris, uint32_t* nt);
int rtx_geo_box(const float center[3], const float half[3], const float quat_xyzw[4],
                float* verts6, uint32_t* nv, uint32_t* tris, uint32_t* nt);

/* ---- path tracing --------------------------------------------------------------- */
typedef struct rtx_camera {
    float pos[3], fwd[3], right[3], up[3];
    float tan_half_fov_y;
} rtx_camera;

typedef struct rtx_render_params {
    int width, height, spp;
    int max_bounce;   /* segments per path = max_bounce + 1 (src/main.cpp:156) */
    int rr_bounce;    /* Russian roulette when bounce >= rr_bounce (reference: 3, main.cpp:188) */
    uint32_t seed;    /* sample-table seed */
    int env_mode;     /* 0 = sky gradient (main.cpp:169-172), 1 = constant env_color,
                         2 = equirectangular IBL map ibl_d (IBL::sample, main.cpp:107-117),
                         3 = env_color plus sun_radiance inside the cone dot(d, sun_dir) >= sun_cos */
    float env_color[3];
    const float* ibl_d;   /* env_mode 2: ibl_width * ibl_height * 3 floats in HBM, top row first */
    int ibl_width, ibl_height;
    /* Texture albedo (optional, an extension): uv_d holds 2 floats per input vertex, tex_d a
     * tex_width x tex_height RGB texture (texel (x, y) at 3 * (y * tex_width + x)). When tex_d
     * is set, the albedo at a hit is the nearest texel at the uv interpolated like the shading
     * normal; the albedo argument is unused and rtx_render_adjoint's grad_albedo is per texel. */
    const float* uv_d;
    const float* tex_d;
    int tex_width, tex_height;
    float sun_dir[3];       /* env_mode 3: a distant sun seen by the rays that reach its cone */
    float sun_cos;
    float sun_radiance[3];
} rtx_render_params;

/* Wavefront path tracer: replaces the raster loop + pathTrace (src/main.cpp:145-204,
 * 339-365) with every random number drawn from the counter-based sample table.
 * albedo_d / emission_d: 3 floats per ORIGINAL triangle (NULL = reference constant 0.8 /
 * no emission).  image_d: width*height*3 floats, row y = 0 at the bottom like the
 * reference's pre-flip buffer.  rtx_render, rtx_render_primal, rtx_render_adjoint and
 * rtx_render_adjoint_replay queue their work on `stream` and return without waiting for it
 * (the parameters are copied at the call); the caller synchronizes before reading results
 * on the host. */
int rtx_render(const rtx_bvh* bvh, const rtx_camera* cam, const rtx_render_params* p,
               const float* albedo_d, const float* emission_d, float* image_d, void* stream);

/* rtx_render that also stores each path's record (the primal of the path-replay adjoint):
 * path_records_d (16-byte aligned) holds 8 floats per path in two halves, N = width*height*
 * spp paths: path g = sample*width*height + k has its radiance L at 4g..4g+2 (4g+3 is 0) and
 * the hit of its camera ray at 4N+4g..4N+4g+3 (t, u, v, triangle index bits; 0xFFFFFFFF on a
 * miss).  Slot k of a sample is a pixel in 8 x 8 tile order: bands of 8 rows from y = 0,
 * tiles of 8 columns from x = 0 along a band, row-major inside a tile (the last band and a
 * band's last tile may be narrower).  32 bytes per path whatever max_bounce is; with the
 * camera rays' hits stored the replay does not trace them again.  The image equals
 * rtx_render's.  No reference counterpart. */
int rtx_render_primal(const rtx_bvh* bvh, const rtx_camera* cam, const rtx_render_params* p,
                      const float* albedo_d, const float* emission_d, float* image_d,
                      float* path_records_d, void* stream);

/* Path-replay adjoint of rtx_render for the loss sum (I - target)^2: accumulates
 * dLoss/dalbedo and dLoss/demission (3 floats per original triangle, added to the
 * buffers) given d_image_d = dLoss/dI.  Paths are re-traced with the same sample table,
 * not stored: per batch of paths it runs a primal pass for each path's radiance L and
 * camera-ray hit (a 32-byte record per path), then the replay, which at the k-th surface hit of a path adds dL/da_k = (L - L_upto_k) / a_k
 * and dL/dLe_k = throughput arriving at k, where L_upto_k is the radiance gathered up to
 * and including that hit.  Only the current bounce is held per path, so memory does not
 * grow with max_bounce.  Requires rr_bounce > max_bounce (no Russian roulette) so the path
 * set does not depend on albedo.  The albedo gradient is exact for albedo > 0; a path
 * vertex on a zero albedo channel adds 0 to that channel (the replay holds no radiance past
 * it), so an optimiser should keep albedo in (0, 1].  No reference counterpart. */
int rtx_render_adjoint(const rtx_bvh* bvh, const rtx_camera* cam, const rtx_render_params* p,
                       const float* albedo_d, const float* emission_d, const float* d_image_d,
                       float* grad_albedo_d, float* grad_emission_d, void* stream);
/* The replay of rtx_render_adjoint alone, with each path's record (radiance and camera-ray
 * hit) taken from an earlier rtx_render_primal of the same scene, camera, parameters and
 * materials (so the primal pass and the camera rays' trace are saved).  No reference
 * counterpart. */
int rtx_render_adjoint_replay(const rtx_bvh* bvh, const rtx_camera* cam, const rtx_render_params* p,
                              const float* albedo_d, const float* emission_d, const float* path_records_d,
                              const float* d_image_d, float* grad_albedo_d, float* grad_emission_d, void* stream);

/* Batched wavefront render of many scenes.  scenes[i] renders with cams[i]/params[i]
 * into images_d[i].  mode 0 = one launch sequence per scene (plain), 1 = all scenes in
 * shared persistent launches captured once in a hipGraph (batched).  No reference
 * counterpart (the reference renders one scene per process). */
int rtx_render_batched(const rtx_bvh* const* scenes, int nscenes, const rtx_camera* cams,
                       const rtx_render_params* params, const float* const* albedo_d,
                       const float* const* emission_d, float* const* images_d, int mode, void* stream);

/* Prepared multi-scene render.  mode 0 = one launch sequence per scene (plain),
 * 1 = all scenes share every launch and the whole sequence is captured once into a
 * hipGraph at creation (batched), 2 = shared launches without the graph.  Buffers passed
 * at creation must stay valid; launch may be repeated (each launch re-renders). */
typedef struct rtx_batch rtx_batch;
int rtx_batch_create(const rtx_bvh* const* scenes, int nscenes, const rtx_camera* cams,
                     const rtx_render_params* params, const float* const* albedo_d,
                     const float* const* emission_d, float* const* images_d, int mode, rtx_batch** out);
int rtx_batch_launch(rtx_batch* batch, void* stream);
void rtx_batch_free(rtx_batch* batch);

/* Launch shapes (process-wide; chosen by the project autotuner; none changes a result).
 *
 * Trace entry points (rtx_trace_closest / _any / _counts / _closest_host):
 *   trace_lds_depth    LDS stack entries per lane: 8, 12 or 16;
 *   trace_loop         0 = one node/leaf loop, 1 = while-while (triangle tests run with the
 *                      wave mostly converged), 2 = while-while in a persistent grid of
 *                      trace_grid_blocks blocks where idle lanes fetch new rays (calls
 *                      with fewer than 4096 rays run loop 1). Closest- and any-hit calls
 *                      of at most 16,384 rays take the one-wave-per-ray walk instead,
 *                      whatever this field (rtx_trace_counts never does);
 *   ray_sort           1 = sort rays (>= 65536 of them) by a Morton key of their origins
 *                      and trace in that order, writing results back at each ray's index.
 *                      The sort's scratch lives on the BVH handle and grows to the largest
 *                      ray count seen. Trace calls that use the handle's scratch (sorting
 *                      or trace_loop 2) are ordered on the GPU: each call's stream waits for
 *                      the previous such call on that handle, from any stream or thread;
 *   wait_lanes         0..64: the while-while leaf phase starts once at most this many lanes
 *                      of a wave are still at internal nodes (0 = wait for all);
 *   trace_grid_blocks  blocks of the persistent trace kernel;
 *   refill_lanes       1..64: a wave of the persistent trace kernel fetches new rays once
 *                      at least this many of its lanes are idle (incoherent closest-hit
 *                      rays are fastest near 18, coherent and any-hit rays near 32);
 *   trace_tree         0 = walk the binary tree in the reference's visiting order (hits equal
 *                      the reference's, including which triangle wins an exact-t tie);
 *                      1 = walk its 4-wide collapse (128-byte nodes of four child boxes, built
 *                      with every tree): the same t for every ray, and at an exact-t tie
 *                      between triangles of different leaves possibly another of the tied
 *                      triangles.
 * Renderer (rtx_render, rtx_render_adjoint, rtx_batch_*):
 *   batch_paths        paths in flight per wavefront batch (>= 1024; default 2^26, 92
 *                      bytes of path state per path, 32 more for the adjoint's records);
 *   grid_blocks        blocks of the wavefront trace and shade kernels;
 *   shade_threads      threads per shade block (multiple of 64, <= 1024);
 *   render_lds_depth   LDS stack entries per lane: 8, 12 or 16;
 *   render_loop        0 or 1 as trace_loop; 2 = while-while in a persistent grid of
 *                      trace_grid_blocks blocks whose waves refill idle lanes from the
 *                      bounce queue, for batches of at least trace_grid_blocks * 512
 *                      paths (smaller batches take the loop of 1);
 *   render_wait_lanes  as wait_lanes.
 * Both:
 *   leaf_coop          1 = the while-while loops of the trace entry points (trace_loop
 *                      1 and 2) run cooperative leaf phases: the triangle tests
 *                      of all lanes waiting at leaves are spread over the whole wave, 64 per
 *                      round, instead of each lane testing its own leaf;
 *   render_leaf_coop   the same for the renderer's while-while loops (render_loop 1 and 2). */
typedef struct rtx_tuning {
    uint32_t batch_paths;
    uint32_t grid_blocks;
    uint32_t shade_threads;
    uint32_t trace_lds_depth;
    uint32_t trace_loop;
    uint32_t ray_sort;
    uint32_t wait_lanes;
    uint32_t trace_grid_blocks;
    uint32_t render_lds_depth;
    uint32_t render_loop;
    uint32_t render_wait_lanes;
    uint32_t leaf_coop;
    uint32_t render_leaf_coop;
    uint32_t refill_lanes;
    uint32_t trace_tree;
} rtx_tuning;
int rtx_set_tuning(const rtx_tuning* t);
int rtx_get_tuning(rtx_tuning* t);

/* Kernel timing for measurement tools (no counterpart in the reference). Between
 * rtx_profile_begin() and rtx_profile_end(), every trace entry point records a HIP event
 * pair on the caller's stream around its traversal kernel alone (ray sorting excluded).
 * rtx_profile_end waits for the last pair and returns the summed kernel milliseconds and
 * the number of launches, then stops recording. Process-wide; calls from several threads
 * in that window are serialized. */
int rtx_profile_begin(void);
int rtx_profile_end(double* kernel_ms, uint32_t* launches);

#ifdef __cplusplus
}
#endif
#endif
