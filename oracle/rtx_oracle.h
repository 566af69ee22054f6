/* CPU oracle for the ray-tracer hot path (BVH build, traversal, path-trace loop).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * (librtx) is compared against. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product never links or calls it.
 *
 * Every function restates the reference algorithm serially, with the reference's
 * float operation order (no FMA, IEEE divide/sqrt, SSE min/max and std::min/max
 * NaN/tie semantics). Citations are to minsuk0519/ray-tracer source files.
 */
#ifndef RTX_ORACLE_H
#define RTX_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference BVHNode (src/bvh/bvh.hpp:15-30): 48 bytes.
 * AABB as two 16-byte vec3 (w lane carried), isLeaf at byte 32, union at 36. */
typedef struct orc_node48 {
    float    bmin[4];
    float    bmax[4];
    uint8_t  is_leaf;
    uint8_t  pad0[3];
    uint32_t a;   /* childID[0] or beginTriIndex */
    uint32_t b;   /* childID[1] or triSize */
    uint32_t pad1;
} orc_node48;

/* 32-byte ray used at every C boundary of this project. */
typedef struct orc_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} orc_ray;

/* Closest-hit record. tri = index into the reordered triangle array
 * (the reference Hit.triIndex), 0xFFFFFFFF on a miss. */
typedef struct orc_hit {
    float t, u, v;
    uint32_t tri;
} orc_hit;

typedef struct orc_tree orc_tree;

/* Scene AABB + 63-bit Morton codes per input triangle (bvh_morton.cpp:40-105). */
int orc_morton(const float* verts6, uint32_t nv, const uint32_t* tris, uint32_t nt,
               uint64_t* codes_out, float* scene_box6_out);
/* Stable 8-pass LSD radix sort of triangle indices by code (bvh_morton.cpp:109-159). */
int orc_sort(const uint64_t* codes, uint32_t nt, uint32_t* sorted_out);

/* Full bake pipeline minus file I/O (bvh_bake.cpp:180-194). */
int orc_build(const float* verts6, uint32_t nv, const uint32_t* tris, uint32_t nt, orc_tree** out);
void orc_tree_free(orc_tree* t);
/* counts[0]=nodes, [1]=tris (reordered, duplicates included), [2]=triIndexSize,
 * [3]=sortedTrisSize, [4]=max depth, [5]=spatial splits accepted, [6]=budget rejections. */
void orc_tree_counts(const orc_tree* t, uint32_t* counts);
void orc_tree_nodes(const orc_tree* t, orc_node48* out);
void orc_tree_tris(const orc_tree* t, uint32_t* out3);
void orc_tree_prims(const orc_tree* t, uint32_t* out);       /* original triangle id per reordered triangle */
void orc_tree_sorted(const orc_tree* t, uint32_t* out);      /* s_sortedTris[0..N) after the Morton sort */
void orc_tree_bfs_nodes(const orc_tree* t, orc_node48* out); /* node array before DFS renumbering */

/* Closest-hit / any-hit traversal over reference-layout arrays (bvh_traverse.cpp:22-178).
 * counts (optional): per ray {internal visits, triangle tests}. normals (optional): 3 floats per ray.
 * any_hit != 0: out hit.tri = first hit found, traversal stops (occlusion query). */
int orc_trace(const orc_node48* nodes, uint32_t nn, const uint32_t* tris, uint32_t ntri,
              const float* verts6, uint32_t nv, const orc_ray* rays, size_t n,
              orc_hit* hits, float* normals, uint32_t* counts, int any_hit, int nthreads);

/* Counter-based sample table value (seed, pixel, sample, dim) in [0,1). */
float orc_sample(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim);
void  orc_sincos(float x, float* s, float* c);
/* atan2 of the IBL lookup (octant reduction + degree-7 polynomial; shared with the GPU). */
float orc_atan2(float y, float x);

typedef struct orc_camera {
    float pos[3], fwd[3], right[3], up[3];
    float tan_half_fov_y;
} orc_camera;

typedef struct orc_render_params {
    int width, height, spp, max_bounce;
    int rr_bounce;      /* Russian roulette applies when bounce >= rr_bounce (reference: 3) */
    uint32_t seed;
    int env_mode;       /* 0 = reference sky gradient, 1 = constant env_color, 2 = IBL map,
                           3 = env_color plus sun_rad inside the cone dot(d, sun_dir) >= sun_cos */
    float env_color[3];
    const float* ibl;   /* env_mode 2: ibl_w * ibl_h * 3 floats, top row first */
    int ibl_w, ibl_h;
    const float* uv;    /* texture albedo: 2 floats per vertex */
    const float* tex;   /* tex_w * tex_h * 3 floats; when set, albedo is the nearest texel at the
                           interpolated uv and the albedo argument is unused */
    int tex_w, tex_h;
    float sun_dir[3];
    float sun_cos;
    float sun_rad[3];
} orc_render_params;

/* Path-trace loop of main.cpp:145-204 / 339-365 driven by the sample table.
 * albedo/emission: 3 floats per ORIGINAL triangle (prims maps reordered -> original);
 * albedo == NULL means the reference's constant 0.8, emission == NULL means none. */
int orc_render(const orc_node48* nodes, uint32_t nn, const uint32_t* tris, const uint32_t* prims,
               uint32_t ntri, const float* verts6, uint32_t nv, const float* albedo,
               const float* emission, const orc_camera* cam, const orc_render_params* p,
               float* image, int nthreads);

/* Loss = sum over pixels and channels of (I - target)^2 where the image is rendered with
 * float32 path geometry and float64 throughput/radiance (used for finite differences). */
double orc_render_loss_f64(const orc_node48* nodes, uint32_t nn, const uint32_t* tris,
                           const uint32_t* prims, uint32_t ntri, const float* verts6, uint32_t nv,
                           const float* albedo, const float* emission, const orc_camera* cam,
                           const orc_render_params* p, const float* target, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
