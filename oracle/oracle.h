/*
 * oracle.h -- CPU restatement of turner's pathtracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product (libturner_amd.so) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates.
 * Arithmetic follows the reference operation by operation (FP32, no FMA
 * contraction, IEEE division) so that results are bit-identical to the
 * reference's x86-64 build.
 *
 * Pinning (what checks each function against something independent):
 *   or_precompute        KATs of tests/test_triangle.cpp (unit normal, perpendicular),
 *                        an FP64 recomputation (tests) and bit equality with the
 *                        product's own precompute (tn_triangle_records_host)
 *   or_clip_triangle     KATs of tests/test_clipping.cpp:55-80
 *   or_ray_box           KATs of tests/test_intersection.cpp:15-36
 *   or_kdtree_build      KATs of tests/test_kdtree.cpp:16-63,162-186 (height,
 *                        node counts), identical node arrays from the product's
 *                        independent C++ builder (std::sort), and hits == brute force
 *   or_intersect         KATs of tests/test_kdtree.cpp:23-63,188-245 + brute force
 *                        over all triangles on random scenes + FP64 MT for r
 *   or_introsort         std::sort (libstdc++) on random inputs with ties (tests)
 *   or_xorshift64s       closed form of xorshift64* (Vigna 2016 / Marsaglia 2003)
 *   or_render / or_trace_rays (path tracer, pathtracer.cpp:14-102):
 *                        white-furnace closed form (rho * background); the
 *                        point-light term at depth 0 against rho max(0, l.n) Lc / pi
 *                        with l from p, computed in FP64 (tests); the shadow
 *                        predicate (lit iff no hit or r > |L - p2|) with blockers
 *                        nearer than, exactly at and beyond the light; depth 1
 *                        under a background c: rho c exactly with cosine sampling,
 *                        in expectation with the reference's uniform sampling
 *   extended estimator   furnace closed form (rho * Le) and the unbiasedness
 *                        relation cosine vs uniform sampling (tests)
 *   or_exposure/gamma    KATs of tests/test_effects.cpp
 *   or_ppm               KAT of tests/test_raster.cpp:44-57
 */
#ifndef TURNER_ORACLE_H
#define TURNER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_INVALID_ID (1u << 30)

/* Per-triangle precompute, 16 floats: v0.xyz n.xyz u.xyz v.xyz uv vv uu denom */
void or_precompute(const float* verts9, uint32_t n, float* out16);

/* Sutherland-Hodgman clip of one triangle against a box; writes the clipped
 * bbox (6 floats) and returns the number of polygon points (<2: quirk box). */
int or_clip_triangle(const float* tri9, const float* box6, float* out6);

int or_ray_box(const float* o, const float* d, const float* box6, float* tmin,
               float* tmax);

/* SAH kd-tree build (WH06 Alg. 4 as in the reference) + flatten into the
 * reference's 8-byte FlatNode format. Returns node count; *nodes is malloc'ed. */
size_t or_kdtree_build(const float* verts9, uint32_t n, uint64_t** nodes,
                       float* box6);
uint32_t or_kdtree_height(const uint64_t* nodes, size_t n_nodes);
void or_free(void* p);

/* Reference closest-hit traversal (no early exit). Per ray: id (OR_INVALID_ID
 * on miss), r/s/t (r = FLT_MAX, s = t = 0 on miss).  Counters (optional) get
 * inner nodes visited, leaves visited, triangle tests per ray. */
void or_intersect(const float* pre16, const uint64_t* nodes, const float* box6,
                  const float* o3, const float* d3, uint32_t nrays,
                  uint32_t* id, float* r, float* s, float* t, int nthreads);
/* Same traversal but with near-first early exit and the short-stack layout of
 * the GPU kernel; counts the work the GPU does (nodes, leaves, refs, tests). */
void or_intersect_count(const float* pre16, const uint64_t* nodes,
                        const float* box6, const float* o3, const float* d3,
                        uint32_t nrays, uint64_t* counters4, int nthreads);
/* Brute force: leaf test over all triangles in index order (strict <). */
void or_intersect_brute(const float* pre16, uint32_t ntris, const float* o3,
                        const float* d3, uint32_t nrays, uint32_t* id, float* r,
                        float* s, float* t, int nthreads);
/* FP64 Moeller-Trumbore distance/barycentrics of given triangle per ray. */
void or_mt64(const float* verts9, const uint32_t* id, const float* o3,
             const float* d3, uint32_t nrays, double* r, double* b1,
             double* b2);

/* xorshift64* (reference lib/xorshift.h:36-41,57-62) */
uint64_t or_xorshift64s_next(uint64_t* state);
float or_xorshift64s_float(uint64_t* state);

/* Camera (reference lib/types.h:89-123). */
typedef struct {
    float pos[3];
    float trafo[9]; /* row-major 3x3 (aiMatrix3x3 a1..c3) */
    float delta_x, delta_y;
} or_camera;
void or_raster2cam(const or_camera* cam, float px, float py, float w, float h,
                   float* dir3);

typedef struct {
    float pos[3];
    float color[4];
} or_light;

typedef struct {
    uint32_t width, height;
    int32_t max_depth;
    int32_t mc_samples;    /* --monte-carlo-samples */
    int32_t pixel_samples; /* --pixel-samples */
    uint32_t flags;        /* OR_FLAG_* */
    uint64_t seed;
    float bg[4];
    int32_t nthreads;
    uint32_t row_step; /* render rows y % row_step == row_offset only */
    uint32_t row_offset;
    uint32_t renderer;      /* OR_RENDERER_* */
    float max_visibility;   /* raycaster */
    float shadow_intensity; /* raytracer */
} or_render_params;

#define OR_RENDERER_PATHTRACER 0u /* pathtracer.cpp:14-102 */
#define OR_RENDERER_RAYCASTER 1u  /* raycaster.cpp:7-24 */
#define OR_RENDERER_RAYTRACER 2u  /* raytracer.cpp:6-67 (needs one light) */

#define OR_FLAG_COSINE 1u     /* cosine-weighted hemisphere sampling */
#define OR_FLAG_AREA_LIGHTS 2u /* emissive triangles: NEE + camera-visible emission */
#define OR_FLAG_REF_RNG 4u    /* reference RNG layout: per-row gen(42), one stream seed 4 */

/* Scene for trace: shading normals 9N, albedo 4N (rgba), emission 4N. */
typedef struct {
    const float* pre16;
    const float* verts9;
    const float* normals9;
    const float* albedo4;
    const float* emission4;
    uint32_t ntris;
    const uint64_t* nodes;
    float box[6];
    const float* reflective4;  /* raytracer: RGBA per triangle (NULL: no reflection) */
    const float* reflectivity; /* raytracer: per triangle */
} or_scene;

/* Renders linear RGBA (sum over samples / p, before exposure) into out
 * (w*h*4), sum of squares of per-sample radiance into out_sq (optional),
 * and counts closest-hit segments and shadow rays. */
void or_render(const or_scene* sc, const or_camera* cam, const or_light* lights,
               uint32_t nlights, const or_render_params* prm, float* out,
               float* out_sq, uint64_t* n_rays, uint64_t* n_shadow);

/* trace(ray, tree, lights, depth, conf) (trace.h:23-25) for n caller rays:
 * out4 (n x RGBA) receives each ray's radiance, entering at recursion level
 * `depth`.  Ray j draws its own xorshift64* stream seeded from (seed, j).
 * prm: max_depth, mc_samples, flags, seed, bg, nthreads, renderer and the
 * plugin options are used (width, height, pixel samples and rows are not). */
void or_trace_rays(const or_scene* sc, const or_light* lights, uint32_t nlights,
                   const or_render_params* prm, const float* o3, const float* d3, uint32_t n,
                   int32_t depth, float* out4, uint64_t* n_rays, uint64_t* n_shadow);

/* Tone map (reference lib/effects.h:15-49) and PPM P3 (lib/raster.h:79-100). */
float or_exposure(float light, float value);
float or_gamma(float light, float inverse_gamma);
/* Writes PPM text into buf (cap bytes); returns length (excluding NUL). */
size_t or_ppm(const float* rgba, uint32_t w, uint32_t h, char* buf, size_t cap);

/* FlatNode encoding (lib/kdtree.h:62-154) as the builder writes it. */
uint64_t or_flat_inner(int axis, float split, uint32_t right);
uint64_t or_flat_leaf(uint32_t a, uint32_t b);

/* libstdc++ std::sort restated on 32-bit keys (test hook for the event sort). */
void or_introsort_u64(uint64_t* keys, size_t n);

#ifdef __cplusplus
}
#endif
#endif
