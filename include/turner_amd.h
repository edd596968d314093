/*
 * turner_amd.h -- C ABI of the MI355X (gfx950) wavefront path tracer that
 * replaces turner's pathtracer hot path: trace() (pathtracer.cpp:14-102)
 * over KDTree / KDTreeIntersection::intersect (lib/kdtree.h:187-309,
 * lib/kdtree.cpp:474-607).
 *
 * Plain C types only: host arrays are const float* / uint64_t*, device arrays
 * are raw device pointers obtained from hipMalloc (or torch), streams are
 * hipStream_t passed as void*.  Every entry point returns 0 (TN_OK) or a
 * negative status and never throws; tn_last_error() gives a thread-local
 * message for the last failure.  Scene handles are immutable after creation
 * and may be shared by host threads.
 */
#ifndef TURNER_AMD_H
#define TURNER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TN_OK 0
#define TN_ERR_INVALID (-1)  /* bad handle or argument */
#define TN_ERR_OOM (-2)      /* host or device allocation failed */
#define TN_ERR_HIP (-3)      /* HIP runtime error (message in tn_last_error) */
#define TN_ERR_QUEUE (-4)    /* a device ray queue would overflow its capacity */
#define TN_ERR_NODEVICE (-5) /* no usable GPU */
#define TN_ERR_LIMIT (-6)    /* size beyond a documented limit */
#define TN_ERR_INCOMPLETE (-7) /* a traversal hit its iteration cap */

/* Invalid triangle id, as OptionalId{} of lib/kdtree.h:156-181 (MAX_TRIANGLE_ID). */
#define TN_INVALID_ID (1u << 30)
/* Largest triangle count (ids are 30 bits, lib/kdtree.h:66). */
#define TN_MAX_TRIANGLES ((1u << 30) - 1u)

typedef struct tn_scene tn_scene;

typedef struct {
    uint32_t num_triangles; /* KDTree::num_triangles  lib/kdtree.h:219 */
    uint32_t num_nodes;     /* KDTree::num_nodes (reference FlatNode count) lib/kdtree.h:218 */
    uint32_t height;        /* KDTree::height  lib/kdtree.h:197-216 */
    uint32_t device_nodes;  /* nodes of the device layout (one per inner node / leaf) */
    uint32_t device_refs;   /* triangle references held by device leaves */
    uint32_t num_emitters;  /* triangles with emission > 0 */
    float box[6];           /* KDTree::box  lib/kdtree.h:221 (min xyz, max xyz) */
    double build_ms;        /* host SAH build time */
    uint32_t refined_nodes; /* TN_SCENE_REFINED: nodes of the refined tree (else 0) */
    uint32_t refined_refs;  /* TN_SCENE_REFINED: triangle references of the refined tree */
} tn_scene_info;

/* Scene creation flags */
#define TN_SCENE_SERIAL_BUILD 1u /* build the kd-tree on one host thread */
/* Rendering tree: the reference's SAH build with two quirks removed.  A
 * triangle that clipping removes from a node entirely keeps, in the
 * reference, an unbounded box (lib/clipping.h:210-212) and so sinks into
 * every descendant, filling leaves with triangles that do not touch them; a
 * split whose one side holds no triangle is dropped (lib/kdtree.cpp:166-171),
 * so the other subtree also spans the empty space.  With this flag such
 * triangles are dropped and empty sides become empty leaves, and the SAH
 * prices a node step at 60 instead of 15 (a wave tests a leaf's triangles
 * together, so shallower trees with larger leaves render faster).  Closest hits
 * equal the reference tree's except between exactly equal distances (visit
 * order differs) and, rarely (of order one ray in 10^7), where the
 * thick-plane clip leaves a sliver of a triangle out of a leaf the ray
 * crosses; the scene has no reference FlatNode array (tn_scene_get_flat_nodes
 * fails).  The reference tree stays the default. */
#define TN_SCENE_EMPTY_LEAVES 2u
/* Refined reference tree, for renders whose hit ids are the reference's.  The
 * scene keeps the reference tree (tn_intersect, tn_scene_get_flat_nodes) and
 * adds a refined copy of it: the same inner nodes, leaves reached in the same
 * order, each leaf's references filtered to triangles whose bounding box
 * reaches the leaf (the others were sunk into it by the reference's
 * unbounded-box quirk) and larger leaves split further by rendering-variant
 * sub-trees whose every leaf also holds each of those triangles that meets its
 * cell.  Renders (and tn_intersect with TN_ISECT_REFINED) traverse the
 * refined tree, detect every ray whose closest hit ties with another triangle
 * at exactly the same distance -- only there can the order inside a leaf
 * decide the id -- and re-trace those rays, and rays with a zero direction
 * component, on the reference tree.  Not combinable with
 * TN_SCENE_EMPTY_LEAVES. */
#define TN_SCENE_REFINED 4u

/*
 * Replaces KDTree::KDTree(Triangles) (lib/kdtree.h:195, lib/kdtree.cpp:474-490):
 * precomputes the triangles (lib/triangle.h:22-39), builds the SAH kd-tree on
 * the host with the reference's algorithm and constants (lib/kdtree.cpp:124-467)
 * and uploads nodes, triangle references, triangle records and shading data
 * to the device.  Device memory: 8 B per node, 4 B per triangle reference,
 * 64 B per triangle record plus a 64-byte copy of the record per triangle
 * reference (reference order, read by the traversal), 68 B of shading data per
 * triangle.  verts9: 9 floats per triangle (v0 v1 v2).  normals9: vertex
 * normals (NULL: face normals).  albedo4: RGBA diffuse per triangle (NULL:
 * 0.8 grey, alpha 1).  emission4: RGBA emission per triangle (NULL: none).
 */
int tn_scene_create(const float* verts9, const float* normals9, const float* albedo4,
                    const float* emission4, uint32_t num_triangles, uint32_t flags,
                    tn_scene** out);
/* As tn_scene_create, but with a tree the caller already holds in the
 * reference's FlatNode format (lib/kdtree.h:62-154) and its root box, e.g.
 * from tn_tree_cache_load.  The tree is validated before upload. */
int tn_scene_create_with_tree(const float* verts9, const float* normals9,
                              const float* albedo4, const float* emission4,
                              uint32_t num_triangles, const uint64_t* flat_nodes,
                              size_t num_nodes, const float* box6, tn_scene** out);
/* Reflective colour (RGBA, 4 per triangle) and reflectivity (1 per triangle)
 * of Triangle (lib/triangle.h:23-33, main.cpp:40-47), read by the raytracer
 * (raytracer.cpp:37-46).  Without this call every reflectivity is 0. */
int tn_scene_set_reflection(tn_scene* scene, const float* reflective4, const float* reflectivity);
int tn_scene_destroy(tn_scene* scene);
int tn_scene_get_info(const tn_scene* scene, tn_scene_info* info);
/* Copies the reference-format FlatNode array (cap = capacity in nodes). */
int tn_scene_get_flat_nodes(const tn_scene* scene, uint64_t* out, size_t cap);
/* Copies the 16-float triangle records: v0 n u v uv vv uu denom. */
int tn_scene_get_triangle_records(const tn_scene* scene, float* out16, size_t cap);

/* Host-only (no GPU needed): the SAH build of tn_scene_create, returning the
 * reference-format FlatNode array (free with tn_buffer_free), its root box
 * and KDTree::height.  flags: TN_SCENE_SERIAL_BUILD. */
int tn_kdtree_build(const float* verts9, uint32_t num_triangles, uint32_t flags,
                    uint64_t** flat_nodes, size_t* num_nodes, float* box6, uint32_t* height);
void tn_buffer_free(void* p);
/* Host-only: the triangle records tn_scene_create uploads (16 floats per
 * triangle, as tn_scene_get_triangle_records returns them), i.e. the
 * precompute of Triangle (lib/triangle.h:22-39): v0, the face normal
 * normalize(cross(u, v)) with the cross product in double, u = v1 - v0,
 * v = v2 - v0, then uv, vv, uu and denom = uv^2 - uu vv. */
int tn_triangle_records_host(const float* verts9, uint32_t num_triangles, float* out16);
/* Host-only: the device layout of the tree tn_scene_create would upload (with
 * flags TN_SCENE_SERIAL_BUILD / TN_SCENE_EMPTY_LEAVES; with TN_SCENE_REFINED the
 * refined tree): nodes2 holds two
 * uint32 per node -- inner (split bits, right << 2 | axis), leaf (offset into
 * refs, count << 2 | 3), left child = next node -- and refs the leaves'
 * triangle ids.  Free both with tn_buffer_free. */
int tn_kdtree_build_device(const float* verts9, uint32_t num_triangles, uint32_t flags,
                           uint32_t** nodes2, size_t* num_nodes, uint32_t** refs, size_t* num_refs,
                           float* box6, uint32_t* height);

/* kd-tree cache replacing ./kdtree.cache (main.cpp:144-167).  Unlike the
 * reference's cache, the file is versioned and keyed by a 64-bit hash of the
 * triangle data: tn_tree_cache_load returns TN_ERR_INVALID for a cache that
 * belongs to other geometry (or is truncated/corrupt), so a stale cache is
 * rebuilt instead of silently rendering the wrong scene. */
int tn_tree_cache_save(const char* path, const float* verts9, uint32_t num_triangles,
                       const uint64_t* flat_nodes, size_t num_nodes, const float* box6);
int tn_tree_cache_load(const char* path, const float* verts9, uint32_t num_triangles,
                       uint64_t** flat_nodes, size_t* num_nodes, float* box6);

/* Intersection flags */
/* Visit every leaf the ray crosses, in the reference's order, as
 * KDTreeIntersection::intersect does (kdtree.cpp:515-607).  Without it a ray
 * stops after the first leaf whose exit lies beyond its best hit by a 2^-16
 * margin: the same hits on scenes of about unit scale (checked on every test
 * and benchmark batch), but not always where coordinates are far larger than
 * the triangles (a mesh of unit size moved 1e2..1e4 from the origin: a few in
 * 10^4 of rays aimed at vertices) or where the clipping tolerance 1e-5 is not
 * small against the triangles (a unit mesh scaled by 1e-3): there a later
 * leaf can still hold a triangle hit a little before the best one.  Use this
 * flag for such scenes. */
#define TN_ISECT_EXHAUSTIVE 1u
#define TN_ISECT_MT 2u         /* FP32 Moeller-Trumbore leaf test instead of the reference's */
#define TN_ISECT_COUNT 4u      /* accumulate work counters (tn_get_counters) */
/* Traversal schedules (default: wave-cooperative leaf tests -- the wave pools
 * the triangle references of all its lanes' current leaves and tests them 64
 * at a time).  Results are identical under every schedule. */
#define TN_ISECT_TICK 8u       /* each lane: a few node steps or one triangle test per step */
#define TN_ISECT_LEAF 16u      /* each lane: one leaf (all of its triangles) per step */
/* TN_SCENE_REFINED scenes: traverse the refined tree; rays tied at their
 * closest distance and rays with a zero direction component are re-traced on
 * the reference tree, so results equal the default traversal's. */
#define TN_ISECT_REFINED 32u
/* Bits 24..27: occupancy variant of the default closest-hit kernel (waves
 * per SIMD 5..8, register budget accordingly; 0 = the tuned default). */
#define TN_ISECT_WAVES(w) ((uint32_t)(w) << 24)

/*
 * Batched twin of KDTreeIntersection::intersect(ray, r, a, b) (lib/kdtree.h:291,
 * lib/kdtree.cpp:515-607).  Device pointers: o3, d3 hold n rays as xyz
 * triples; outputs id (TN_INVALID_ID on miss), r (ray parameter of the hit,
 * FLT_MAX on miss), s, t (barycentrics of v1, v2; 0 on miss).  Closest hit
 * with the reference's strict-< tie rule in near-first visit order.
 * stream: hipStream_t or NULL for the null stream.  Asynchronous.  At most
 * 2^31 rays per call (TN_ERR_LIMIT).  Calls of more than 256 rays per
 * resident wave use one of the scene's 16 work slots (its stream may wait for
 * the launch that last used the slot); calls may run concurrently on several
 * streams.
 */
int tn_intersect(tn_scene* scene, const float* o3, const float* d3, uint32_t n,
                 uint32_t* id, float* r, float* s, float* t, uint32_t flags, void* stream);
/* tn_intersect for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops): no reference counterpart.  Same arguments,
 * results and flags as tn_intersect; allocates nothing and records no event.
 * Large batches (more than 256 rays per resident wave) draw their chunks from
 * one set of work counters held in a static device array (16 x 9 words per
 * device, kept until the process ends), also under stream capture, where
 * tn_intersect falls back to equal static shares.  The counters are cleared
 * by a kernel at the start of every call, so calls on one device must not
 * overlap: concurrent calls on different streams of one device race on the
 * counters.  The scene's device must be current.  At most 2^31 rays per call
 * (TN_ERR_LIMIT). */
int tn_intersect_noalloc(tn_scene* scene, const float* o3, const float* d3, uint32_t n,
                         uint32_t* id, float* r, float* s, float* t, uint32_t flags, void* stream);
/* Host-pointer variant (copies in and out, synchronous). */
int tn_intersect_host(tn_scene* scene, const float* o3, const float* d3, uint32_t n,
                      uint32_t* id, float* r, float* s, float* t, uint32_t flags);
/* Shadow / visibility query with the reference's predicate
 * (pathtracer.cpp:49-53): occluded[i] = 1 iff some triangle is hit at
 * 0 <= r <= tmax[i], i.e. iff !(!hit || r_closest > tmax[i]).  Device
 * pointers.  Asynchronous. */
int tn_occluded(tn_scene* scene, const float* o3, const float* d3, const float* tmax,
                uint32_t n, uint8_t* occluded, uint32_t flags, void* stream);
/* Work counters of the launches made with TN_ISECT_COUNT since the last reset:
 * [0] inner nodes visited, [1] leaves visited, [2] triangle references read,
 * [3] ray-triangle tests, [4] rays, [5] stack restarts, [6] tests that read
 * the whole 64-byte triangle record (the rest stop after 32 bytes), [7] the
 * most leaves any single ray visited.
 * reset != 0 zeroes them.  Synchronises the device. */
int tn_get_counters(tn_scene* scene, uint64_t* out8, int reset);

/* Camera as lib/types.h:89-144: ray direction =
 * trafo * (-delta_x (1 - 2 px / w), delta_y (1 - 2 py / h), -1), origin pos. */
typedef struct {
    float pos[3];
    float trafo[9]; /* row-major 3x3 */
    float delta_x, delta_y;
} tn_camera;

/* Point light as lib/types.h:81-84 */
typedef struct {
    float pos[3];
    float color[4];
} tn_light;

/* Render flags */
#define TN_RENDER_COSINE 1u      /* cosine-weighted hemisphere (default: uniform, as the reference) */
/* emissive triangles: camera-visible emission + one next-event sample per hit: an emitter by solid
 * angle (scenes with <= 4 emitter triangles) or by area, then a direction uniform in its solid angle
 * (Arvo 1995), or a point uniform by area when it subtends < 1e-3 sr */
#define TN_RENDER_AREA_LIGHTS 2u
#define TN_RENDER_REF_JITTER 4u  /* pixel jitter of main.cpp:201-206 (xorshift64* seeded 42 per row) */
#define TN_RENDER_GUIDED 8u      /* learned guiding proposal (needs a tn_guide) */
#define TN_RENDER_EXHAUSTIVE 16u /* reference traversal order without early exit */
#define TN_RENDER_GRAPH 32u      /* capture the per-chunk launch sequence once in a hipGraph, replay per chunk */
#define TN_RENDER_AUTOTUNE 64u   /* batches: chunk size, grid and launch mode picked by timing probes (cached) */
/* Plain path tracing (no guide, no training records, camera samples, not a
 * TN_SCENE_REFINED scene) in one cooperative launch of a persistent kernel
 * that walks every chunk and level, its stages separated by grid barriers
 * instead of kernel boundaries (no reference counterpart).  Same paths and
 * counts as the launch-per-stage sequence; slower on MI355X (one pipeline,
 * no shadow/trace overlap, the union of the stages' registers), so the
 * autotuner does not pick it.  Ignored where it does not apply. */
#define TN_RENDER_PERSISTENT 128u
/* tn_render_guided: the records of a training render are not trained at the
 * end of its call but during the next tn_render_guided call with this flag
 * (on the guide's own stream, beside that call's render, which samples from a
 * copy of the weights taken before that training starts), or by
 * tn_guide_flush / tn_guide_train_recorded.  Every call still returns with
 * its work done.  No reference counterpart. */
#define TN_RENDER_DEFER_TRAINING 256u

typedef struct {
    uint32_t width, height;
    int32_t max_depth;     /* --max-depth */
    int32_t mc_samples;    /* --monte-carlo-samples */
    int32_t pixel_samples; /* --pixel-samples */
    uint32_t flags;        /* TN_RENDER_* */
    uint64_t seed;
    float bg[4];           /* --background (alpha 1) */
    float exposure;        /* --exposure */
    float inverse_gamma;   /* --inverse-gamma */
    int32_t gamma_enabled; /* !--no-gamma-correction */
    uint32_t chunk_paths;  /* primary samples per wavefront chunk (0: automatic) */
    uint32_t sample_offset;/* first pixel-sample index (progressive rendering) */
    uint32_t renderer;     /* TN_RENDERER_*: which of the reference's trace() plugins */
    float max_visibility;  /* raycaster --max-visibility (reference default 2) */
    float shadow_intensity;/* raytracer --shadow (reference default 0.5) */
    uint32_t reserved;
} tn_render_params;

/* The reference's trace() plugins (trace.h:23-25; CMakeLists.txt:62-67 links
 * one per binary).  All three run on the same traversal and queues. */
#define TN_RENDERER_PATHTRACER 0u /* pathtracer.cpp:14-102 */
#define TN_RENDERER_RAYCASTER 1u  /* raycaster.cpp:7-24: diffuse colour, alpha fading with distance */
#define TN_RENDERER_RAYTRACER 2u  /* raytracer.cpp:6-67: Lambert + mirror reflection + shadow darkening */

typedef struct {
    uint64_t rays;        /* closest-hit segments, the reference's num_rays (pathtracer.cpp:21) */
    uint64_t primary_rays;/* num_prim_rays (main.cpp:211) */
    uint64_t shadow_rays; /* visibility queries */
    double ms;            /* wall time of the render loop (device-synchronised) */
} tn_render_stats;

/*
 * Replaces the pixel loop of main.cpp:187-226 calling trace() per pixel
 * sample.  linear_rgba (host, w*h*4, may be NULL): mean radiance over the
 * pixel samples before exposure, as image(x,y) after main.cpp:216.
 * display_rgba (host, may be NULL): after exposure and gamma (main.cpp:218-223).
 */
int tn_render(tn_scene* scene, const tn_camera* cam, const tn_light* lights,
              uint32_t num_lights, const tn_render_params* params, float* linear_rgba,
              float* display_rgba, tn_render_stats* stats);

/* Same as tn_render but leaves the result in device memory: accum (device,
 * w*h*4 floats) receives the SUM over pixel samples (not divided) and is
 * accumulated into, not overwritten. */
int tn_render_accumulate(tn_scene* scene, const tn_camera* cam, const tn_light* lights,
                         uint32_t num_lights, const tn_render_params* params, float* accum,
                         tn_render_stats* stats, void* stream);

/*
 * Batched twin of trace(ray, tree_intersection, lights, depth, conf)
 * (pathtracer.cpp:14-102, trace.h:23-25): traces n caller rays (device
 * arrays o3, d3) starting at recursion level `depth` and ADDS each ray's
 * radiance (RGBA) to rgba[4 i .. 4 i + 3] (device).  The estimator, flags,
 * background, m (mc_samples) and max_depth come from params (width, height,
 * pixel_samples and jitter are ignored).  depth > max_depth traces nothing
 * (the reference returns Color{}).  Asynchronous on `stream`.
 */
int tn_trace(tn_scene* scene, const float* o3, const float* d3, uint32_t n, int32_t depth,
             const tn_light* lights, uint32_t num_lights, const tn_render_params* params,
             float* rgba, tn_render_stats* stats, void* stream);

/* Tone map (lib/effects.h:15-49) of device RGBA sums: out = gamma(exposure(sum/spp)). */
int tn_tonemap(const float* accum, float* out, uint32_t num_pixels, float inv_spp,
               float exposure, float inverse_gamma, int gamma_enabled, void* stream);

/* PPM "P3" text of an RGBA image as lib/raster.h:79-100 writes it
 * (int(clamp(255 * c * a)) per channel).  Returns the byte count written
 * (excluding the terminating NUL) or a negative status; buf may be NULL to
 * query the size. */
long tn_write_ppm(const float* rgba, uint32_t width, uint32_t height, char* buf, size_t cap);

/*
 * Learned path guiding.  No reference counterpart: the reference samples the
 * hemisphere uniformly (lib/sampling.h:20-32, pathtracer.cpp:64-101); these
 * entry points are an addition behind the same render call.
 *
 * A tn_guide holds an MLP (78 -> 64 -> 64 -> 64 -> 40, ReLU) mapping a hit
 * point and its shading normal to an 8-lobe von Mises-Fisher mixture q(w):
 * FP32 master parameters, Adam state, BF16 operand copies for the MFMA
 * kernels and the training set the guided renderer fills.  Flat parameter
 * layout: per layer W (out x in, row major) then b.  Network outputs per
 * point: 8 x (logit, m.xyz, kraw); pi = softmax(logit), mu = m / |m|,
 * kappa = min(softplus(kraw) + 0.01, 1e4).
 */
typedef struct tn_guide tn_guide;
#define TN_GUIDE_PARAMS 15976
#define TN_GUIDE_OUTPUTS 40

/* Positions are encoded relative to the box [bmin, bmax] (the scene box). */
int tn_guide_create(const float bmin[3], const float bmax[3], uint64_t seed, tn_guide** out);
int tn_guide_destroy(tn_guide* g);
/* FP32 master parameters to / from host memory (TN_GUIDE_PARAMS floats);
 * set_params resets the Adam state. */
int tn_guide_get_params(tn_guide* g, float* theta);
int tn_guide_set_params(tn_guide* g, const float* theta);
/* Adam learning rate (1e-3), mini-batch size (16384, a multiple of 64), the
 * one-sample MIS probability of drawing from the guide (0.5) and the epochs
 * of each training pass (1): a pass trains on a uniform subset of
 * train_samples of its recorded samples, epochs * train_samples / batch Adam
 * steps. */
int tn_guide_set_options(tn_guide* g, float lr, uint32_t batch, float mis_guided, uint32_t epochs);
/* Network outputs (40 per point, device) for n points (device p3, n3). */
int tn_guide_forward(tn_guide* g, const float* p3, const float* n3, uint32_t n, float* out,
                     void* stream);
/* Mixture density q(w | p, n) (device arrays). */
int tn_guide_pdf(tn_guide* g, const float* p3, const float* n3, const float* w3, uint32_t n,
                 float* pdf, void* stream);
/* Directions drawn from q with the uniforms u3 (device, 3 per point: lobe,
 * cosine, azimuth) and their density q(w). */
int tn_guide_sample(tn_guide* g, const float* p3, const float* n3, const float* u3, uint32_t n,
                    float* w3, float* pdf, void* stream);
/* Loss mean_i(c_i / mean(c) * -log q(w_i | p_i, n_i)) of a batch (device
 * arrays) and its gradient (device, TN_GUIDE_PARAMS floats).  The loss is
 * returned to the host (synchronises the stream). */
int tn_guide_loss_grad(tn_guide* g, const float* p3, const float* n3, const float* w3,
                       const float* c, uint32_t n, float* grad, float* loss, void* stream);
/* One Adam step on a batch (device arrays); loss may be NULL (asynchronous). */
int tn_guide_train_step(tn_guide* g, const float* p3, const float* n3, const float* w3,
                        const float* c, uint32_t n, float* loss, void* stream);
/* Samples in the training set of the last training render, Adam steps taken
 * and the loss of the last step. */
int tn_guide_info(tn_guide* g, uint32_t* train_samples, uint32_t* steps, float* last_loss);
/* One more training pass on the set the last training render recorded: a new
 * uniform subset of train_samples records, epochs * train_samples / batch
 * Adam steps (as tn_render_guided trains after rendering).  Lets one render
 * supply the data of several training passes.  Asynchronous on `stream`. */
int tn_guide_train_recorded(tn_guide* g, uint32_t train_samples, void* stream);
/* Trains the records a TN_RENDER_DEFER_TRAINING render left untrained (if
 * any), on stream; afterwards nothing is pending.  Asynchronous on stream.
 * No reference counterpart. */
int tn_guide_flush(tn_guide* g, void* stream);
/* Copies up to cap samples of that training set to host memory, 12 floats
 * each: p.xyz, c, n.xyz, cos/pdf, w.xyz, tag (uint32 bits: the path's sample
 * index within its render call, modulo 2^26, << 6 | bounce level).  Returns
 * the number copied. */
long tn_guide_get_train_set(tn_guide* g, float* out, size_t cap);

/*
 * tn_render_accumulate with the guide as proposal: every indirect direction
 * is drawn from q with probability mis_guided and from the cosine lobe
 * otherwise, weighted by the exact one-sample MIS density
 * mis_guided * q(w) + (1 - mis_guided) cos(w) / pi, so the estimator of
 * trace() stays unbiased.  With train_samples > 0 (needs mc_samples == 1)
 * the render also records, for every bounce, the radiance its path brought
 * back, and afterwards trains on a uniform subset of train_samples of those
 * records (epochs * train_samples / batch Adam steps, tn_guide_set_options).
 */
int tn_render_guided(tn_scene* scene, const tn_camera* cam, const tn_light* lights,
                     uint32_t num_lights, const tn_render_params* params, tn_guide* guide,
                     uint32_t train_samples, float* accum, tn_render_stats* stats, void* stream);

/*
 * Batched heterogeneous rendering (no reference counterpart: the reference
 * renders one scene per process).  tn_batch_create packs the trees and
 * shading data of n scenes into one forest (the scenes may be destroyed
 * afterwards); tn_batch_render renders one job per scene -- camera, at most
 * one point light, image size, pixel samples, background -- through one
 * wavefront launch sequence per chunk of samples, where chunks cut across
 * scenes and every ray finds its scene from its pixel.  accum (device)
 * receives the per-pixel SUMS of all images back to back: job j's image starts
 * at float offset 4 * sum_{i<j} width_i * height_i.  From params only
 * max_depth, mc_samples, flags, seed, chunk_paths and sample_offset are used.
 */
typedef struct tn_batch tn_batch;
typedef struct {
    tn_camera cam;
    tn_light light;
    uint32_t num_lights; /* 0 or 1 */
    uint32_t width, height;
    int32_t pixel_samples;
    float bg[4];
} tn_batch_job;
typedef struct {
    uint32_t chunk_paths;   /* primary samples per chunk */
    uint32_t blocks_per_cu; /* grid of the persistent trace / shade / shadow kernels */
    uint32_t graph;         /* chunk sequence replayed from a hipGraph */
    uint32_t candidates;    /* launch shapes timed */
    double probe_ms;        /* time spent tuning */
    double best_rate;       /* samples per ms of the chosen shape on the probe */
} tn_batch_tuning;
int tn_batch_create(tn_scene* const* scenes, uint32_t num_scenes, tn_batch** out);
int tn_batch_destroy(tn_batch* batch);
int tn_batch_render(tn_batch* batch, const tn_batch_job* jobs, const tn_render_params* params,
                    float* accum, tn_render_stats* stats, void* stream);
/* The launch shape the autotuner chose for the last TN_RENDER_AUTOTUNE render. */
int tn_batch_get_tuning(const tn_batch* batch, tn_batch_tuning* out);

/*
 * Test hook (no reference counterpart), host only: the shade kernel's
 * solid-angle sampler of a triangle seen from a point, on the CPU.  p3: the
 * point; tri9: the vertices; u2: n pairs (u1, u2) in [0, 1); writes n unit
 * directions to dir3 and the sampler's solid angle (sr) to *omega and the
 * triple-product solid angle used to pick among emitters to *omega_vos.
 * TN_ERR_INVALID for a degenerate projection.
 */
int tn_sample_triangle_host(const float* p3, const float* tri9, const float* u2, uint32_t n, float* dir3,
                            float* omega, float* omega_vos);

const char* tn_last_error(void);
/* Device name and compute-unit count of the current HIP device. */
int tn_device_info(char* name, size_t cap, int* compute_units);

#ifdef __cplusplus
}
#endif
#endif
