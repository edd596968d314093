// render.hip -- wavefront path tracer: the pixel loop of main.cpp:187-226 over
// trace() (pathtracer.cpp:14-102), unrolled into device queues.
//
// trace() computes  L = rho * (direct / pi + (1/m) sum_j 2 cos_j L_j)  with one
// point light (no 1/r^2, pathtracer.cpp:44-58), a uniform hemisphere per
// indirect sample (lib/sampling.h:20-32) and m samples per hit; levels
// 0..max_depth are traced and a miss returns the background.  Unrolled, a path
// segment at depth k carries the product W of the factors rho * 2 cos / m of
// its ancestors and adds W * (miss ? bg : rho * direct / pi) to its pixel.
// One chunk of primary samples is processed level by level:
//
//   raygen -> [trace -> shade -> shadow] x (max_depth + 1)
//
//   trace : closest hit for every queued segment (lane-refill traversal);
//   shade : miss / emission / next-event estimation (shadow queue) and m
//           children per hit (next queue), pushed with wave-aggregated
//           atomics (ballot + mbcnt prefix);
//   shadow: visibility of the shadow queue (reference predicate
//           !hit || r > dist, pathtracer.cpp:53); unoccluded contributions
//           are added to the pixel's FP32 accumulator.
//
// All counts live in device memory, so the per-chunk launch sequence needs no
// host synchronisation.  RNG: one xorshift128 (Marsaglia 2003) state per
// segment, seeded by splitmix64 from (seed, sample index) and split for the
// children.  The reference's single sequential xorshift64* stream
// (lib/sampling.h:12) cannot be reproduced by a parallel tracer; images agree
// with the reference estimator statistically.  With TN_RENDER_REF_JITTER the
// pixel jitter is the reference's exactly (xorshift64* seeded 42 per row,
// main.cpp:201-206), so primary rays are bit-identical.
//
// Guided rendering (tn_render_guided, no reference counterpart) swaps the
// hemisphere sampler in shade for one-sample MIS between the learned vMF
// mixture q and the cosine lobe: the wave runs the guide MLP on its 64 hits
// (BF16 MFMA, tn_guide.h), each child direction comes from q with
// probability a and from the cosine lobe otherwise, and carries the weight
// rho cos / (pi m (a q(w) + (1 - a) cos / pi)), so E[estimate] is unchanged.
// When training, every bounce of a path (m = 1) leaves a record (point,
// normal, direction, cos / pdf, rho f) and every segment the radiance it
// found itself (background, unoccluded next-event estimates); after the
// chunk, k_collect walks each path backwards, L_k = X_{k+1} + g_{k+1} L_{k+1},
// and appends (p, n, w, lum(L_k) cos / pdf) to the guide's training set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "tn_guide.h"
#include "tn_sample.h"
#include "tn_scene.h"

namespace tn {
namespace {

constexpr int kRBlock = 256;
// LDS short-stack entries per lane of the renderer's traversal kernels (16
// measured no faster on small or large launches than 8)
constexpr int kRStack = kStack;
constexpr uint32_t kMaxLevels = 64;
constexpr uint32_t kFixCap = 1u << 16;
constexpr uint32_t kFixHits = 4 * kMaxLevels + 2, kFixShadows = 5 * kMaxLevels + 2;
constexpr uint64_t kSceneSeedStep = 0x9E3779B97F4A7C15ull;
// scenes with at most this many emitter triangles pick one by solid angle
constexpr uint32_t kSolidAngleSelect = 4;
// concurrent render pipelines (queues + streams) that take alternate chunks;
// 3 and 4 measured slower than 2 (1920x1080: -7 %, -4 %)
constexpr int kPipes = 2;

struct Queues {
    float4* o[2];   // o.xyz, pixel (bits)
    float4* d[2];   // d.xyz, depth (bits)
    float4* w[2];   // throughput weight (RGBA)
    uint4* rng[2];  // xorshift128 state
    uint4* hit;     // id, r, s, t bits
    // shadow queues, two: level k uses [k & 1], so the shadow test of level
    // k can run (on a second stream) while level k + 1 is traced and shaded
    float4* so[2];  // shadow: origin, pixel (bits)
    float4* sd[2];  // shadow: direction, tmax
    float4* sc[2];  // shadow: contribution
    float4* sx[2];  // shadow: radiance at the segment (training records), slot (bits)
    uint32_t cap;   // capacity of each path queue (entries)
    uint32_t scap;  // capacity of the shadow queue
    // device counters: [0..kMaxLevels) path counts, [kMaxLevels..2k) shadow
    // counts, [2k..3k) trace work, [3k..4k) shadow work, [4k] overflow,
    // [4k + 1] raycaster hits, [4k + 2..5k + 2) segments and [5k + 2..6k + 2)
    // shadow rays a refined traversal handed to the reference tree
    uint32_t* cnt;
    // lists of those rays: [0] the segments of the level being traced,
    // [1 + (level & 1)] the shadow rays of a level; kFixCap entries each (a
    // list that overflows is replaced by a scan of the whole queue)
    uint32_t* fix;
    // chunk counters of the trace / shadow kernels: [kind][level][XCD], 16
    // words apart (one per 64-byte line)
    uint32_t* xwork;
    unsigned long long* totals;  // [0] segments, [1] shadow rays
    uint32_t* bar;  // grid barrier words of the persistent kernel (kBarWords)
};

// Per-scene parameters of a render: one entry per scene of a batch (one for
// a single-scene render).  Global sample j of the call belongs to the last
// scene whose sample_begin <= j; its pixel p to the last whose pixel_begin <= p.
struct SceneParams {
    float4 cam_pos;
    float cam_m[9];
    float cam_dx, cam_dy;
    uint32_t width, height, spp, nlights;
    float4 light_pos, light_col, bg;
    unsigned long long sample_begin;  // first global sample index
    uint32_t pixel_begin;             // first accumulator pixel
    uint32_t emit_begin, nemit;       // emitter range (emit_ids / emit_cdf)
    float emit_area;
};

// Chunk bounds kept on the device, so that one captured launch sequence
// serves every chunk: k_chunk_begin advances them.
struct ChunkState {
    unsigned long long next, total, first;
    unsigned long long stride;  // samples from one chunk of this pipeline to its next
    uint32_t chunk, count;
};

struct RenderArgs {
    SceneView sv;        // the tree the traversal kernels walk (refined tree of a TN_SCENE_REFINED scene)
    SceneView sv_exact;  // the reference tree behind a refined one (rays the refined traversal marks)
    bool refined;        // sv is a refined tree
    const float* normals;  // 9 per triangle
    const float4* albedo;
    const float4* emission;
    const float4* trirec;  // device triangle records (geometric normal in rec[1])
    const uint32_t* emit_ids;
    const float* emit_cdf;  // per scene: inclusive prefix of its emitters' areas
    const SceneParams* sp;
    uint32_t nscenes;
    ChunkState* cs;
    uint32_t* status;  // [2]: traversals abandoned as incomplete
    int32_t max_depth, mc;
    uint32_t flags;
    uint64_t seed;
    const float* jitter;  // 2 * width * spp (REF_JITTER, single scene) or null
    float4* accum;
    uint64_t seed_pass;  // seed of this call's sample range (sample_offset)
    // guided sampling
    GuideView gv;
    float mis;  // probability of drawing from the guide
    // training records, slot = path * rstride + level (null unless training)
    float4* rec_p;  // hit point, cos / pdf
    float4* rec_n;  // shading normal, valid (1) when the bounce spawned its child
    float4* rec_w;  // direction
    float4* rec_g;  // rho * f: the child's weight relative to its parent
    float4* rec_x;  // radiance found by the segment itself (its level), local units
    uint32_t rstride;
    float4* ts;  // guide training set (3 float4 per sample)
    uint32_t* ts_count;
    uint32_t ts_cap;
    float accept;  // probability that a record enters the training set
    // raycaster / raytracer
    float max_vis;    // raycaster max_visibility
    float shadow_si;  // raytracer shadow_intensity
    const float4* reflective;   // raytracer: reflective colour per triangle
    const float* reflectivity;  // raytracer: null = no reflection
};

__device__ __forceinline__ uint32_t scene_of_pixel(const RenderArgs& A, uint32_t pixel) {
    if (A.nscenes <= 1) return 0;
    uint32_t lo = 0, hi = A.nscenes - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (A.sp[mid].pixel_begin <= pixel) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t scene_of_sample(const RenderArgs& A, unsigned long long j) {
    if (A.nscenes <= 1) return 0;
    uint32_t lo = 0, hi = A.nscenes - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (A.sp[mid].sample_begin <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint4 rng_seed(uint64_t a, uint64_t b) {
    const uint64_t h1 = splitmix64(a ^ splitmix64(b));
    const uint64_t h2 = splitmix64(h1 + 0x632BE59BD9B4E019ull);
    uint4 s = make_uint4((uint32_t)h1, (uint32_t)(h1 >> 32), (uint32_t)h2, (uint32_t)(h2 >> 32));
    if ((s.x | s.y | s.z | s.w) == 0) s.x = 1;
    return s;
}

// xorshift128 (Marsaglia 2003)
__device__ __forceinline__ uint32_t rng_next(uint4& s) {
    uint32_t t = s.x ^ (s.x << 11);
    s.x = s.y;
    s.y = s.z;
    s.z = s.w;
    s.w = s.w ^ (s.w >> 19) ^ (t ^ (t >> 8));
    return s.w;
}
__device__ __forceinline__ float rng_uniform(uint4& s) { return (float)(rng_next(s) >> 8) * 0x1p-24f; }

// Stream of child c: two words of the parent's stream, high word first (named
// locals fix the order, which an expression with two calls would leave open).
__device__ __forceinline__ uint4 child_seed(uint4& parent, uint32_t c) {
    const uint64_t hi = rng_next(parent);
    const uint64_t lo = rng_next(parent);
    return rng_seed((hi << 32) | lo, c);
}

__device__ __forceinline__ float4 mul4(float4 a, float4 b) {
    return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}
__device__ __forceinline__ float4 scale4(float s, float4 a) {
    return make_float4(s * a.x, s * a.y, s * a.z, s * a.w);
}
__device__ __forceinline__ void accumulate(float4* acc, uint32_t pixel, float4 c) {
    float* p = reinterpret_cast<float*>(acc + pixel);
    atomicAdd(p + 0, c.x);
    atomicAdd(p + 1, c.y);
    atomicAdd(p + 2, c.z);
    atomicAdd(p + 3, c.w);
}

// Starts the next chunk: [first, first + count) of the call's samples; zeroes
// the level counters and sets the level-0 count.
__global__ void k_chunk_begin(uint32_t* cnt, uint32_t* xwork, ChunkState* cs) {
    const uint32_t i = threadIdx.x;
    const unsigned long long first = cs->next;
    const unsigned long long left = cs->total > first ? cs->total - first : 0ull;
    const uint32_t count = (uint32_t)(left < cs->chunk ? left : cs->chunk);
    // the overflow word [4 kMaxLevels] lives for the whole render
    if (i < 4 * kMaxLevels || i == 4 * kMaxLevels + 1) cnt[i] = (i == 0) ? count : 0u;
    for (uint32_t k = kFixHits + i; k < 6 * kMaxLevels + 2; k += blockDim.x) cnt[k] = 0u;
    for (uint32_t k = i; k < 2 * kMaxLevels * kXcds; k += blockDim.x) xwork[16 * k] = 0u;
    __syncthreads();
    if (i == 0) {
        cs->first = first;
        cs->count = count;
        cs->next = first + cs->stride;
    }
}

__global__ void k_zero_totals(unsigned long long* totals, uint32_t* cnt, uint32_t* incomplete) {
    if (threadIdx.x < 2) totals[threadIdx.x] = 0;
    if (threadIdx.x == 2) cnt[4 * kMaxLevels] = 0;  // queue overflow
    if (threadIdx.x == 3 && incomplete) *incomplete = 0;
}

// levels < 0: the raycaster, whose rays are its hits (cnt[4 kMaxLevels + 1])
__global__ void k_chunk_totals(const uint32_t* cnt, unsigned long long* totals, int levels) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long a = levels < 0 ? cnt[4 * kMaxLevels + 1] : 0, b = 0;
        for (int k = 0; k < levels; ++k) {
            a += cnt[k];
            b += cnt[kMaxLevels + k];
        }
        totals[0] += a;
        totals[1] += b;
    }
}

// Primary samples: within a scene, j = (y * width + x) * spp + i
// (main.cpp:193-215 order); scenes follow each other in the sample space.
__device__ __forceinline__ void raygen_one(const RenderArgs& A, const Queues& Q, uint32_t i) {
    const unsigned long long j = A.cs->first + i;
    const uint32_t si = scene_of_sample(A, j);
    const SceneParams& S = A.sp[si];
    const unsigned long long local = j - S.sample_begin;
    const uint32_t sample = (uint32_t)(local % S.spp);
    const uint64_t pix = local / S.spp;
    const uint32_t x = (uint32_t)(pix % S.width), y = (uint32_t)(pix / S.width);
    // scene s of a batch draws the streams of a single render seeded seed + s * kSceneSeedStep
    uint4 rng = rng_seed(A.seed_pass + si * kSceneSeedStep, local);
    float dx, dy;
    if (A.jitter) {
        const size_t k = 2 * ((size_t)x * S.spp + sample);
        dx = A.jitter[k];
        dy = A.jitter[k + 1];
    } else {
        dx = rng_uniform(rng);
        dy = rng_uniform(rng);
    }
    // lib/types.h:119-123
    const float px = (float)x + dx, py = (float)y + dy;
    const float w = (float)S.width, h = (float)S.height;
    const float cx = -S.cam_dx * (1 - 2 * px / w);
    const float cy = S.cam_dy * (1 - 2 * py / h);
    const float cz = -1;
    const float* m = S.cam_m;
    const float3 d = make_float3(m[0] * cx + m[1] * cy + m[2] * cz, m[3] * cx + m[4] * cy + m[5] * cz,
                                 m[6] * cx + m[7] * cy + m[8] * cz);
    Q.o[0][i] = make_float4(S.cam_pos.x, S.cam_pos.y, S.cam_pos.z, __uint_as_float(S.pixel_begin + (uint32_t)pix));
    Q.d[0][i] = make_float4(d.x, d.y, d.z, __uint_as_float(i));  // .w: path within the chunk
    Q.w[0][i] = make_float4(1.f, 1.f, 1.f, 1.f);
    Q.rng[0][i] = rng;
}

__global__ __launch_bounds__(kRBlock) void k_raygen(RenderArgs A, Queues Q) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.cs->count) return;
    raygen_one(A, Q, i);
}

// Caller rays (tn_trace): ray j accumulates into accum[j]; seeds from (seed, j).
__global__ __launch_bounds__(kRBlock) void k_raygen_rays(RenderArgs A, Queues Q, const float* o3,
                                                         const float* d3) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.cs->count) return;
    const uint64_t j = A.cs->first + i;
    Q.o[0][i] = make_float4(o3[3 * j], o3[3 * j + 1], o3[3 * j + 2], __uint_as_float((uint32_t)j));
    Q.d[0][i] = make_float4(d3[3 * j], d3[3 * j + 1], d3[3 * j + 2], __uint_as_float(i));
    Q.w[0][i] = make_float4(1.f, 1.f, 1.f, 1.f);
    Q.rng[0][i] = rng_seed(A.seed_pass, j);
}

// Per-wave chunks of a device-counted queue.  The chunk shrinks with the
// queue (rays_per_wave over a quarter of the grid's waves, so a wave takes
// about four chunks) to keep every wave busy on short queues.  A queue of at
// most one chunk per wave is dealt statically (wave w takes chunk w) with no
// atomics: every wave hitting one device-wide counter at once serialises,
// and on 1-spp passes (262 k paths) dropping those atomics made a pass 21 %
// faster.  Longer queues hand out chunks by atomic counters, one per XCD
// (blocks are dealt to the XCDs round-robin): XCD x takes chunks x, x + 8,
// x + 16, ... so neighbouring chunks still spread over the XCDs.  (A static
// first chunk per wave measured 5 % slower on long queues.)
struct DynamicChunks {
    uint32_t n;
    uint32_t* xw;  // this queue's eight counters, 16 words apart
    uint32_t size, nchunks, nwaves, wave;
    bool first;
    __device__ DynamicChunks(uint32_t n_, uint32_t* xw_) : n(n_), xw(xw_), first(true) {
        nwaves = (gridDim.x * blockDim.x) >> 6;
        wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        uint32_t c = (uint32_t)(((uint64_t)n + 16ull * nwaves - 1) / (16ull * nwaves));
        c = (c + 63u) & ~63u;
        size = c < 64u ? 64u : (c > (uint32_t)kChunk ? (uint32_t)kChunk : c);
        nchunks = (n + size - 1) / size;
    }
    __device__ bool operator()(uint32_t& b, uint32_t& e) {
        uint32_t idx;
        if (nchunks <= nwaves) {
            if (!first) return false;
            first = false;
            idx = wave;
        } else {
            const uint32_t groups = gridDim.x >= kXcds ? kXcds : 1u;  // every group has waves
            const uint32_t x = blockIdx.x % groups;
            uint32_t k = 0;
            if ((threadIdx.x & 63u) == 0) k = atomicAdd(xw + 16 * x, 1u);
            idx = x + groups * __shfl(k, 0, 64);
        }
        if (idx >= nchunks) return false;
        b = idx * size;
        e = min(b + size, n);
        return true;
    }
};

// An unoccluded shadow ray: its contribution goes to the pixel (and, when
// training, its radiance to the segment's record).
__device__ __forceinline__ void shadow_lit(const RenderArgs& A, const Queues& Q, int sb, uint32_t k) {
    accumulate(A.accum, __float_as_uint(Q.so[sb][k].w), Q.sc[sb][k]);
    if (A.rec_x) {
        const float4 x = Q.sx[sb][k];
        float* dst = reinterpret_cast<float*>(A.rec_x + __float_as_uint(x.w));
        atomicAdd(dst + 0, x.x);
        atomicAdd(dst + 1, x.y);
        atomicAdd(dst + 2, x.z);
    }
}

// REF: the refined tree of a TN_SCENE_REFINED scene; rays tied at their
// closest distance and rays with a zero direction component get the hit
// record kFixId, which k_fix_hits replaces before anything reads it, and go
// on the level's fix list.
template <bool EXH, bool REF>
__device__ __forceinline__ void trace_body(const RenderArgs& A, const Queues& Q, int level, int buf, uint2* stk,
                                           std::conditional_t<REF, CoopLdsTie, CoopLds>* wl) {
    const uint32_t n = min(Q.cnt[level], Q.cap);  // overflow is reported, never read
    const float4* qo = Q.o[buf];
    const float4* qd = Q.d[buf];
    uint4* hit = Q.hit;
    TraversalCounters cnt;
    coop_loop<kRStack, EXH, false, false, false, REF>(
        A.sv, stk + threadIdx.x, kRBlock, wl + threadIdx.x / 64,
        DynamicChunks{n, Q.xwork + 16 * kXcds * level},
        [&](uint32_t k, float3& o, float3& d, float&, uint32_t& scene) {
            const float4 a = qo[k], b = qd[k];
            o = make_float3(a.x, a.y, a.z);
            d = make_float3(b.x, b.y, b.z);
            scene = scene_of_pixel(A, __float_as_uint(a.w));
        },
        [&](uint32_t k, const Trav& T, bool traversed) {
            if (REF && (T.tie || needs_reference(T.d))) {
                hit[k] = make_uint4(kFixId, 0u, 0u, 0u);
                const uint32_t slot = atomicAdd(Q.cnt + kFixHits + level, 1u);
                if (slot < kFixCap) Q.fix[slot] = k;
                return;
            }
            const bool h = traversed && T.best_id != kInvalidId;
            hit[k] = make_uint4(h ? T.best_id : kInvalidId, __float_as_uint(h ? T.best_r : FLT_MAX),
                                __float_as_uint(h ? T.best_s : 0.f), __float_as_uint(h ? T.best_t : 0.f));
        },
        cnt);
    if (cnt.incomplete) atomicAdd(A.status + 2, 1u);
}

template <bool EXH, bool REF>
__global__ __launch_bounds__(kRBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_trace(
    RenderArgs A, Queues Q, int level, int buf) {
    __shared__ uint2 stk[kRStack * kRBlock];
    __shared__ std::conditional_t<REF, CoopLdsTie, CoopLds> wl[kRBlock / 64];
    trace_body<EXH, REF>(A, Q, level, buf, stk, wl);
}

// The segments of one level that k_trace<..., true> marked, traced one per
// lane on the reference tree (the reference's leaves and order decide them):
// the entries of the fix list, or every marked entry of the queue when the
// list overflowed.
template <bool EXH>
__global__ __launch_bounds__(kRBlock) void k_fix_hits(RenderArgs A, Queues Q, int level, int buf) {
    __shared__ uint2 stk[kRStack * kRBlock];
    const uint32_t m = Q.cnt[kFixHits + level];
    if (m == 0) return;
    const bool scan = m > kFixCap;
    const uint32_t n = scan ? min(Q.cnt[level], Q.cap) : m;
    TraversalCounters cnt;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t i = scan ? j : Q.fix[j];
        if (scan && Q.hit[i].x != kFixId) continue;
        const float4 a = Q.o[buf][i], b = Q.d[buf][i];
        float r = FLT_MAX, s = 0.f, t = 0.f;
        const uint32_t id = traverse<kRStack, EXH, false, false, false>(
            A.sv_exact, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), 0.f, r, s, t, stk + threadIdx.x,
            kRBlock, cnt);
        const bool h = id != kInvalidId;
        Q.hit[i] = make_uint4(h ? id : kInvalidId, __float_as_uint(h ? r : FLT_MAX), __float_as_uint(h ? s : 0.f),
                              __float_as_uint(h ? t : 0.f));
    }
    if (cnt.incomplete) atomicAdd(A.status + 2, 1u);
}

// Shadow rays (any hit up to tmax).  REF: the refined tree; a ray with a zero
// direction component is marked (sign bit of its tmax) and listed for
// k_fix_shadows.
template <bool EXH, bool REF>
__device__ __forceinline__ void shadow_body(const RenderArgs& A, const Queues& Q, int level, uint2* stk,
                                            CoopLds* wl) {
    const uint32_t n = min(Q.cnt[kMaxLevels + level], Q.scap);
    const int sb = level & 1;
    TraversalCounters cnt;
    coop_loop<kRStack, EXH, false, true, false>(
        A.sv, stk + threadIdx.x, kRBlock, wl + threadIdx.x / 64,
        DynamicChunks{n, Q.xwork + 16 * kXcds * (kMaxLevels + level)},
        [&](uint32_t k, float3& o, float3& d, float& tmax, uint32_t& scene) {
            const float4 a = Q.so[sb][k], b = Q.sd[sb][k];
            o = make_float3(a.x, a.y, a.z);
            d = make_float3(b.x, b.y, b.z);
            tmax = b.w;
            scene = scene_of_pixel(A, __float_as_uint(a.w));
        },
        [&](uint32_t k, const Trav& T, bool traversed) {
            if (REF && needs_reference(T.d)) {
                float* w = reinterpret_cast<float*>(Q.sd[sb] + k) + 3;
                *w = __uint_as_float(__float_as_uint(*w) | 0x80000000u);
                const uint32_t slot = atomicAdd(Q.cnt + kFixShadows + level, 1u);
                if (slot < kFixCap) Q.fix[(1 + sb) * kFixCap + slot] = k;
                return;
            }
            if (!traversed || T.best_id == kInvalidId) shadow_lit(A, Q, sb, k);
        },
        cnt);
    if (cnt.incomplete) atomicAdd(A.status + 2, 1u);
}

template <bool EXH, bool REF>
__global__ __launch_bounds__(kRBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_shadow(
    RenderArgs A, Queues Q, int level) {
    __shared__ uint2 stk[kRStack * kRBlock];
    __shared__ CoopLds wl[kRBlock / 64];
    shadow_body<EXH, REF>(A, Q, level, stk, wl);
}

// The shadow rays of one level that k_shadow<..., true> marked, tested one
// per lane on the reference tree: the fix list's entries, or every marked
// entry of the queue when the list overflowed.
template <bool EXH>
__global__ __launch_bounds__(kRBlock) void k_fix_shadows(RenderArgs A, Queues Q, int level) {
    __shared__ uint2 stk[kRStack * kRBlock];
    const uint32_t m = Q.cnt[kFixShadows + level];
    if (m == 0) return;
    const bool scan = m > kFixCap;
    const int sb = level & 1;
    const uint32_t n = scan ? min(Q.cnt[kMaxLevels + level], Q.scap) : m;
    TraversalCounters cnt;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t i = scan ? j : Q.fix[(1 + sb) * kFixCap + j];
        const float4 b = Q.sd[sb][i];
        const uint32_t wbits = __float_as_uint(b.w);
        if (!(wbits >> 31)) continue;
        const float4 a = Q.so[sb][i];
        float r = FLT_MAX, s = 0.f, t = 0.f;
        const uint32_t id = traverse<kRStack, EXH, false, true, false>(
            A.sv_exact, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), __uint_as_float(wbits & 0x7FFFFFFFu),
            r, s, t, stk + threadIdx.x, kRBlock, cnt);
        if (id == kInvalidId) shadow_lit(A, Q, sb, i);
    }
    if (cnt.incomplete) atomicAdd(A.status + 2, 1u);
}

// Orthonormal basis around n (Duff et al. 2017, "Building an Orthonormal
// Basis, Revisited").
__device__ __forceinline__ void onb(float3 n, float3& b1, float3& b2) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    b1 = make_float3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b2 = make_float3(b, sign + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ void push_shadow(const Queues& Q, int sb, uint32_t k, float3 o, uint32_t pixel,
                                            float3 dir, float tmax, float4 contrib, float4 local, uint32_t slot,
                                            bool records) {
    Q.so[sb][k] = make_float4(o.x, o.y, o.z, __uint_as_float(pixel));
    Q.sd[sb][k] = make_float4(dir.x, dir.y, dir.z, tmax);
    Q.sc[sb][k] = contrib;
    if (records) Q.sx[sb][k] = make_float4(local.x, local.y, local.z, __uint_as_float(slot));
}

// ------------------------------------------------ raycaster and raytracer
// The reference's other two trace() plugins on the same queues and traversal.

// Raycaster shade (raycaster.cpp:7-24), level 0 only: a miss gives the
// background; a hit the triangle's diffuse colour with alpha
// clamp(1 - r / max_visibility, 0, 1).  Hits are counted in cnt[4 kMaxLevels + 1]
// (the reference counts num_rays on hits only).
__global__ __launch_bounds__(kRBlock) void k_cast(RenderArgs A, Queues Q, int level, int buf) {
    const uint32_t n = min(Q.cnt[level], Q.cap);
    uint32_t hits = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 qo = Q.o[buf][i];
        const float4 w = Q.w[buf][i];
        const uint32_t pixel = __float_as_uint(qo.w);
        const SceneParams& S = A.sp[scene_of_pixel(A, pixel)];
        const uint4 h = Q.hit[i];
        if (h.x == kInvalidId) {
            accumulate(A.accum, pixel, mul4(w, S.bg));
            continue;
        }
        ++hits;
        float4 c = A.albedo[h.x];
        const float a = 1.f - __uint_as_float(h.y) / A.max_vis;
        c.w = a < 0.f ? 0.f : (a > 1.f ? 1.f : a);  // clamp of src/turner.h:28-36
        accumulate(A.accum, pixel, mul4(w, c));
    }
    for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off, 64);
    if ((threadIdx.x & 63u) == 0 && hits) atomicAdd(Q.cnt + 4 * kMaxLevels + 1, hits);
}

// Raytracer shade (raytracer.cpp:6-67) of one level.  A segment carries the
// product W of the factors reflectivity * reflective (and shadow darkening)
// of its ancestors.  A hit contributes W * colour, colour = Lambert
// max(0, L.n) diffuse light (lib/lambertian.h:15-19), times (1 - reflectivity)
// when it reflects, and spawns the mirror ray d - 2 (n.d) n from p2 with
// weight W * reflectivity * reflective.  The contribution waits in the shadow
// queue with the child's index: k_rt_shadow darkens both by the shadow
// intensity when the light is blocked (raytracer.cpp:57-64).  Children beyond
// max_depth are counted (the reference counts the call that returns {})
// but not queued.
__global__ __launch_bounds__(kRBlock) void k_rt_shade(RenderArgs A, Queues Q, int level, int buf) {
    const uint32_t n = min(Q.cnt[level], Q.cap);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const int ob = buf ^ 1;
    for (uint32_t base = wave * 64u; base < n; base += nwaves * 64u) {
        const uint32_t i = base + lane;
        bool has_hit = false, spawn = false, shadow = false;
        float3 p2 = make_float3(0, 0, 0), rdir = p2, sdir = p2;
        float sdist = 0.f;
        float4 contrib = make_float4(0, 0, 0, 0), cw = contrib;
        uint32_t pixel = 0;
        float pathbits = 0.f;
        if (i < n) {
            const float4 qo = Q.o[buf][i], qd = Q.d[buf][i];
            const float4 w = Q.w[buf][i];
            pixel = __float_as_uint(qo.w);
            pathbits = qd.w;
            const SceneParams& S = A.sp[scene_of_pixel(A, pixel)];
            const uint4 h = Q.hit[i];
            if (h.x == kInvalidId) {
                accumulate(A.accum, pixel, mul4(w, S.bg));  // raytracer.cpp:19-21
            } else {
                has_hit = true;
                const uint32_t id = h.x;
                const float r = __uint_as_float(h.y), s = __uint_as_float(h.z), t = __uint_as_float(h.w);
                const float3 p = make_float3(qo.x + r * qd.x, qo.y + r * qd.y, qo.z + r * qd.z);
                // raytracer.cpp:24-26: light direction from p
                const float3 l = make_float3(S.light_pos.x - p.x, S.light_pos.y - p.y, S.light_pos.z - p.z);
                const float il = 1.0f / sqrtf(l.x * l.x + l.y * l.y + l.z * l.z);
                const float3 ld = make_float3(il * l.x, il * l.y, il * l.z);
                // raytracer.cpp:29-30 (lib/triangle.h:54-56)
                const float* nv = A.normals + 9 * (size_t)id;
                const float b0 = 1.f - s - t;
                const float3 ni = make_float3(b0 * nv[0] + s * nv[3] + t * nv[6], b0 * nv[1] + s * nv[4] + t * nv[7],
                                              b0 * nv[2] + s * nv[5] + t * nv[8]);
                const float inv = 1.0f / sqrtf(ni.x * ni.x + ni.y * ni.y + ni.z * ni.z);
                const float3 nn = make_float3(inv * ni.x, inv * ni.y, inv * ni.z);
                // raytracer.cpp:33-34
                const float cosa = fmaxf(0.f, ld.x * nn.x + ld.y * nn.y + ld.z * nn.z);
                const float4 direct = mul4(scale4(cosa, A.albedo[id]), S.light_col);
                // raytracer.cpp:37
                p2 = make_float3(p.x + nn.x * 0.0001f, p.y + nn.y * 0.0001f, p.z + nn.z * 0.0001f);
                float4 color = direct;
                const float refl = A.reflectivity ? A.reflectivity[id] : 0.f;
                if (refl > 0) {  // raytracer.cpp:42-52
                    const float dn = 2.f * (nn.x * qd.x + nn.y * qd.y + nn.z * qd.z);
                    rdir = make_float3(qd.x - dn * nn.x, qd.y - dn * nn.y, qd.z - dn * nn.z);
                    color = scale4(1.f - refl, direct);
                    cw = mul4(w, scale4(refl, A.reflective[id]));
                    spawn = true;
                }
                contrib = mul4(w, color);
                // raytracer.cpp:55-60: shadow ray from p2 toward the light
                const float3 l2 = make_float3(S.light_pos.x - p2.x, S.light_pos.y - p2.y, S.light_pos.z - p2.z);
                sdist = sqrtf(l2.x * l2.x + l2.y * l2.y + l2.z * l2.z);
                const float il2 = 1.0f / sdist;
                sdir = make_float3(il2 * l2.x, il2 * l2.y, il2 * l2.z);
                shadow = sdist > 0.f;
            }
        }
        // children (mirror rays)
        const uint64_t mc = __ballot(spawn);
        uint32_t child = ~0u;
        if (mc) {
            uint32_t cbase = 0;
            if (lane == 0) cbase = atomicAdd(&Q.cnt[level + 1], (uint32_t)__popcll(mc));
            cbase = __shfl(cbase, 0);
            if (spawn && level < A.max_depth) {
                const uint32_t k = cbase + lanes_below(mc);
                if (k < Q.cap) {
                    Q.o[ob][k] = make_float4(p2.x, p2.y, p2.z, __uint_as_float(pixel));
                    Q.d[ob][k] = make_float4(rdir.x, rdir.y, rdir.z, pathbits);
                    Q.w[ob][k] = cw;
                    child = k;
                } else {
                    atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
                }
            }
        }
        // lit test; a hit whose light sits at p2 itself is lit
        if (has_hit && !shadow) accumulate(A.accum, pixel, contrib);
        const uint64_t ms = __ballot(shadow);
        if (ms) {
            uint32_t sbase = 0;
            if (lane == 0) sbase = atomicAdd(&Q.cnt[kMaxLevels + level], (uint32_t)__popcll(ms));
            sbase = __shfl(sbase, 0);
            if (shadow) {
                const uint32_t k = sbase + lanes_below(ms);
                if (k < Q.scap) {
                    // shadowed iff the closest hit is nearer than the light
                    // (strict): any hit with r <= the float below sdist
                    const float tmax = __uint_as_float(__float_as_uint(sdist) - 1u);
                    const int sb = level & 1;
                    Q.so[sb][k] = make_float4(p2.x, p2.y, p2.z, __uint_as_float(pixel));
                    Q.sd[sb][k] = make_float4(sdir.x, sdir.y, sdir.z, tmax);
                    Q.sc[sb][k] = contrib;
                    Q.sx[sb][k] = make_float4(__uint_as_float(child), 0.f, 0.f, 0.f);
                } else {
                    atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
                }
            }
        }
    }
}

// Raytracer shadow test (raytracer.cpp:55-64): a blocked light darkens the
// segment's colour, and so the weight of its mirror child, by
// c -= c * shadow_intensity.
template <bool EXH>
__global__ __launch_bounds__(kRBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_rt_shadow(
    RenderArgs A, Queues Q, int level, int buf) {
    __shared__ uint2 stk[kRStack * kRBlock];
    __shared__ CoopLds wl[kRBlock / 64];
    const uint32_t n = min(Q.cnt[kMaxLevels + level], Q.scap);
    const int ob = buf ^ 1;
    const int sb = level & 1;
    TraversalCounters cnt;
    // the raytracer's shadow test darkens the mirror child too: it runs on the
    // reference tree itself (no marks to resolve afterwards)
    const SceneView& sv = A.refined ? A.sv_exact : A.sv;
    coop_loop<kRStack, EXH, false, true, false>(
        sv, stk + threadIdx.x, kRBlock, wl + threadIdx.x / 64,
        DynamicChunks{n, Q.xwork + 16 * kXcds * (kMaxLevels + level)},
        [&](uint32_t k, float3& o, float3& d, float& tmax, uint32_t& scene) {
            const float4 a = Q.so[sb][k], b = Q.sd[sb][k];
            o = make_float3(a.x, a.y, a.z);
            d = make_float3(b.x, b.y, b.z);
            tmax = b.w;
            scene = scene_of_pixel(A, __float_as_uint(a.w));
        },
        [&](uint32_t k, const Trav& T, bool traversed) {
            float4 c = Q.sc[sb][k];
            if (traversed && T.best_id != kInvalidId) {
                const float si = A.shadow_si;
                c = make_float4(c.x - c.x * si, c.y - c.y * si, c.z - c.z * si, c.w - c.w * si);
                const uint32_t child = __float_as_uint(Q.sx[sb][k].x);
                if (child != ~0u) {
                    const float4 w = Q.w[ob][child];
                    Q.w[ob][child] = make_float4(w.x - w.x * si, w.y - w.y * si, w.z - w.z * si, w.w - w.w * si);
                }
            }
            accumulate(A.accum, __float_as_uint(Q.so[sb][k].w), c);
        },
        cnt);
    if (cnt.incomplete) atomicAdd(A.status + 2, 1u);
}

// Shade one level.  GUIDED kernels run one wave per block (the guide's
// forward pass is separated by block barriers).
template <bool GUIDED>
__device__ __forceinline__ void shade_body(const RenderArgs& A, const Queues& Q, int level, int buf,
                                           uint16_t* glds) {
    const uint32_t n = min(Q.cnt[level], Q.cap);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const bool spawn = level < A.max_depth && A.mc > 0;
    const uint32_t m = (uint32_t)max(A.mc, 0);
    const bool cosine = A.flags & TN_RENDER_COSINE;
    const bool area_flag = (A.flags & TN_RENDER_AREA_LIGHTS) != 0;
    const bool records = GUIDED && A.rec_p != nullptr;
    const float inv_pi = 0.318309886183790671538f;
    for (uint32_t base = wave * 64u; base < n; base += nwaves * 64u) {
        const uint32_t i = base + lane;
        const bool active = i < n;
        bool has_hit = false, has_point = false, has_area = false;
        float4 w = make_float4(0, 0, 0, 0), rho = w;
        float3 p = make_float3(0, 0, 0), p2 = p, nn = p, ldir = p, adir = p;
        float ldist = 0.f, adist = 0.f;
        float4 lcontrib = w, acontrib = w, llocal = w, alocal = w;
        uint32_t pixel = 0, slot = 0;
        uint4 rng = make_uint4(1, 0, 0, 0);
        if (active) {
            const float4 qo = Q.o[buf][i], qd = Q.d[buf][i];
            w = Q.w[buf][i];
            rng = Q.rng[buf][i];
            pixel = __float_as_uint(qo.w);
            slot = __float_as_uint(qd.w) * A.rstride + (uint32_t)level;
            const SceneParams& S = A.sp[scene_of_pixel(A, pixel)];
            const uint4 h = Q.hit[i];
            if (h.x == kInvalidId) {
                accumulate(A.accum, pixel, mul4(w, S.bg));  // pathtracer.cpp:26-28
                if (records) A.rec_x[slot] = S.bg;
            } else {
                has_hit = true;
                const uint32_t id = h.x;
                const float r = __uint_as_float(h.y), s = __uint_as_float(h.z), t = __uint_as_float(h.w);
                // pathtracer.cpp:30-36
                p = make_float3(qo.x + r * qd.x, qo.y + r * qd.y, qo.z + r * qd.z);
                const float* nv = A.normals + 9 * (size_t)id;
                const float b0 = 1.f - s - t;
                float3 ni = make_float3(b0 * nv[0] + s * nv[3] + t * nv[6], b0 * nv[1] + s * nv[4] + t * nv[7],
                                        b0 * nv[2] + s * nv[5] + t * nv[8]);
                const float inv = 1.0f / sqrtf(ni.x * ni.x + ni.y * ni.y + ni.z * ni.z);
                nn = make_float3(inv * ni.x, inv * ni.y, inv * ni.z);
                p2 = make_float3(p.x + 0.0001f * nn.x, p.y + 0.0001f * nn.y, p.z + 0.0001f * nn.z);
                rho = A.albedo[id];
                const float4 wr = mul4(w, rho);
                if (S.nlights) {  // pathtracer.cpp:44-58
                    float3 l = make_float3(S.light_pos.x - p.x, S.light_pos.y - p.y, S.light_pos.z - p.z);
                    const float il = 1.0f / sqrtf(l.x * l.x + l.y * l.y + l.z * l.z);
                    ldir = make_float3(il * l.x, il * l.y, il * l.z);
                    const float3 l2 = make_float3(S.light_pos.x - p2.x, S.light_pos.y - p2.y, S.light_pos.z - p2.z);
                    ldist = sqrtf(l2.x * l2.x + l2.y * l2.y + l2.z * l2.z);
                    const float c = fmaxf(0.f, ldir.x * nn.x + ldir.y * nn.y + ldir.z * nn.z);
                    lcontrib = scale4(inv_pi, mul4(wr, scale4(c, S.light_col)));
                    if (records) llocal = scale4(inv_pi, mul4(rho, scale4(c, S.light_col)));
                    has_point = true;
                }
                if (area_flag && S.nemit > 0) {
                    const float4 le = A.emission[id];
                    const float4 g = A.trirec[4 * (size_t)id + 1];  // geometric normal
                    if (level == 0 && (le.x > 0 || le.y > 0 || le.z > 0) &&
                        g.x * qd.x + g.y * qd.y + g.z * qd.z < 0)
                        accumulate(A.accum, pixel, mul4(w, le));
                    const float u0 = rng_uniform(rng), u1 = rng_uniform(rng), u2 = rng_uniform(rng);
                    // one emitter k with probability pk: by solid angle among
                    // the front-facing ones of a scene with few emitters, else
                    // by area
                    uint32_t k = 0;
                    float pk = 0.f;
                    if (S.nemit <= kSolidAngleSelect) {
                        float wk[kSolidAngleSelect];
                        float wsum = 0.f;
#pragma unroll
                        for (uint32_t j = 0; j < kSolidAngleSelect; ++j) {
                            wk[j] = 0.f;
                            if (j < S.nemit) {
                                const uint32_t kj = A.emit_ids[S.emit_begin + j];
                                const float4 r0 = A.trirec[4 * (size_t)kj], rn = A.trirec[4 * (size_t)kj + 1],
                                             ru = A.trirec[4 * (size_t)kj + 2], rv = A.trirec[4 * (size_t)kj + 3];
                                const float3 a = make_float3(r0.x - p.x, r0.y - p.y, r0.z - p.z);
                                if (-(a.x * rn.x + a.y * rn.y + a.z * rn.z) > 0.f)  // p in front of the emitter
                                    wk[j] = tri_solid_angle(a, v3axpy(1.f, make_float3(ru.x, ru.y, ru.z), a),
                                                            v3axpy(1.f, make_float3(rv.x, rv.y, rv.z), a));
                            }
                            wsum += wk[j];
                        }
                        const float x = u0 * wsum;
                        float run = 0.f;
                        bool picked = false;
#pragma unroll
                        for (uint32_t j = 0; j < kSolidAngleSelect; ++j) {
                            run += wk[j];
                            if (!picked && wk[j] > 0.f && (x < run || run >= wsum)) {
                                picked = true;
                                k = A.emit_ids[S.emit_begin + j];
                                pk = wk[j] / wsum;
                            }
                        }
                    } else {
                        const float x = u0 * S.emit_area;
                        const float* cdf = A.emit_cdf + S.emit_begin;
                        uint32_t lo = 0, hi = S.nemit - 1;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (cdf[mid] <= x) lo = mid + 1;
                            else hi = mid;
                        }
                        k = A.emit_ids[S.emit_begin + lo];
                        const float cprev = lo ? cdf[lo - 1] : 0.f;
                        pk = (cdf[lo] - cprev) / S.emit_area;
                    }
                    if (pk > 0.f) {
                        const float4 e0 = A.trirec[4 * (size_t)k], en = A.trirec[4 * (size_t)k + 1],
                                     eu = A.trirec[4 * (size_t)k + 2], ev = A.trirec[4 * (size_t)k + 3];
                        const float3 a = make_float3(e0.x - p.x, e0.y - p.y, e0.z - p.z);
                        const float3 n3 = make_float3(en.x, en.y, en.z);
                        const float3 eu3 = make_float3(eu.x, eu.y, eu.z), ev3 = make_float3(ev.x, ev.y, ev.z);
                        // direction and inverse density per steradian
                        float inv_pdf = 0.f;
                        float3 q;
                        SphTri st;
                        if (sph_tri_setup(a, v3axpy(1.f, eu3, a), v3axpy(1.f, ev3, a), st) &&
                            st.omega >= kMinSolidAngle) {
                            adir = sph_tri_sample(st, u1, u2);
                            const float dn = v3dot(adir, n3);  // = -cos_l
                            const float tq = dn < 0.f ? v3dot(a, n3) / dn : 0.f;
                            q = v3axpy(tq, adir, p);
                            inv_pdf = dn < 0.f ? st.omega : 0.f;
                        } else {  // uniform point by area: density d^2 / (cos_l area)
                            const float su = sqrtf(u1);
                            const float bu = su * (1.0f - u2), bv = su * u2;
                            q = make_float3(e0.x + bu * eu.x + bv * ev.x, e0.y + bu * eu.y + bv * ev.y,
                                            e0.z + bu * eu.z + bv * ev.z);
                            const float3 dl = v3sub(q, p);
                            const float d2 = v3dot(dl, dl);
                            adir = v3scale(1.0f / sqrtf(d2), dl);
                            const float cl = -v3dot(adir, n3);
                            const float3 cr = v3cross(eu3, ev3);
                            inv_pdf = cl > 0.f ? cl * 0.5f * sqrtf(v3dot(cr, cr)) / d2 : 0.f;
                        }
                        const float cosx = v3dot(adir, nn);
                        if (cosx > 0 && inv_pdf > 0) {
                            const float3 q2 = v3sub(q, p2);
                            adist = sqrtf(v3dot(q2, q2)) * (1.0f - 1e-4f);
                            const float4 lek = A.emission[k];
                            const float gterm = cosx * inv_pdf / pk;
                            acontrib = scale4(inv_pi, mul4(wr, scale4(gterm, lek)));
                            if (records) alocal = scale4(inv_pi, mul4(rho, scale4(gterm, lek)));
                            has_area = true;
                        }
                    }
                }
            }
        }
        // shadow rays: wave-aggregated push
        const uint64_t mp = __ballot(has_point), ma = __ballot(has_area);
        const uint32_t tot = (uint32_t)(__popcll(mp) + __popcll(ma));
        if (tot) {
            uint32_t sbase = 0;
            if (lane == 0) sbase = atomicAdd(&Q.cnt[kMaxLevels + level], tot);
            sbase = __shfl(sbase, 0);
            if (has_point) {
                const uint32_t k = sbase + lanes_below(mp);
                if (k < Q.scap) push_shadow(Q, level & 1, k, p2, pixel, ldir, ldist, lcontrib, llocal, slot, records);
                else atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
            }
            if (has_area) {
                const uint32_t k = sbase + (uint32_t)__popcll(mp) + lanes_below(ma);
                if (k < Q.scap) push_shadow(Q, level & 1, k, p2, pixel, adir, adist, acontrib, alocal, slot, records);
                else atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
            }
        }
        if (!spawn) continue;
        const int ob = buf ^ 1;
        const float invm = 1.0f / (float)m;
        const float4 wr = mul4(w, rho);
        const float pathbits = Q.d[buf][active ? i : 0].w;
        float3 b1, b2;
        onb(nn, b1, b2);
        if (!GUIDED) {
            // children: m per hit (pathtracer.cpp:72-85)
            const uint64_t mc = __ballot(has_hit);
            if (!mc) continue;
            uint32_t cbase = 0;
            if (lane == 0) cbase = atomicAdd(&Q.cnt[level + 1], (uint32_t)__popcll(mc) * m);
            cbase = __shfl(cbase, 0);
            if (!has_hit) continue;
            const uint32_t first = cbase + lanes_below(mc) * m;
            for (uint32_t c = 0; c < m; ++c) {
                const float u1 = rng_uniform(rng), u2 = rng_uniform(rng);
                float lx, ly, lz, f;
                const float phi = 6.28318530717958647692f * u2;
                float sp, cp;
                __sincosf(phi, &sp, &cp);
                if (cosine) {
                    const float rr = sqrtf(u1);
                    lx = rr * cp;
                    ly = rr * sp;
                    lz = sqrtf(fmaxf(0.f, 1.f - u1));
                    f = invm;
                } else {  // lib/sampling.h:20-32
                    lz = u1;
                    const float rr = sqrtf(fmaxf(0.f, 1.f - lz * lz));
                    lx = rr * cp;
                    ly = rr * sp;
                    f = 2.f * u1 * invm;
                }
                const float3 dir = make_float3(lx * b1.x + ly * b2.x + lz * nn.x, lx * b1.y + ly * b2.y + lz * nn.y,
                                               lx * b1.z + ly * b2.z + lz * nn.z);
                const uint32_t k = first + c;
                if (k < Q.cap) {
                    Q.o[ob][k] = make_float4(p2.x, p2.y, p2.z, __uint_as_float(pixel));
                    Q.d[ob][k] = make_float4(dir.x, dir.y, dir.z, pathbits);
                    Q.w[ob][k] = scale4(f, wr);
                    Q.rng[ob][k] = child_seed(rng, c);
                } else {
                    atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
                }
            }
        } else {
            // guided: the wave's 64 hits through the guide MLP
            float* gout = reinterpret_cast<float*>(glds);
            guide_forward_wave(A.gv, p, nn, has_hit, gout, glds);
            Mixture M;
            mixture_from(gout + lane * kGOut, M);
            __syncthreads();  // gout is overwritten by the next iteration's forward pass
            for (uint32_t c = 0; c < m; ++c) {
                const float us = rng_uniform(rng), u0 = rng_uniform(rng), u1 = rng_uniform(rng),
                            u2 = rng_uniform(rng);
                float3 dir;
                if (us < A.mis) {
                    dir = mixture_sample(M, u0, u1, u2);
                } else {
                    const float rr = sqrtf(u1);
                    float sp, cp;
                    __sincosf(6.28318530717958647692f * u2, &sp, &cp);
                    const float lx = rr * cp, ly = rr * sp, lz = sqrtf(fmaxf(0.f, 1.f - u1));
                    dir = make_float3(lx * b1.x + ly * b2.x + lz * nn.x, lx * b1.y + ly * b2.y + lz * nn.y,
                                      lx * b1.z + ly * b2.z + lz * nn.z);
                }
                const float cosn = dir.x * nn.x + dir.y * nn.y + dir.z * nn.z;
                const bool ok = has_hit && cosn > 0.f;
                float f = 0.f, cos_pdf = 0.f;
                if (ok) {
                    const float pdf = A.mis * mixture_pdf(M, dir) + (1.0f - A.mis) * cosn * inv_pi;
                    cos_pdf = cosn / pdf;
                    f = inv_pi * cos_pdf * invm;
                }
                const uint64_t mc = __ballot(ok);
                uint32_t cbase = 0;
                if (mc && lane == 0) cbase = atomicAdd(&Q.cnt[level + 1], (uint32_t)__popcll(mc));
                cbase = __shfl(cbase, 0);
                if (records && has_hit) {
                    A.rec_p[slot] = make_float4(p.x, p.y, p.z, cos_pdf);
                    A.rec_n[slot] = make_float4(nn.x, nn.y, nn.z, ok ? 1.f : 0.f);
                    A.rec_w[slot] = make_float4(dir.x, dir.y, dir.z, 0.f);
                    A.rec_g[slot] = scale4(f, rho);
                }
                if (!ok) continue;
                const uint32_t k = cbase + lanes_below(mc);
                if (k < Q.cap) {
                    Q.o[ob][k] = make_float4(p2.x, p2.y, p2.z, __uint_as_float(pixel));
                    Q.d[ob][k] = make_float4(dir.x, dir.y, dir.z, pathbits);
                    Q.w[ob][k] = scale4(f, wr);
                    Q.rng[ob][k] = child_seed(rng, c);
                } else {
                    atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
                }
            }
        }
    }
}

template <bool GUIDED>
__global__ __launch_bounds__(kRBlock) void k_shade(RenderArgs A, Queues Q, int level, int buf) {
    __shared__ __align__(16) uint16_t glds[GUIDED ? 64 * (kGStrideIn + kGStrideHid) : 1];
    shade_body<GUIDED>(A, Q, level, buf, glds);
}

// The unguided shade capped at 72 VGPRs (7 waves per SIMD, 68 bytes of
// spill) instead of its natural 99 (5 waves): at 1920x1080, depth 3, m 4
// the render went 1151 -> 1191 Mrays/s on the rendering tree (6 waves, 80
// VGPRs: 1157).
__global__ __launch_bounds__(kRBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_shade_plain(
    RenderArgs A, Queues Q, int level, int buf) {
    __shared__ uint16_t glds[1];
    shade_body<false>(A, Q, level, buf, glds);
}

// ---------------------------------------------------------- persistent
// Grid barrier of the persistent kernel (all blocks resident: cooperative
// launch), two-level so that no counter sees more than a few hundred
// atomics: the blocks count in at one of eight group counters (block b in
// group b % 8, the XCD it is dealt to), the last of a group at the group
// total, and the last group publishes the next generation, which the others
// poll.  Words, each on its own 128-byte line: [0] generation, [32] set when
// a wait ran out, [64] groups arrived, [96 + 32 g] arrivals in group g.  A
// wait is bounded (2^22 polls of about 64 cycles): then [32] is set, and
// every block stops at its next barrier (grid_sync returns false).
constexpr uint32_t kBarrierPolls = 1u << 22;
constexpr uint32_t kBarGen = 0, kBarOut = 32, kBarGroups = 64, kBarGroup = 96;
constexpr uint32_t kBarWords = kBarGroup + 32 * kXcds;

__device__ __forceinline__ bool grid_sync(uint32_t* bar, uint32_t& gen) {
    __shared__ uint32_t ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t good = __hip_atomic_load(bar + kBarOut, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
        if (good) {
            __threadfence();
            const uint32_t target = gen + 1u;
            const uint32_t groups = gridDim.x < kXcds ? gridDim.x : kXcds;
            const uint32_t g = blockIdx.x % groups;
            const uint32_t members = gridDim.x / groups + (g < gridDim.x % groups ? 1u : 0u);
            bool released = false;
            if (atomicAdd(bar + kBarGroup + 32 * g, 1u) == members - 1u) {
                atomicExch(bar + kBarGroup + 32 * g, 0u);  // reset before the group counts in
                __threadfence();
                if (atomicAdd(bar + kBarGroups, 1u) == groups - 1u) {
                    atomicExch(bar + kBarGroups, 0u);
                    __threadfence();
                    atomicExch(bar + kBarGen, target);
                    released = true;
                }
            }
            if (!released) {
                for (uint32_t polls = 0;
                     __hip_atomic_load(bar + kBarGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target; ++polls) {
                    if (polls >= kBarrierPolls) {
                        atomicExch(bar + kBarOut, 1u);
                        good = 0u;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __threadfence();
        }
        ok = good;
    }
    __syncthreads();
    ++gen;
    return ok != 0u;
}

// k_chunk_begin for one block of any size (block 0 of the persistent kernel).
__device__ __forceinline__ void chunk_begin_block(uint32_t* cnt, uint32_t* xwork, ChunkState* cs) {
    const unsigned long long first = cs->next;
    const unsigned long long left = cs->total > first ? cs->total - first : 0ull;
    const uint32_t count = (uint32_t)(left < cs->chunk ? left : cs->chunk);
    for (uint32_t i = threadIdx.x; i < 4 * kMaxLevels + 2; i += blockDim.x)
        if (i != 4 * kMaxLevels) cnt[i] = (i == 0) ? count : 0u;  // the overflow word lives for the whole render
    for (uint32_t k = kFixHits + threadIdx.x; k < 6 * kMaxLevels + 2; k += blockDim.x) cnt[k] = 0u;
    for (uint32_t k = threadIdx.x; k < 2 * kMaxLevels * kXcds; k += blockDim.x) xwork[16 * k] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
        cs->first = first;
        cs->count = count;
        cs->next = first + cs->stride;
    }
}

// Plain path tracing (no guide, no training records, no refined tree) of
// every chunk of a call in one cooperative launch: per chunk, raygen, then
// per level the same trace, shade and shadow stages as the wavefront
// kernels, separated by grid barriers instead of kernel boundaries.  One
// pipeline; the shadow test of a level no longer overlaps the next level's
// tracing.  The chunk loop ends when a chunk is empty; every barrier wait is
// bounded.
template <bool EXH>
__global__ __launch_bounds__(kRBlock) void k_persist(RenderArgs A, Queues Q, int levels) {
    __shared__ uint2 stk[kRStack * kRBlock];
    __shared__ CoopLds wl[kRBlock / 64];
    __shared__ uint16_t glds[1];
    uint32_t gen = 0;
    for (;;) {
        if (blockIdx.x == 0) chunk_begin_block(Q.cnt, Q.xwork, A.cs);
        if (!grid_sync(Q.bar, gen)) return;
        const uint32_t count = __hip_atomic_load(&A.cs->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (count == 0u) return;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
            raygen_one(A, Q, i);
        if (!grid_sync(Q.bar, gen)) return;
        int buf = 0;
        for (int lv = 0; lv < levels; ++lv) {
            trace_body<EXH, false>(A, Q, lv, buf, stk, wl);
            if (!grid_sync(Q.bar, gen)) return;
            shade_body<false>(A, Q, lv, buf, glds);
            if (!grid_sync(Q.bar, gen)) return;
            shadow_body<EXH, false>(A, Q, lv, stk, wl);
            if (!grid_sync(Q.bar, gen)) return;
            buf ^= 1;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {  // k_chunk_totals
            unsigned long long a = 0, b = 0;
            for (int k = 0; k < levels; ++k) {
                a += Q.cnt[k];
                b += Q.cnt[kMaxLevels + k];
            }
            Q.totals[0] += a;
            Q.totals[1] += b;
        }
        __syncthreads();  // the counts are read before block 0 clears them for the next chunk
    }
}

__global__ void k_zero_bar(uint32_t* bar) {
    for (uint32_t i = threadIdx.x; i < kBarWords; i += blockDim.x) bar[i] = 0u;
}

// Clears the training records of a chunk: valid flags and found radiance.
__global__ void k_rec_clear(RenderArgs A) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.cs->count * A.rstride) return;
    A.rec_n[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    A.rec_x[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Walks each path of the chunk backwards (L_k = X_{k+1} + g_{k+1} L_{k+1})
// and appends (p, n, w, lum(L_k) cos / pdf) of every bounce that brought
// light back to the training set (kept with probability A.accept).  Sample
// layout: (p, c), (n, cos / pdf), (w, sample << 6 | level as bits), sample =
// the path's sample index within the call modulo 2^26.
__global__ __launch_bounds__(kRBlock) void k_collect(RenderArgs A, int levels_with_records) {
    const uint32_t npaths = A.cs->count;
    const uint64_t salt = A.seed_pass ^ (A.cs->first * 0x9E3779B97F4A7C15ull);
    const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool act = pid < npaths;
    float3 L = make_float3(0.f, 0.f, 0.f), gprev = L;
    bool have = false;
    for (int k = levels_with_records - 1; k >= 0; --k) {
        const uint32_t slot = (act ? pid : 0u) * A.rstride + (uint32_t)k;
        bool emit = false;
        float4 rp, rn, rw;
        float c = 0.f;
        if (act) {
            rn = A.rec_n[slot];
            if (rn.w == 0.f) {
                have = false;
            } else {
                const float4 x = A.rec_x[slot + 1];
                float3 Lk = make_float3(x.x, x.y, x.z);
                if (have) {
                    Lk.x += gprev.x * L.x;
                    Lk.y += gprev.y * L.y;
                    Lk.z += gprev.z * L.z;
                }
                const float4 g = A.rec_g[slot];
                rp = A.rec_p[slot];
                rw = A.rec_w[slot];
                L = Lk;
                gprev = make_float3(g.x, g.y, g.z);
                have = true;
                c = (0.2126f * Lk.x + 0.7152f * Lk.y + 0.0722f * Lk.z) * rp.w;
                if (c > 0.f && isfinite(c)) {
                    const uint64_t h = splitmix64(salt ^ ((uint64_t)pid << 8) ^ (uint64_t)k);
                    emit = (float)(h >> 40) * 0x1p-24f < A.accept;
                }
            }
        }
        const uint64_t me = __ballot(emit);
        if (!me) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(A.ts_count, (uint32_t)__popcll(me));
        base = __shfl(base, 0);
        if (!emit) continue;
        const uint32_t t = base + lanes_below(me);
        if (t < A.ts_cap) {
            A.ts[3 * (size_t)t] = make_float4(rp.x, rp.y, rp.z, c);
            A.ts[3 * (size_t)t + 1] = make_float4(rn.x, rn.y, rn.z, rp.w);
            const uint32_t sample = (uint32_t)(A.cs->first + pid) & ((1u << 26) - 1u);
            A.ts[3 * (size_t)t + 2] = make_float4(rw.x, rw.y, rw.z, __uint_as_float((sample << 6) | (uint32_t)k));
        }
    }
}

// The reference jitter sequence of one row: xorshift64* seeded 42,
// float = low 24 bits * 2^-24 (lib/xorshift.h:36-62, main.cpp:201-206).
std::vector<float> reference_jitter(uint32_t width, uint32_t spp) {
    std::vector<float> j(2 * (size_t)width * spp);
    uint64_t s = 42;
    for (float& v : j) {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        const uint64_t out = s * 2685821657736338717ull;
        v = std::ldexp((float)(out & 0xFFFFFFull), -24);
    }
    return j;
}

struct RenderBuffers {
    Queues q{};
    size_t bytes = 0;
    char* base = nullptr;
    ChunkState* cs = nullptr;
    SceneParams* sp = nullptr;
    uint32_t sp_cap = 0;
    // the shadow tests run on a second stream (default priority), ordered
    // against the main stream by one event per level and kind
    hipStream_t side = nullptr;
    hipEvent_t ev_shade[kMaxLevels] = {};
    hipEvent_t ev_shadow[kMaxLevels] = {};
};

int ensure_side_stream(RenderBuffers& rb) {
    if (rb.side) return TN_OK;
    TN_HIP(hipStreamCreateWithFlags(&rb.side, hipStreamNonBlocking));
    for (uint32_t k = 0; k < kMaxLevels; ++k) {
        TN_HIP(hipEventCreateWithFlags(&rb.ev_shade[k], hipEventDisableTiming));
        TN_HIP(hipEventCreateWithFlags(&rb.ev_shadow[k], hipEventDisableTiming));
    }
    return TN_OK;
}

// Queues, streams and events of the renders on one device.  A render call
// holds the context's lock from its first enqueued operation until its
// streams have drained, so concurrent calls on one device take turns; the
// buffers only grow (no allocation once a call of the same size has run).
constexpr int kMaxDevices = 64;

// Carves the queue buffers out of one allocation; with base == nullptr it
// only measures (every piece is rounded up to 256 bytes).
size_t carve(Queues& q, ChunkState** cs, char* base, uint32_t cap) {
    size_t off = 0;
    auto take = [&](size_t b) {
        char* r = base ? base + off : nullptr;
        off += (b + 255) & ~(size_t)255;
        return r;
    };
    for (int k = 0; k < 2; ++k) {
        q.o[k] = (float4*)take((size_t)cap * 16);
        q.d[k] = (float4*)take((size_t)cap * 16);
        q.w[k] = (float4*)take((size_t)cap * 16);
        q.rng[k] = (uint4*)take((size_t)cap * 16);
    }
    q.hit = (uint4*)take((size_t)cap * 16);
    q.scap = 2 * cap;
    for (int k = 0; k < 2; ++k) {
        q.so[k] = (float4*)take((size_t)q.scap * 16);
        q.sd[k] = (float4*)take((size_t)q.scap * 16);
        q.sc[k] = (float4*)take((size_t)q.scap * 16);
        q.sx[k] = (float4*)take((size_t)q.scap * 16);
    }
    q.cnt = (uint32_t*)take((6 * kMaxLevels + 2) * sizeof(uint32_t));
    q.fix = (uint32_t*)take(3 * (size_t)kFixCap * sizeof(uint32_t));
    q.xwork = (uint32_t*)take(2 * kMaxLevels * kXcds * 16 * sizeof(uint32_t));
    q.totals = (unsigned long long*)take(2 * sizeof(unsigned long long));
    q.bar = (uint32_t*)take(kBarWords * sizeof(uint32_t));
    *cs = (ChunkState*)take(sizeof(ChunkState));
    q.cap = cap;
    return off;
}

int alloc_queues(RenderBuffers& rb, uint32_t cap, uint32_t nscenes) {
    if (!rb.sp || rb.sp_cap < nscenes) {
        if (rb.sp) (void)hipFree(rb.sp);
        rb.sp = nullptr;
        rb.sp_cap = 0;
        const uint32_t want = std::max(nscenes, 64u);
        TN_HIP(hipMalloc(&rb.sp, (size_t)want * sizeof(SceneParams)));
        rb.sp_cap = want;
    }
    Queues probe{};
    ChunkState* pcs = nullptr;
    const size_t need = carve(probe, &pcs, nullptr, cap);
    if (rb.base && rb.bytes >= need && rb.q.cap >= cap) return TN_OK;
    if (rb.base) (void)hipFree(rb.base);
    rb.base = nullptr;
    rb.bytes = 0;
    TN_HIP(hipMalloc(&rb.base, need));
    rb.bytes = need;
    const size_t used = carve(rb.q, &rb.cs, rb.base, cap);
    return used <= need ? TN_OK : set_error(TN_ERR_INVALID, "queue layout exceeds its allocation");
}

struct RecordBuffers {
    float4* base = nullptr;
    size_t slots = 0;
};

int alloc_records(RecordBuffers& rb, size_t slots) {
    if (rb.base && rb.slots >= slots) return TN_OK;
    if (rb.base) (void)hipFree(rb.base);
    rb.base = nullptr;
    rb.slots = 0;
    TN_HIP(hipMalloc(&rb.base, slots * 5 * sizeof(float4)));
    rb.slots = slots;
    return TN_OK;
}

struct RenderContext {
    std::mutex mu;
    RenderBuffers rbs[kPipes];
    RecordBuffers recs[kPipes];
    // main streams of pipelines 1.. (pipeline 0 runs on the caller's stream)
    hipStream_t pipe_stream[kPipes] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kPipes] = {};
    float* jitter = nullptr;  // reference pixel jitter (TN_RENDER_REF_JITTER)
    size_t jitter_cap = 0;
    // graph replay (TN_RENDER_GRAPH): the capture stream and the executable
    // graph of the last captured chunk sequence, updated in place while the
    // topology stays the same
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t graph_instantiations = 0, graph_updates = 0;
};

RenderContext& render_context(int device) {
    static RenderContext ctx[kMaxDevices];
    return ctx[device];
}

// Pipeline p of n renders chunks p, p + n, p + 2n, ... of the call's samples.
__global__ void k_chunk_setup(ChunkState* cs, unsigned long long total, uint32_t chunk, uint32_t p, uint32_t n) {
    if (threadIdx.x == 0) {
        cs->next = (unsigned long long)p * chunk;
        cs->total = total;
        cs->first = 0;
        cs->stride = (unsigned long long)n * chunk;
        cs->chunk = chunk;
        cs->count = 0;
    }
}

// The geometry a render traverses and shades: one scene, or the packed
// forest of a batch.
struct RenderGeom {
    SceneView sv;
    SceneView exact;       // TN_SCENE_REFINED: the reference tree behind the refined sv
    bool refined = false;
    const float* normals;
    const float4* albedo;
    const float4* emission;
    const float4* trirec;
    const uint32_t* emit_ids;
    const float* emit_cdf;
    const float4* reflective = nullptr;
    const float* reflectivity = nullptr;
    int device;
    int cus;
    uint32_t* status;  // 4 words; [2] counts incomplete render traversals
};

RenderGeom geom_of(const tn_scene* sc) {
    RenderGeom g;
    g.sv = sc->refined ? sc->fview() : sc->view();
    g.exact = sc->view();
    g.refined = sc->refined;
    g.normals = sc->d_normals;
    g.albedo = sc->d_albedo;
    g.emission = sc->d_emission;
    g.trirec = sc->d_tris;
    g.emit_ids = sc->d_emit_ids;
    g.emit_cdf = sc->d_emit_cdf;
    g.reflective = sc->d_reflective;
    g.reflectivity = sc->d_reflectivity;
    g.device = sc->device;
    g.cus = sc->cus;
    g.status = sc->d_status;
    return g;
}

SceneParams scene_params(const tn_camera* cam, const tn_light* lights, uint32_t nl, uint32_t width,
                         uint32_t height, uint32_t spp, const float bg[4], uint32_t emit_begin, uint32_t nemit,
                         float emit_area) {
    SceneParams S{};
    if (cam) {
        S.cam_pos = make_float4(cam->pos[0], cam->pos[1], cam->pos[2], 0);
        std::memcpy(S.cam_m, cam->trafo, sizeof S.cam_m);
        S.cam_dx = cam->delta_x;
        S.cam_dy = cam->delta_y;
    }
    S.width = width;
    S.height = height;
    S.spp = spp;
    S.nlights = nl;
    if (nl) {
        S.light_pos = make_float4(lights[0].pos[0], lights[0].pos[1], lights[0].pos[2], 0);
        S.light_col = make_float4(lights[0].color[0], lights[0].color[1], lights[0].color[2], lights[0].color[3]);
    }
    S.bg = make_float4(bg[0], bg[1], bg[2], bg[3]);
    S.emit_begin = emit_begin;
    S.nemit = nemit;
    S.emit_area = emit_area;
    return S;
}

}  // namespace

// Launch shape of the wavefront loop: primary samples per chunk (0: from the
// queue budget), blocks per CU of the persistent trace / shade / shadow
// kernels, and whether the per-chunk launch sequence is captured once in a
// hipGraph and replayed for every chunk.
struct LaunchShape {
    uint32_t chunk = 0;
    uint32_t blocks_per_cu = 8;
    bool graph = false;
    bool persistent = false;  // one cooperative launch walks every chunk and level (k_persist)
};

int check_render_params(const tn_render_params* P) {
    if (P->max_depth < 0 || P->mc_samples < 0) return set_error(TN_ERR_INVALID, "invalid render parameters");
    if ((uint32_t)P->max_depth + 2 > kMaxLevels) return set_error(TN_ERR_LIMIT, "max_depth too large");
    return TN_OK;
}

// Chunk size for a render: the queue of a chunk's deepest level holds
// chunk * m^max_depth segments; the default keeps it near 2^23.
int chunk_size(const tn_render_params* P, uint32_t requested, uint32_t& chunk, uint32_t& cap) {
    const uint32_t m = (uint32_t)P->mc_samples;
    double fan = 1;
    for (int k = 0; k < P->max_depth; ++k) fan *= std::max(1u, m);
    chunk = requested ? requested : P->chunk_paths;
    if (chunk == 0) chunk = (uint32_t)std::max(1.0, std::min((double)(1u << 23) / fan, 1.0 * (1u << 22)));
    if (fan * chunk > 2147483647.0) return set_error(TN_ERR_LIMIT, "queue would exceed 2^31 entries");
    cap = (uint32_t)std::max(1.0, fan * chunk);
    return TN_OK;
}

// The wavefront loop over `total` samples of the scenes in `table` (sample
// and pixel ranges filled in by the caller), adding per-pixel radiance sums
// into accum.  Caller rays (ray_o/ray_d) replace the camera for tn_trace.
int render_core(const RenderGeom& G, const std::vector<SceneParams>& table, uint64_t total,
                const tn_render_params* P, const LaunchShape& shape, float* accum, tn_render_stats* st,
                hipStream_t stream, const float* ray_o, const float* ray_d, const float* jitter_host,
                size_t jitter_len, tn_guide* guide, uint32_t train_samples) {
    const bool rays = ray_o != nullptr;
    const bool guided = guide != nullptr;
    const bool training = guided && train_samples > 0;
    int rc = check_render_params(P);
    if (rc) return rc;
    if ((P->flags & TN_RENDER_GUIDED) && !guided)
        return set_error(TN_ERR_INVALID, "TN_RENDER_GUIDED needs a guide (tn_render_guided)");
    if (training && P->mc_samples != 1)
        return set_error(TN_ERR_INVALID, "training records need mc_samples == 1 (one path per sample)");
    if (guided && guide->device != G.device)
        return set_error(TN_ERR_INVALID, "guide and scene live on different devices");
    const uint32_t renderer = P->renderer;
    if (renderer > TN_RENDERER_RAYTRACER) return set_error(TN_ERR_INVALID, "unknown renderer %u", renderer);
    if (renderer != TN_RENDERER_PATHTRACER && guided)
        return set_error(TN_ERR_INVALID, "guided rendering needs the path tracer");
    if (renderer == TN_RENDERER_RAYTRACER)
        for (const SceneParams& S : table)
            if (S.nlights != 1)
                return set_error(TN_ERR_INVALID, "the raytracer needs one point light (raytracer.cpp:14)");
    // queue sizing: the raycaster traces one level, the raytracer spawns at
    // most one (mirror) child per hit
    tn_render_params q = *P;
    if (renderer == TN_RENDERER_RAYCASTER) q.max_depth = 0;
    if (renderer != TN_RENDERER_PATHTRACER) q.mc_samples = 1;
    uint32_t chunk = 0, cap = 0;
    if ((rc = chunk_size(&q, shape.chunk, chunk, cap))) return rc;
    // Two pipelines, each with its own queues and streams, render alternate
    // chunks concurrently, so that one's launch tails (the slowest rays of a
    // level) overlap the other's work.  A render of one chunk is split evenly
    // between them when each half keeps at least 2^19 paths (smaller halves
    // measured slower than one pipeline).  A graph-replayed render uses one
    // pipeline: capturing the second pipeline's stream, which itself forks
    // its shadow stream, crashed the host process, so that combination
    // stays off.
    // the persistent kernel serves plain path tracing of camera samples
    const bool persistent = shape.persistent && renderer == TN_RENDERER_PATHTRACER && !guided && !rays && !G.refined;
    int npipes = (shape.graph || persistent) ? 1 : kPipes;
    if (npipes > 1 && total <= chunk) {
        if (total >= (uint64_t)npipes << 19) {
            const uint64_t part = ((total + npipes - 1) / npipes + 255) & ~255ull;
            if ((rc = chunk_size(&q, (uint32_t)part, chunk, cap))) return rc;
        } else {
            npipes = 1;
        }
    }
    if (st) *st = tn_render_stats{};
    if (total == 0) return TN_OK;

    if (G.device < 0 || G.device >= kMaxDevices) return set_error(TN_ERR_LIMIT, "device %d", G.device);
    RenderContext& C = render_context(G.device);
    std::lock_guard<std::mutex> lock(C.mu);
    RenderBuffers* rbs = C.rbs;
    hipStream_t* pipe_stream = C.pipe_stream;
    // Every exit after the first enqueued operation drains the streams this
    // call used, so that no later call (once the lock is released) reuses
    // queues or records that kernels of this one may still read.
    bool enqueued = false;
    bool deferred = false;  // a deferred training set trains on the guide's stream during this call
    auto fail = [&](int code) {
        if (deferred) (void)hipStreamSynchronize(guide->tstream);
        if (enqueued) {
            (void)hipStreamSynchronize(stream);
            for (int p = 0; p < kPipes; ++p) {
                if (pipe_stream[p]) (void)hipStreamSynchronize(pipe_stream[p]);
                if (rbs[p].side) (void)hipStreamSynchronize(rbs[p].side);
            }
        }
        return code;
    };
#define TN_RHIP(call)                                                    \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) return fail(::tn::hip_error(e_, #call));   \
    } while (0)
    for (int p = 0; p < npipes; ++p) {
        if ((rc = alloc_queues(rbs[p], cap, (uint32_t)table.size()))) return rc;
        if ((rc = ensure_side_stream(rbs[p]))) return rc;
    }
    if (!C.ev_fork) {
        TN_HIP(hipEventCreateWithFlags(&C.ev_fork, hipEventDisableTiming));
        for (int p = 1; p < kPipes; ++p) {
            TN_HIP(hipStreamCreateWithFlags(&pipe_stream[p], hipStreamNonBlocking));
            TN_HIP(hipEventCreateWithFlags(&C.ev_join[p], hipEventDisableTiming));
        }
    }
    if (jitter_host && C.jitter_cap < jitter_len) {
        if (C.jitter) (void)hipFree(C.jitter);
        C.jitter = nullptr;
        C.jitter_cap = 0;
        TN_HIP(hipMalloc(&C.jitter, jitter_len * sizeof(float)));
        C.jitter_cap = jitter_len;
    }
    for (int p = 0; p < npipes && training; ++p)
        if ((rc = alloc_records(C.recs[p], (size_t)chunk * ((uint32_t)P->max_depth + 1)))) return rc;
    enqueued = true;
    TN_RHIP(hipMemcpyAsync(rbs[0].sp, table.data(), table.size() * sizeof(SceneParams), hipMemcpyHostToDevice,
                           stream));

    RenderArgs A0{};
    A0.sv = G.sv;
    A0.sv_exact = G.exact;
    A0.refined = G.refined;
    A0.normals = G.normals;
    A0.albedo = G.albedo;
    A0.emission = G.emission;
    A0.trirec = G.trirec;
    A0.emit_ids = G.emit_ids;
    A0.emit_cdf = G.emit_cdf;
    A0.sp = rbs[0].sp;
    A0.nscenes = (uint32_t)table.size();
    A0.status = G.status;
    A0.max_depth = q.max_depth;
    A0.mc = P->mc_samples;
    A0.max_vis = P->max_visibility;
    A0.shadow_si = P->shadow_intensity;
    A0.reflective = G.reflective;
    A0.reflectivity = G.reflectivity;
    A0.flags = P->flags;
    A0.seed = P->seed;
    A0.accum = (float4*)accum;
    A0.seed_pass = P->seed + (uint64_t)P->sample_offset * 0xD1B54A32D192ED03ull;
    // deferred training: the records of the last render train on the
    // guide's stream while this one renders from a snapshot of the weights
    const bool defer = guided && (P->flags & TN_RENDER_DEFER_TRAINING);
    if (guided && guide->pending) {
        if (defer) rc = guide_defer_begin(guide, stream, training, &deferred);
        else rc = guide_train_pending(guide, stream);
        if (rc) return fail(rc);
    }
    if (guided) {
        A0.gv = guide->view();
        if (deferred) {
            A0.gv.wbf = guide->wbf_snap;
            A0.gv.theta = guide->theta_snap;
        }
        A0.mis = guide->mis_guided;
    }
    if (training) {
        A0.rstride = (uint32_t)P->max_depth + 1;
        // every bounce of every path is a candidate; keep about kTrainSet of them
        constexpr uint32_t kTrainSet = 1u << 21;
        if ((rc = guide_reserve_set(guide, kTrainSet, stream))) return fail(rc);
        const double candidates = (double)total * std::max(1, P->max_depth);
        A0.accept = (float)std::min(1.0, 0.95 * kTrainSet / std::max(1.0, candidates));
        A0.ts = guide->ts;
        A0.ts_count = guide->ts_count;
        A0.ts_cap = guide->ts_cap;
    }
    if (jitter_host) {
        TN_RHIP(hipMemcpyAsync(C.jitter, jitter_host, jitter_len * sizeof(float), hipMemcpyHostToDevice, stream));
        A0.jitter = C.jitter;
    }
    RecordBuffers* recs = C.recs;
    RenderArgs A[kPipes];
    Queues Q[kPipes];
    for (int p = 0; p < npipes; ++p) {
        A[p] = A0;
        A[p].cs = rbs[p].cs;
        Q[p] = rbs[p].q;
        if (training) {
            const size_t ns = recs[p].slots;
            A[p].rec_p = recs[p].base;
            A[p].rec_n = recs[p].base + ns;
            A[p].rec_w = recs[p].base + 2 * ns;
            A[p].rec_g = recs[p].base + 3 * ns;
            A[p].rec_x = recs[p].base + 4 * ns;
        }
        k_zero_totals<<<1, 64, 0, stream>>>(Q[p].totals, Q[p].cnt, p == 0 ? G.status + 2 : nullptr);
        k_chunk_setup<<<1, 64, 0, stream>>>(rbs[p].cs, total, chunk, (uint32_t)p, (uint32_t)npipes);
    }
    const bool exh = P->flags & TN_RENDER_EXHAUSTIVE;
    const int levels = q.max_depth + 1;
    const uint32_t bpc = std::max(1u, shape.blocks_per_cu);
    const dim3 tgrid((uint32_t)G.cus * bpc), tblock(kRBlock);
    const uint32_t rgrid = (chunk + kRBlock - 1) / kRBlock;
    // One chunk of pipeline p: raygen, then per level trace and shade on the
    // main stream s and the level's shadow test on the pipeline's side
    // stream, so that the shadow test of level k overlaps the tracing of
    // level k + 1.  Shade of level k writes shadow queue k & 1, so it waits
    // for the shadow test of level k - 2; the raytracer's shade also reads
    // the child weights the shadow test of level k - 1 darkens, so it waits
    // for that one.
    auto chunk_sequence = [&](int p, hipStream_t s) -> hipError_t {
        hipError_t e = hipSuccess;
        auto chk = [&](hipError_t x) {
            if (e == hipSuccess) e = x;
        };
        RenderBuffers& rb = rbs[p];
        const RenderArgs& Ap = A[p];
        const Queues& Qp = Q[p];
        const hipStream_t side = rb.side;
        k_chunk_begin<<<1, 4 * kMaxLevels + 64, 0, s>>>(Qp.cnt, Qp.xwork, rb.cs);
        if (training) k_rec_clear<<<(uint32_t)(((size_t)chunk * Ap.rstride + 255) / 256), 256, 0, s>>>(Ap);
        if (rays) k_raygen_rays<<<rgrid, kRBlock, 0, s>>>(Ap, Qp, ray_o, ray_d);
        else k_raygen<<<rgrid, kRBlock, 0, s>>>(Ap, Qp);
        const bool shadows = renderer != TN_RENDERER_RAYCASTER;
        int buf = 0;
        const bool refd = G.refined;
        const dim3 fgrid(32u);  // fix lists are short (a scan after an overflow is rare)
        for (int lv = 0; lv < levels; ++lv) {
            if (refd) {
                if (exh) k_trace<true, true><<<tgrid, tblock, 0, s>>>(Ap, Qp, lv, buf);
                else k_trace<false, true><<<tgrid, tblock, 0, s>>>(Ap, Qp, lv, buf);
                if (exh) k_fix_hits<true><<<fgrid, kRBlock, 0, s>>>(Ap, Qp, lv, buf);
                else k_fix_hits<false><<<fgrid, kRBlock, 0, s>>>(Ap, Qp, lv, buf);
            } else {
                if (exh) k_trace<true, false><<<tgrid, tblock, 0, s>>>(Ap, Qp, lv, buf);
                else k_trace<false, false><<<tgrid, tblock, 0, s>>>(Ap, Qp, lv, buf);
            }
            if (renderer == TN_RENDERER_RAYCASTER) {
                k_cast<<<tgrid, tblock, 0, s>>>(Ap, Qp, lv, buf);
                buf ^= 1;
                continue;
            }
            const int dep = renderer == TN_RENDERER_RAYTRACER ? lv - 1 : lv - 2;
            if (dep >= 0) chk(hipStreamWaitEvent(s, rb.ev_shadow[dep], 0));
            if (renderer == TN_RENDERER_RAYTRACER) k_rt_shade<<<tgrid, tblock, 0, s>>>(Ap, Qp, lv, buf);
            else if (guided) k_shade<true><<<tgrid.x * (kRBlock / 64), 64, 0, s>>>(Ap, Qp, lv, buf);
            else k_shade_plain<<<tgrid, tblock, 0, s>>>(Ap, Qp, lv, buf);
            chk(hipEventRecord(rb.ev_shade[lv], s));
            chk(hipStreamWaitEvent(side, rb.ev_shade[lv], 0));
            if (renderer == TN_RENDERER_RAYTRACER) {
                if (exh) k_rt_shadow<true><<<tgrid, tblock, 0, side>>>(Ap, Qp, lv, buf);
                else k_rt_shadow<false><<<tgrid, tblock, 0, side>>>(Ap, Qp, lv, buf);
            } else if (refd) {
                if (exh) k_shadow<true, true><<<tgrid, tblock, 0, side>>>(Ap, Qp, lv);
                else k_shadow<false, true><<<tgrid, tblock, 0, side>>>(Ap, Qp, lv);
                if (exh) k_fix_shadows<true><<<fgrid, kRBlock, 0, side>>>(Ap, Qp, lv);
                else k_fix_shadows<false><<<fgrid, kRBlock, 0, side>>>(Ap, Qp, lv);
            } else {
                if (exh) k_shadow<true, false><<<tgrid, tblock, 0, side>>>(Ap, Qp, lv);
                else k_shadow<false, false><<<tgrid, tblock, 0, side>>>(Ap, Qp, lv);
            }
            chk(hipEventRecord(rb.ev_shadow[lv], side));
            buf ^= 1;
        }
        // every shadow test is done before the counts are read (the side
        // stream runs in order, so its last event covers all of them)
        if (shadows) chk(hipStreamWaitEvent(s, rb.ev_shadow[levels - 1], 0));
        // the raytracer also counts the mirror rays of its last level
        // (trace() is entered and returns {} beyond max_depth)
        k_chunk_totals<<<1, 64, 0, s>>>(Qp.cnt, Qp.totals,
                                          renderer == TN_RENDERER_RAYCASTER   ? -1
                                          : renderer == TN_RENDERER_RAYTRACER ? levels + 1
                                                                              : levels);
        if (training) k_collect<<<rgrid, kRBlock, 0, s>>>(Ap, P->max_depth);
        return e;
    };
    // One round: pipeline 0 on s0, pipeline 1 forked onto its own stream and
    // joined back (the fork and join are events, so a captured round is a
    // graph with the same concurrency).
    auto round = [&](hipStream_t s0) -> hipError_t {
        if (npipes == 1) return chunk_sequence(0, s0);
        hipError_t e = hipEventRecord(C.ev_fork, s0);
        for (int p = 1; p < npipes && e == hipSuccess; ++p) e = hipStreamWaitEvent(pipe_stream[p], C.ev_fork, 0);
        if (e == hipSuccess) e = chunk_sequence(0, s0);
        for (int p = 1; p < npipes && e == hipSuccess; ++p) e = chunk_sequence(p, pipe_stream[p]);
        for (int p = 1; p < npipes && e == hipSuccess; ++p) {
            e = hipEventRecord(C.ev_join[p], pipe_stream[p]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s0, C.ev_join[p], 0);
        }
        return e;
    };
    const uint64_t nchunks = (total + chunk - 1) / chunk;
    const uint64_t rounds = (nchunks + npipes - 1) / npipes;
    auto t0 = std::chrono::steady_clock::now();
    uint32_t barrier_out = 0;
    if (persistent) {
        // every block resident at once: the occupancy the kernel allows, at
        // most the shape's blocks per CU
        const void* kfn = exh ? (const void*)k_persist<true> : (const void*)k_persist<false>;
        int per_cu = 0;
        TN_RHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kRBlock, 0));
        if (per_cu < 1) return fail(set_error(TN_ERR_LIMIT, "the persistent kernel does not fit a CU"));
        const dim3 pgrid((uint32_t)G.cus * std::min<uint32_t>((uint32_t)per_cu, bpc));
        k_zero_bar<<<1, 256, 0, stream>>>(Q[0].bar);
        TN_RHIP(hipGetLastError());
        RenderArgs pa = A[0];
        Queues pq = Q[0];
        int plevels = levels;
        void* args[] = {&pa, &pq, &plevels};
        TN_RHIP(hipLaunchCooperativeKernel(kfn, pgrid, dim3(kRBlock), args, 0, stream));
        TN_RHIP(hipMemcpyAsync(&barrier_out, Q[0].bar + kBarOut, 4, hipMemcpyDeviceToHost, stream));
    } else if (shape.graph && rounds > 1) {
        // The chunk sequence is captured on every call (host-side recording;
        // its kernel arguments change from call to call) and applied to the
        // executable graph kept from the last call by hipGraphExecUpdate,
        // which needs the same topology; only a changed topology (levels,
        // renderer, guided or training work) instantiates a new one.
        if (!C.cap_stream) TN_RHIP(hipStreamCreateWithFlags(&C.cap_stream, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(C.cap_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            const hipError_t es = round(C.cap_stream);  // the other streams join and leave the capture
            e = hipStreamEndCapture(C.cap_stream, &graph);
            if (e == hipSuccess) e = es;
        }
        if (e == hipSuccess && C.exec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult why = hipGraphExecUpdateSuccess;
            if (hipGraphExecUpdate(C.exec, graph, &bad, &why) != hipSuccess || why != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(C.exec);
                C.exec = nullptr;
            } else {
                ++C.graph_updates;
            }
        }
        if (e == hipSuccess && !C.exec) {
            e = hipGraphInstantiate(&C.exec, graph, nullptr, nullptr, 0);
            if (e == hipSuccess) ++C.graph_instantiations;
            else C.exec = nullptr;
        }
        for (uint64_t c = 0; e == hipSuccess && c < rounds; ++c) e = hipGraphLaunch(C.exec, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(hip_error(e, "captured chunk sequence"));
    } else {
        for (uint64_t c = 0; c < rounds; ++c) {
            const hipError_t e = round(stream);
            if (e != hipSuccess) return fail(hip_error(e, "chunk sequence"));
        }
    }
    TN_RHIP(hipGetLastError());
    if (training) {
        if (defer) guide->pending = train_samples;  // trains during the next deferred call (or tn_guide_flush)
        else if ((rc = guide_train_from_set(guide, train_samples, stream))) return fail(rc);
    }
    if (deferred) TN_RHIP(hipStreamWaitEvent(stream, guide->ev_join, 0));
    unsigned long long tot[kPipes][2] = {};
    uint32_t overflow[kPipes] = {}, incomplete = 0;
    for (int p = 0; p < npipes; ++p) {
        TN_RHIP(hipMemcpyAsync(tot[p], Q[p].totals, 16, hipMemcpyDeviceToHost, stream));
        TN_RHIP(hipMemcpyAsync(&overflow[p], Q[p].cnt + 4 * kMaxLevels, 4, hipMemcpyDeviceToHost, stream));
    }
    TN_RHIP(hipMemcpyAsync(&incomplete, G.status + 2, 4, hipMemcpyDeviceToHost, stream));
    TN_RHIP(hipStreamSynchronize(stream));
#undef TN_RHIP
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int p = 0; p < npipes; ++p)
        if (overflow[p]) return set_error(TN_ERR_QUEUE, "a ray queue overflowed");
    if (barrier_out) return set_error(TN_ERR_INCOMPLETE, "a grid barrier wait of the persistent kernel ran out");
    if (incomplete) return set_error(TN_ERR_INCOMPLETE, "%u render traversals hit the visit cap", incomplete);
    if (st) {
        st->rays = st->shadow_rays = 0;
        for (int p = 0; p < npipes; ++p) {
            st->rays += tot[p][0];
            st->shadow_rays += tot[p][1];
        }
        st->primary_rays = rays ? 0 : total;
        st->ms = ms;
    }
    return TN_OK;
}

// Renders params->width x height x spp samples of one scene, adding the
// per-pixel SUM of sample radiance into accum (device, RGBA).
int render_into(tn_scene* sc, const tn_camera* cam, const tn_light* lights, uint32_t nl,
                const tn_render_params* P, float* accum, tn_render_stats* st, hipStream_t stream,
                const float* ray_o = nullptr, const float* ray_d = nullptr, uint64_t nrays = 0,
                tn_guide* guide = nullptr, uint32_t train_samples = 0) {
    const bool rays = ray_o != nullptr;
    if ((!rays && (P->width == 0 || P->height == 0 || P->pixel_samples < 1)) || P->max_depth < 0 ||
        P->mc_samples < 0)
        return set_error(TN_ERR_INVALID, "invalid render parameters");
    if (nl > 1) return set_error(TN_ERR_LIMIT, "at most one point light (main.cpp:123)");
    if (!rays && (size_t)P->width * P->height >= (1ull << 32))
        return set_error(TN_ERR_LIMIT, "image too large");
    if (rays && nrays >= (1ull << 32)) return set_error(TN_ERR_LIMIT, "too many rays");
    const uint32_t spp = rays ? 1u : (uint32_t)P->pixel_samples;
    std::vector<SceneParams> table(1, scene_params(cam, lights, nl, rays ? 1u : P->width, rays ? 1u : P->height,
                                                   spp, P->bg, 0, sc->nemit, sc->emit_area));
    const uint64_t total = rays ? nrays : (uint64_t)P->width * P->height * spp;
    std::vector<float> jit;
    if (!rays && (P->flags & TN_RENDER_REF_JITTER)) jit = reference_jitter(P->width, spp);
    LaunchShape shape;
    shape.graph = (P->flags & TN_RENDER_GRAPH) != 0;
    shape.persistent = (P->flags & TN_RENDER_PERSISTENT) != 0;
    return render_core(geom_of(sc), table, total, P, shape, accum, st, stream, ray_o, ray_d,
                       jit.empty() ? nullptr : jit.data(), jit.size(), guide, train_samples);
}

// ---------------------------------------------------------------- batches
namespace {

// Node fixup for the packed forest: right-child indices move by the scene's
// node base, leaf reference offsets by its reference base.
void append_tree(const HostTree& t, uint32_t node_base, uint32_t ref_base, uint32_t tri_base,
                 std::vector<DevNode>& nodes, std::vector<uint32_t>& refs) {
    for (DevNode n : t.nodes) {
        if ((n.y & 3u) == 3u) n.x += ref_base;
        else n.y += node_base << 2;
        nodes.push_back(n);
    }
    for (uint32_t r : t.refs) refs.push_back(r + tri_base);
}

}  // namespace
}  // namespace tn

// Scenes packed into one forest: node, reference, triangle and shading arrays
// concatenated (indices rebased), one root per scene.
struct tn_batch {
    int device = 0, cus = 256;
    uint32_t nscenes = 0;
    uint2* nodes = nullptr;
    uint32_t* refs = nullptr;
    float4* tris = nullptr;
    float4* lrec = nullptr;  // records per reference (first halves, then second halves)
    size_t nrefs = 0;
    float* normals = nullptr;
    float4* albedo = nullptr;
    float4* emission = nullptr;
    uint32_t* emit_ids = nullptr;
    float* emit_cdf = nullptr;
    tn::SceneRoot* roots = nullptr;
    uint32_t* status = nullptr;
    std::vector<uint32_t> emit_begin, nemit;
    std::vector<float> emit_area;
    struct Tuned {
        int32_t depth, mc;
        uint32_t flags;
        int total_log2;
        tn::LaunchShape shape;
        tn_batch_tuning info;
    };
    std::vector<Tuned> tuned;
    tn_batch_tuning last{};

    tn::RenderGeom geom() const {
        tn::RenderGeom g;
        g.sv.nodes = nodes;
        g.sv.refs = refs;
        g.sv.tris = tris;
        g.sv.lra = lrec;
        g.sv.lrb = lrec + 2 * nrefs;
        for (int k = 0; k < 3; ++k) g.sv.bmin[k] = g.sv.bmax[k] = 0.f;
        g.sv.roots = roots;
        g.normals = normals;
        g.albedo = albedo;
        g.emission = emission;
        g.trirec = tris;
        g.emit_ids = emit_ids;
        g.emit_cdf = emit_cdf;
        g.device = device;
        g.cus = cus;
        g.status = status;
        return g;
    }
};

namespace tn {
namespace {

void free_batch(tn_batch* b) {
    for (void* p : {(void*)b->nodes, (void*)b->refs, (void*)b->tris, (void*)b->lrec, (void*)b->normals,
                    (void*)b->albedo,
                    (void*)b->emission, (void*)b->emit_ids, (void*)b->emit_cdf, (void*)b->roots,
                    (void*)b->status})
        if (p) (void)hipFree(p);
}

int batch_table(const tn_batch* b, const tn_batch_job* jobs, std::vector<SceneParams>& table, uint64_t& total,
                uint64_t& pixels) {
    table.resize(b->nscenes);
    total = 0;
    pixels = 0;
    for (uint32_t s = 0; s < b->nscenes; ++s) {
        const tn_batch_job& J = jobs[s];
        if (J.width == 0 || J.height == 0 || J.pixel_samples < 1 || J.num_lights > 1)
            return set_error(TN_ERR_INVALID, "invalid job %u", s);
        table[s] = scene_params(&J.cam, &J.light, J.num_lights, J.width, J.height, (uint32_t)J.pixel_samples,
                                J.bg, b->emit_begin[s], b->nemit[s], b->emit_area[s]);
        table[s].sample_begin = total;
        table[s].pixel_begin = (uint32_t)pixels;
        total += (uint64_t)J.width * J.height * (uint32_t)J.pixel_samples;
        pixels += (uint64_t)J.width * J.height;
        if (pixels >= (1ull << 32)) return set_error(TN_ERR_LIMIT, "batch images exceed 2^32 pixels");
    }
    return TN_OK;
}

// Times candidate launch shapes on the first samples of the batch (into a
// scratch accumulator) and returns the fastest.
int autotune(tn_batch* b, const std::vector<SceneParams>& table, uint64_t total, uint64_t pixels,
             const tn_render_params* P, hipStream_t stream, LaunchShape& best, tn_batch_tuning& info) {
    // the probe spans at least two of the largest chunks, so both pipelines
    // of a direct-launch shape are busy as in the full render
    const uint64_t probe = std::min<uint64_t>(total, 1ull << 23);
    float* scratch = nullptr;
    TN_HIP(hipMalloc(&scratch, (size_t)pixels * 16));
    const uint32_t chunks[] = {1u << 18, 1u << 20, 1u << 22};
    const uint32_t grids[] = {4, 8, 16};
    double best_rate = -1;
    info = tn_batch_tuning{};
    auto t0 = std::chrono::steady_clock::now();
    int rc = TN_OK;
    for (uint32_t c : chunks) {
        uint32_t ch = 0, cap = 0;
        if (chunk_size(P, c, ch, cap) != TN_OK || cap > (1u << 24)) continue;  // queue budget
        if (c > 2 * probe && c != chunks[0]) continue;
        for (uint32_t g : grids) {
            // graph replay (one pipeline) or direct launches (two
            // pipelines); the persistent launch is no candidate (slower than
            // both on every batch measured)
            for (const bool graph : {true, false}) {
                LaunchShape sh;
                sh.chunk = c;
                sh.blocks_per_cu = g;
                sh.graph = graph;
                tn_render_stats st{};
                rc = render_core(b->geom(), table, probe, P, sh, scratch, &st, stream, nullptr, nullptr, nullptr,
                                 0, nullptr, 0);
                if (rc) break;
                const double rate = (double)probe / std::max(st.ms, 1e-3);
                ++info.candidates;
                if (rate > best_rate) {
                    best_rate = rate;
                    best = sh;
                }
            }
            if (rc) break;
        }
        if (rc) break;
    }
    (void)hipFree(scratch);
    if (rc) return rc;
    if (best_rate < 0) return set_error(TN_ERR_LIMIT, "no launch shape fits the queue budget");
    info.chunk_paths = best.chunk;
    info.blocks_per_cu = best.blocks_per_cu;
    info.graph = best.graph ? 1u : 0u;
    info.best_rate = best_rate;
    info.probe_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return TN_OK;
}

}  // namespace
}  // namespace tn

using namespace tn;

extern "C" {

int tn_sample_triangle_host(const float* p3, const float* tri9, const float* u2, uint32_t n, float* dir3,
                            float* omega, float* omega_vos) {
    if (!p3 || !tri9 || (n && (!u2 || !dir3)) || !omega || !omega_vos)
        return set_error(TN_ERR_INVALID, "NULL argument");
    const float3 p = make_float3(p3[0], p3[1], p3[2]);
    const float3 a = v3sub(make_float3(tri9[0], tri9[1], tri9[2]), p),
                 b = v3sub(make_float3(tri9[3], tri9[4], tri9[5]), p),
                 c = v3sub(make_float3(tri9[6], tri9[7], tri9[8]), p);
    SphTri st;
    if (!sph_tri_setup(a, b, c, st)) return set_error(TN_ERR_INVALID, "degenerate spherical triangle");
    *omega = st.omega;
    *omega_vos = tri_solid_angle(a, b, c);
    for (uint32_t i = 0; i < n; ++i) {
        const float3 d = sph_tri_sample(st, u2[2 * i], u2[2 * i + 1]);
        dir3[3 * i] = d.x;
        dir3[3 * i + 1] = d.y;
        dir3[3 * i + 2] = d.z;
    }
    return TN_OK;
}

int tn_render_accumulate(tn_scene* sc, const tn_camera* cam, const tn_light* lights, uint32_t nl,
                         const tn_render_params* P, float* accum, tn_render_stats* st, void* stream) {
    if (!sc || !cam || !P || !accum || (nl && !lights))
        return set_error(TN_ERR_INVALID, "NULL argument");
    return render_into(sc, cam, lights, nl, P, accum, st, (hipStream_t)stream);
}

int tn_render_guided(tn_scene* sc, const tn_camera* cam, const tn_light* lights, uint32_t nl,
                     const tn_render_params* P, tn_guide* guide, uint32_t train_samples, float* accum,
                     tn_render_stats* st, void* stream) {
    if (!sc || !cam || !P || !guide || !accum || (nl && !lights))
        return set_error(TN_ERR_INVALID, "NULL argument");
    tn_render_params q = *P;
    q.flags |= TN_RENDER_GUIDED;
    return render_into(sc, cam, lights, nl, &q, accum, st, (hipStream_t)stream, nullptr, nullptr, 0, guide,
                       train_samples);
}

int tn_trace(tn_scene* sc, const float* o3, const float* d3, uint32_t n, int32_t depth,
             const tn_light* lights, uint32_t nl, const tn_render_params* P, float* rgba,
             tn_render_stats* st, void* stream) {
    if (!sc || !P || (n && (!o3 || !d3 || !rgba)) || (nl && !lights) || depth < 0)
        return set_error(TN_ERR_INVALID, "NULL or invalid argument");
    if (st) *st = tn_render_stats{};
    if (n == 0) return TN_OK;
    tn_render_params q = *P;
    if (P->renderer == TN_RENDERER_RAYCASTER) {
        q.max_depth = 0;  // raycaster.cpp:7-24 ignores the depth
    } else if (depth > P->max_depth) {  // pathtracer.cpp:17-19, raytracer.cpp:9-13
        if (st && P->renderer == TN_RENDERER_RAYTRACER) st->rays = n;  // counted before the return
        return TN_OK;
    } else {
        q.max_depth = P->max_depth - depth;  // levels depth..max_depth remain
    }
    return render_into(sc, nullptr, lights, nl, &q, rgba, st, (hipStream_t)stream, o3, d3, n);
}

int tn_render(tn_scene* sc, const tn_camera* cam, const tn_light* lights, uint32_t nl,
              const tn_render_params* P, float* linear, float* display, tn_render_stats* st) {
    if (!sc || !cam || !P || (nl && !lights)) return set_error(TN_ERR_INVALID, "NULL argument");
    if (P->width == 0 || P->height == 0 || P->pixel_samples < 1)
        return set_error(TN_ERR_INVALID, "invalid render parameters");
    const size_t np = (size_t)P->width * P->height;
    TN_HIP(hipSetDevice(sc->device));
    std::lock_guard<std::mutex> lock(sc->stage_mu);
    if (int rc = sc->reserve_stage(np * 16 * 2)) return rc;
    float* acc = (float*)sc->d_stage;
    float* out = acc + np * 4;
    int rc = TN_OK;
    hipError_t e = hipMemsetAsync(acc, 0, np * 16, nullptr);
    if (e != hipSuccess) rc = hip_error(e, "hipMemsetAsync");
    if (!rc) rc = render_into(sc, cam, lights, nl, P, acc, st, nullptr);
    const float inv = 1.0f / (float)P->pixel_samples;
    if (!rc && linear) {  // mean over pixel samples (main.cpp:216)
        std::vector<float> tmp(np * 4);
        if (!rc && (e = hipMemcpy(tmp.data(), acc, np * 16, hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_error(e, "hipMemcpy");
        if (!rc)
            for (size_t k = 0; k < np * 4; ++k) linear[k] = tmp[k] * inv;
    }
    if (!rc && display) {
        rc = tn_tonemap(acc, out, (uint32_t)np, inv, P->exposure, P->inverse_gamma, P->gamma_enabled,
                        nullptr);
        if (!rc && (e = hipMemcpy(display, out, np * 16, hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_error(e, "hipMemcpy");
    }
    if (rc) (void)hipDeviceSynchronize();  // no kernel of this call still uses the staging memory
    return rc;
}

int tn_batch_create(tn_scene* const* scenes, uint32_t n, tn_batch** out) {
    if (!scenes || !out || n == 0) return set_error(TN_ERR_INVALID, "NULL argument or empty batch");
    *out = nullptr;
    for (uint32_t s = 0; s < n; ++s)
        if (!scenes[s]) return set_error(TN_ERR_INVALID, "NULL scene %u", s);
    const int dev = scenes[0]->device;
    for (uint32_t s = 0; s < n; ++s)
        if (scenes[s]->device != dev) return set_error(TN_ERR_INVALID, "scenes live on different devices");
    TN_HIP(hipSetDevice(dev));
    std::vector<DevNode> nodes;
    std::vector<uint32_t> refs, eids;
    std::vector<float> ecdf;
    std::vector<SceneRoot> roots(n);
    tn_batch* b = new tn_batch();
    b->device = dev;
    b->cus = scenes[0]->cus;
    b->nscenes = n;
    uint64_t ntri = 0;
    std::vector<uint64_t> tri_base(n);
    for (uint32_t s = 0; s < n; ++s) {
        const tn_scene* sc = scenes[s];
        tri_base[s] = ntri;
        if (ntri + sc->ntris >= TN_INVALID_ID || nodes.size() + sc->tree.nodes.size() >= (1ull << 30) ||
            refs.size() + sc->tree.refs.size() >= (1ull << 32)) {
            delete b;
            return set_error(TN_ERR_LIMIT, "batch exceeds 2^30 triangles or nodes");
        }
        roots[s].node = (uint32_t)nodes.size();
        for (int k = 0; k < 3; ++k) {
            roots[s].bmin[k] = sc->tree.box[k];
            roots[s].bmax[k] = sc->tree.box[3 + k];
        }
        append_tree(sc->tree, (uint32_t)nodes.size(), (uint32_t)refs.size(), (uint32_t)ntri, nodes, refs);
        ntri += sc->ntris;
        b->emit_begin.push_back((uint32_t)eids.size());
        b->nemit.push_back(sc->nemit);
        b->emit_area.push_back(sc->emit_area);
        if (sc->nemit) {
            std::vector<uint32_t> ids(sc->nemit);
            std::vector<float> cdf(sc->nemit);
            hipError_t e = hipMemcpy(ids.data(), sc->d_emit_ids, sc->nemit * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(cdf.data(), sc->d_emit_cdf, sc->nemit * 4, hipMemcpyDeviceToHost);
            if (e != hipSuccess) {
                delete b;
                return hip_error(e, "hipMemcpy");
            }
            for (uint32_t& i : ids) i += (uint32_t)tri_base[s];
            eids.insert(eids.end(), ids.begin(), ids.end());
            ecdf.insert(ecdf.end(), cdf.begin(), cdf.end());
        }
    }
    auto fail = [&](hipError_t e, const char* what) {
        free_batch(b);
        delete b;
        return hip_error(e, what);
    };
    hipError_t e;
    const size_t nt = std::max<uint64_t>(ntri, 1);
    if ((e = hipMalloc(&b->nodes, std::max<size_t>(nodes.size(), 1) * 8)) != hipSuccess ||
        (e = hipMalloc(&b->refs, std::max<size_t>(refs.size(), 1) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->tris, nt * 64)) != hipSuccess || (e = hipMalloc(&b->normals, nt * 36)) != hipSuccess ||
        (e = hipMalloc(&b->albedo, nt * 16)) != hipSuccess || (e = hipMalloc(&b->emission, nt * 16)) != hipSuccess ||
        (e = hipMalloc(&b->emit_ids, std::max<size_t>(eids.size(), 1) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->emit_cdf, std::max<size_t>(ecdf.size(), 1) * 4)) != hipSuccess ||
        (e = hipMalloc(&b->roots, n * sizeof(SceneRoot))) != hipSuccess ||
        (e = hipMalloc(&b->status, 4 * sizeof(uint32_t))) != hipSuccess)
        return fail(e, "hipMalloc");
    if ((e = hipMemset(b->status, 0, 4 * sizeof(uint32_t))) != hipSuccess) return fail(e, "hipMemset");
    if ((e = hipMemcpy(b->nodes, nodes.data(), nodes.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->refs, refs.data(), refs.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->roots, roots.data(), n * sizeof(SceneRoot), hipMemcpyHostToDevice)) != hipSuccess ||
        (!eids.empty() &&
         (e = hipMemcpy(b->emit_ids, eids.data(), eids.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) ||
        (!ecdf.empty() &&
         (e = hipMemcpy(b->emit_cdf, ecdf.data(), ecdf.size() * 4, hipMemcpyHostToDevice)) != hipSuccess))
        return fail(e, "hipMemcpy");
    for (uint32_t s = 0; s < n; ++s) {
        const tn_scene* sc = scenes[s];
        const size_t t0 = tri_base[s], k = sc->ntris;
        if ((e = hipMemcpy(b->tris + 4 * t0, sc->d_tris, k * 64, hipMemcpyDeviceToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->normals + 9 * t0, sc->d_normals, k * 36, hipMemcpyDeviceToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->albedo + t0, sc->d_albedo, k * 16, hipMemcpyDeviceToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->emission + t0, sc->d_emission, k * 16, hipMemcpyDeviceToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy");
    }
    b->nrefs = refs.size();
    if ((e = hipMalloc(&b->lrec, std::max<size_t>(b->nrefs, 1) * 64)) != hipSuccess) return fail(e, "hipMalloc");
    if (int rc = gather_leaf_records(b->tris, b->refs, b->nrefs, b->lrec)) {
        free_batch(b);
        delete b;
        return rc;
    }
    *out = b;
    return TN_OK;
}

int tn_batch_destroy(tn_batch* b) {
    if (!b) return TN_OK;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    free_batch(b);
    delete b;
    return TN_OK;
}

int tn_batch_render(tn_batch* b, const tn_batch_job* jobs, const tn_render_params* P, float* accum,
                    tn_render_stats* st, void* stream) {
    if (!b || !jobs || !P || !accum) return set_error(TN_ERR_INVALID, "NULL argument");
    if (P->flags & (TN_RENDER_GUIDED | TN_RENDER_REF_JITTER))
        return set_error(TN_ERR_INVALID, "batches render without guide or reference jitter");
    if (P->renderer != TN_RENDERER_PATHTRACER) return set_error(TN_ERR_INVALID, "batches render with the path tracer");
    int rc = check_render_params(P);
    if (rc) return rc;
    TN_HIP(hipSetDevice(b->device));
    std::vector<SceneParams> table;
    uint64_t total = 0, pixels = 0;
    if ((rc = batch_table(b, jobs, table, total, pixels))) return rc;
    hipStream_t s = (hipStream_t)stream;
    LaunchShape shape;
    shape.graph = (P->flags & TN_RENDER_GRAPH) != 0;
    shape.persistent = (P->flags & TN_RENDER_PERSISTENT) != 0;
    if ((P->flags & TN_RENDER_AUTOTUNE) && P->chunk_paths == 0) {
        const uint32_t key_flags = P->flags & ~(TN_RENDER_GRAPH | TN_RENDER_AUTOTUNE | TN_RENDER_PERSISTENT);
        int lg = 0;
        while ((1ull << lg) < total) ++lg;
        const tn_batch::Tuned* hit = nullptr;
        for (const auto& t : b->tuned)
            if (t.depth == P->max_depth && t.mc == P->mc_samples && t.flags == key_flags && t.total_log2 == lg)
                hit = &t;
        if (hit) {
            shape = hit->shape;
            b->last = hit->info;
        } else {
            tn_batch_tuning info{};
            if ((rc = autotune(b, table, total, pixels, P, s, shape, info))) return rc;
            b->tuned.push_back({P->max_depth, P->mc_samples, key_flags, lg, shape, info});
            b->last = info;
        }
    }
    return render_core(b->geom(), table, total, P, shape, accum, st, s, nullptr, nullptr, nullptr, 0, nullptr, 0);
}

int tn_batch_get_tuning(const tn_batch* b, tn_batch_tuning* out) {
    if (!b || !out) return set_error(TN_ERR_INVALID, "NULL argument");
    *out = b->last;
    return TN_OK;
}

}  // extern "C"
