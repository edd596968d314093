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
#include <vector>

#include "tn_guide.h"
#include "tn_scene.h"

namespace tn {
namespace {

constexpr int kRBlock = 256;
constexpr uint32_t kMaxLevels = 64;

struct Queues {
    float4* o[2];   // o.xyz, pixel (bits)
    float4* d[2];   // d.xyz, depth (bits)
    float4* w[2];   // throughput weight (RGBA)
    uint4* rng[2];  // xorshift128 state
    uint4* hit;     // id, r, s, t bits
    float4* so;     // shadow: origin, pixel (bits)
    float4* sd;     // shadow: direction, tmax
    float4* sc;     // shadow: contribution
    float4* sx;     // shadow: radiance at the segment (training records), slot (bits)
    uint32_t cap;   // capacity of each path queue (entries)
    uint32_t scap;  // capacity of the shadow queue
    // device counters: [0..kMaxLevels) path counts, [kMaxLevels..2k) shadow
    // counts, [2k..3k) trace work, [3k..4k) shadow work, [4k] overflow
    uint32_t* cnt;
    unsigned long long* totals;  // [0] segments, [1] shadow rays
};

struct RenderArgs {
    SceneView sv;
    const float* normals;  // 9 per triangle
    const float4* albedo;
    const float4* emission;
    const float4* trirec;  // device triangle records (geometric normal in rec[1])
    const uint32_t* emit_ids;
    const float* emit_cdf;
    uint32_t nemit;
    float emit_area;
    float4 cam_pos;
    float cam_m[9];
    float cam_dx, cam_dy;
    uint32_t width, height, spp;
    int32_t max_depth, mc;
    uint32_t flags;
    uint64_t seed;
    float4 bg;
    uint32_t nlights;
    float4 light_pos, light_col;
    const float* jitter;  // 2 * width * spp (REF_JITTER) or null
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
};

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

__global__ void k_chunk_init(uint32_t* cnt, uint32_t n0) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * kMaxLevels + 1) cnt[i] = (i == 0) ? n0 : 0u;
}

__global__ void k_zero_totals(unsigned long long* totals) {
    if (threadIdx.x < 2) totals[threadIdx.x] = 0;
}

__global__ void k_chunk_totals(const uint32_t* cnt, unsigned long long* totals, int levels) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int k = 0; k < levels; ++k) {
            a += cnt[k];
            b += cnt[kMaxLevels + k];
        }
        totals[0] += a;
        totals[1] += b;
    }
}

// Primary samples: j = (y * width + x) * spp + i  (main.cpp:193-215 order).
__global__ __launch_bounds__(kRBlock) void k_raygen(RenderArgs A, Queues Q, uint64_t first,
                                                    uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t j = first + i;
    const uint32_t sample = (uint32_t)(j % A.spp);
    const uint64_t pix = j / A.spp;
    const uint32_t x = (uint32_t)(pix % A.width), y = (uint32_t)(pix / A.width);
    uint4 rng = rng_seed(A.seed_pass, j);
    float dx, dy;
    if (A.jitter) {
        const size_t k = 2 * ((size_t)x * A.spp + sample);
        dx = A.jitter[k];
        dy = A.jitter[k + 1];
    } else {
        dx = rng_uniform(rng);
        dy = rng_uniform(rng);
    }
    // lib/types.h:119-123
    const float px = (float)x + dx, py = (float)y + dy;
    const float w = (float)A.width, h = (float)A.height;
    const float cx = -A.cam_dx * (1 - 2 * px / w);
    const float cy = A.cam_dy * (1 - 2 * py / h);
    const float cz = -1;
    const float* m = A.cam_m;
    const float3 d = make_float3(m[0] * cx + m[1] * cy + m[2] * cz, m[3] * cx + m[4] * cy + m[5] * cz,
                                 m[6] * cx + m[7] * cy + m[8] * cz);
    Q.o[0][i] = make_float4(A.cam_pos.x, A.cam_pos.y, A.cam_pos.z, __uint_as_float((uint32_t)pix));
    Q.d[0][i] = make_float4(d.x, d.y, d.z, __uint_as_float(i));  // .w: path within the chunk
    Q.w[0][i] = make_float4(1.f, 1.f, 1.f, 1.f);
    Q.rng[0][i] = rng;
}

// Caller rays (tn_trace): ray i accumulates into accum[i]; seeds from (seed, i).
__global__ __launch_bounds__(kRBlock) void k_raygen_rays(RenderArgs A, Queues Q, const float* o3,
                                                         const float* d3, uint64_t first,
                                                         uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t j = first + i;
    Q.o[0][i] = make_float4(o3[3 * j], o3[3 * j + 1], o3[3 * j + 2], __uint_as_float((uint32_t)j));
    Q.d[0][i] = make_float4(d3[3 * j], d3[3 * j + 1], d3[3 * j + 2], __uint_as_float(i));
    Q.w[0][i] = make_float4(1.f, 1.f, 1.f, 1.f);
    Q.rng[0][i] = rng_seed(A.seed_pass, j);
}

// Per-wave dynamic chunks of a device-counted queue (one atomic per chunk).
// The chunk shrinks with the queue (rays_per_wave over a quarter of the
// grid's waves, so a wave takes about four chunks) to keep every wave busy
// on short queues.
struct DynamicChunks {
    uint32_t n;
    uint32_t* work;
    uint32_t size;
    __device__ DynamicChunks(uint32_t n_, uint32_t* work_) : n(n_), work(work_) {
        const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
        uint32_t c = (n + 4 * waves - 1) / (4 * waves);
        c = (c + 63u) & ~63u;
        size = c < 64u ? 64u : (c > (uint32_t)kChunk ? (uint32_t)kChunk : c);
    }
    __device__ bool operator()(uint32_t& b, uint32_t& e) const {
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == 0) base = atomicAdd(work, size);
        base = __shfl(base, 0, 64);
        if (base >= n) return false;
        b = base;
        e = min(base + size, n);
        return true;
    }
};

template <bool EXH>
__global__ __launch_bounds__(kRBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_trace(
    RenderArgs A, Queues Q, int level, int buf) {
    __shared__ uint2 stk[kStack * kRBlock];
    __shared__ unsigned long long slots[kRBlock];
    const uint32_t n = min(Q.cnt[level], Q.cap);  // overflow is reported, never read
    const float4* qo = Q.o[buf];
    const float4* qd = Q.d[buf];
    uint4* hit = Q.hit;
    TraversalCounters cnt;
    coop_loop<kStack, EXH, false, false, false>(
        A.sv, stk + threadIdx.x, kRBlock, slots + (threadIdx.x & ~63u),
        DynamicChunks{n, Q.cnt + 2 * kMaxLevels + level},
        [&](uint32_t k, float3& o, float3& d, float&) {
            const float4 a = qo[k], b = qd[k];
            o = make_float3(a.x, a.y, a.z);
            d = make_float3(b.x, b.y, b.z);
        },
        [&](uint32_t k, const Trav& T, bool traversed) {
            const bool h = traversed && T.best_id != kInvalidId;
            hit[k] = make_uint4(h ? T.best_id : kInvalidId, __float_as_uint(h ? T.best_r : FLT_MAX),
                                __float_as_uint(h ? T.best_s : 0.f), __float_as_uint(h ? T.best_t : 0.f));
        },
        cnt);
}

template <bool EXH>
__global__ __launch_bounds__(kRBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_shadow(
    RenderArgs A, Queues Q, int level) {
    __shared__ uint2 stk[kStack * kRBlock];
    __shared__ unsigned long long slots[kRBlock];
    const uint32_t n = min(Q.cnt[kMaxLevels + level], Q.scap);
    TraversalCounters cnt;
    coop_loop<kStack, EXH, false, true, false>(
        A.sv, stk + threadIdx.x, kRBlock, slots + (threadIdx.x & ~63u),
        DynamicChunks{n, Q.cnt + 3 * kMaxLevels + level},
        [&](uint32_t k, float3& o, float3& d, float& tmax) {
            const float4 a = Q.so[k], b = Q.sd[k];
            o = make_float3(a.x, a.y, a.z);
            d = make_float3(b.x, b.y, b.z);
            tmax = b.w;
        },
        [&](uint32_t k, const Trav& T, bool traversed) {
            if (!traversed || T.best_id == kInvalidId) {
                accumulate(A.accum, __float_as_uint(Q.so[k].w), Q.sc[k]);
                if (A.rec_x) {
                    const float4 x = Q.sx[k];
                    float* dst = reinterpret_cast<float*>(A.rec_x + __float_as_uint(x.w));
                    atomicAdd(dst + 0, x.x);
                    atomicAdd(dst + 1, x.y);
                    atomicAdd(dst + 2, x.z);
                }
            }
        },
        cnt);
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

__device__ __forceinline__ void push_shadow(const Queues& Q, uint32_t k, float3 o, uint32_t pixel, float3 dir,
                                            float tmax, float4 contrib, float4 local, uint32_t slot,
                                            bool records) {
    Q.so[k] = make_float4(o.x, o.y, o.z, __uint_as_float(pixel));
    Q.sd[k] = make_float4(dir.x, dir.y, dir.z, tmax);
    Q.sc[k] = contrib;
    if (records) Q.sx[k] = make_float4(local.x, local.y, local.z, __uint_as_float(slot));
}

// Shade one level.  GUIDED kernels run one wave per block (the guide's
// forward pass is separated by block barriers).
template <bool GUIDED>
__global__ __launch_bounds__(kRBlock) void k_shade(RenderArgs A, Queues Q, int level, int buf) {
    __shared__ __align__(16) uint16_t glds[GUIDED ? 64 * (kGStrideIn + kGStrideHid) : 1];
    const uint32_t n = min(Q.cnt[level], Q.cap);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const bool spawn = level < A.max_depth && A.mc > 0;
    const uint32_t m = (uint32_t)max(A.mc, 0);
    const bool cosine = A.flags & TN_RENDER_COSINE;
    const bool area = (A.flags & TN_RENDER_AREA_LIGHTS) && A.nemit > 0;
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
            const uint4 h = Q.hit[i];
            if (h.x == kInvalidId) {
                accumulate(A.accum, pixel, mul4(w, A.bg));  // pathtracer.cpp:26-28
                if (records) A.rec_x[slot] = A.bg;
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
                if (A.nlights) {  // pathtracer.cpp:44-58
                    float3 l = make_float3(A.light_pos.x - p.x, A.light_pos.y - p.y, A.light_pos.z - p.z);
                    const float il = 1.0f / sqrtf(l.x * l.x + l.y * l.y + l.z * l.z);
                    ldir = make_float3(il * l.x, il * l.y, il * l.z);
                    const float3 l2 = make_float3(A.light_pos.x - p2.x, A.light_pos.y - p2.y, A.light_pos.z - p2.z);
                    ldist = sqrtf(l2.x * l2.x + l2.y * l2.y + l2.z * l2.z);
                    const float c = fmaxf(0.f, ldir.x * nn.x + ldir.y * nn.y + ldir.z * nn.z);
                    lcontrib = scale4(inv_pi, mul4(wr, scale4(c, A.light_col)));
                    if (records) llocal = scale4(inv_pi, mul4(rho, scale4(c, A.light_col)));
                    has_point = true;
                }
                if (area) {
                    const float4 le = A.emission[id];
                    const float4 g = A.trirec[4 * (size_t)id + 1];  // geometric normal
                    if (level == 0 && (le.x > 0 || le.y > 0 || le.z > 0) &&
                        g.x * qd.x + g.y * qd.y + g.z * qd.z < 0)
                        accumulate(A.accum, pixel, mul4(w, le));
                    // one emitter by area, uniform point on it
                    const float u0 = rng_uniform(rng), u1 = rng_uniform(rng), u2 = rng_uniform(rng);
                    const float x = u0 * A.emit_area;
                    uint32_t lo = 0, hi = A.nemit - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (A.emit_cdf[mid] <= x) lo = mid + 1;
                        else hi = mid;
                    }
                    const uint32_t k = A.emit_ids[lo];
                    const float4 e0 = A.trirec[4 * (size_t)k], en = A.trirec[4 * (size_t)k + 1],
                                 eu = A.trirec[4 * (size_t)k + 2], ev = A.trirec[4 * (size_t)k + 3];
                    const float su = sqrtf(u1);
                    const float bu = su * (1.0f - u2), bv = su * u2;
                    const float3 q = make_float3(e0.x + bu * eu.x + bv * ev.x, e0.y + bu * eu.y + bv * ev.y,
                                                 e0.z + bu * eu.z + bv * ev.z);
                    const float3 dl = make_float3(q.x - p.x, q.y - p.y, q.z - p.z);
                    const float d2 = dl.x * dl.x + dl.y * dl.y + dl.z * dl.z;
                    const float id_ = 1.0f / sqrtf(d2);
                    adir = make_float3(dl.x * id_, dl.y * id_, dl.z * id_);
                    const float cosx = adir.x * nn.x + adir.y * nn.y + adir.z * nn.z;
                    const float cosl = -(adir.x * en.x + adir.y * en.y + adir.z * en.z);
                    if (cosx > 0 && cosl > 0) {
                        const float3 q2 = make_float3(q.x - p2.x, q.y - p2.y, q.z - p2.z);
                        adist = sqrtf(q2.x * q2.x + q2.y * q2.y + q2.z * q2.z) * (1.0f - 1e-4f);
                        const float4 lek = A.emission[k];
                        const float gterm = cosx * cosl / d2 * A.emit_area;
                        acontrib = scale4(inv_pi, mul4(wr, scale4(gterm, lek)));
                        if (records) alocal = scale4(inv_pi, mul4(rho, scale4(gterm, lek)));
                        has_area = true;
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
                if (k < Q.scap) push_shadow(Q, k, p2, pixel, ldir, ldist, lcontrib, llocal, slot, records);
                else atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
            }
            if (has_area) {
                const uint32_t k = sbase + (uint32_t)__popcll(mp) + lanes_below(ma);
                if (k < Q.scap) push_shadow(Q, k, p2, pixel, adir, adist, acontrib, alocal, slot, records);
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
                    Q.rng[ob][k] = rng_seed(((uint64_t)rng_next(rng) << 32) | rng_next(rng), c);
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
                    Q.rng[ob][k] = rng_seed(((uint64_t)rng_next(rng) << 32) | rng_next(rng), c);
                } else {
                    atomicOr(&Q.cnt[4 * kMaxLevels], 1u);
                }
            }
        }
    }
}

// Clears the training records of a chunk: valid flags and found radiance.
__global__ void k_rec_clear(RenderArgs A, uint32_t slots) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= slots) return;
    A.rec_n[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    A.rec_x[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Walks each path of the chunk backwards (L_k = X_{k+1} + g_{k+1} L_{k+1})
// and appends (p, n, w, lum(L_k) cos / pdf) of every bounce that brought
// light back to the training set (kept with probability A.accept).  Sample
// layout: (p, c), (n, cos / pdf), (w, path << 6 | level as bits).
__global__ __launch_bounds__(kRBlock) void k_collect(RenderArgs A, uint32_t npaths, int levels_with_records,
                                                     uint64_t salt) {
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
            A.ts[3 * (size_t)t + 2] = make_float4(rw.x, rw.y, rw.z, __uint_as_float((pid << 6) | (uint32_t)k));
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
};

std::mutex g_render_mutex;

// Carves the queue buffers out of one allocation; with base == nullptr it
// only measures (every piece is rounded up to 256 bytes).
size_t carve(Queues& q, char* base, uint32_t cap) {
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
    q.so = (float4*)take((size_t)q.scap * 16);
    q.sd = (float4*)take((size_t)q.scap * 16);
    q.sc = (float4*)take((size_t)q.scap * 16);
    q.sx = (float4*)take((size_t)q.scap * 16);
    q.cnt = (uint32_t*)take((4 * kMaxLevels + 1) * sizeof(uint32_t));
    q.totals = (unsigned long long*)take(2 * sizeof(unsigned long long));
    q.cap = cap;
    return off;
}

int alloc_queues(RenderBuffers& rb, uint32_t cap) {
    Queues probe{};
    const size_t need = carve(probe, nullptr, cap);
    if (rb.base && rb.bytes >= need && rb.q.cap >= cap) return TN_OK;
    if (rb.base) (void)hipFree(rb.base);
    rb.base = nullptr;
    rb.bytes = 0;
    TN_HIP(hipMalloc(&rb.base, need));
    rb.bytes = need;
    const size_t used = carve(rb.q, rb.base, cap);
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

}  // namespace

// Renders params->width x height x spp samples, adding the per-pixel SUM of
// sample radiance into accum (device, RGBA).
int render_into(tn_scene* sc, const tn_camera* cam, const tn_light* lights, uint32_t nl,
                const tn_render_params* P, float* accum, tn_render_stats* st, hipStream_t stream,
                const float* ray_o = nullptr, const float* ray_d = nullptr, uint64_t nrays = 0,
                tn_guide* guide = nullptr, uint32_t train_samples = 0) {
    const bool rays = ray_o != nullptr;
    const bool guided = guide != nullptr;
    const bool training = guided && train_samples > 0;
    if ((P->flags & TN_RENDER_GUIDED) && !guided)
        return set_error(TN_ERR_INVALID, "TN_RENDER_GUIDED needs a guide (tn_render_guided)");
    if (training && P->mc_samples != 1)
        return set_error(TN_ERR_INVALID, "training records need mc_samples == 1 (one path per sample)");
    if (guided && guide->device != sc->device)
        return set_error(TN_ERR_INVALID, "guide and scene live on different devices");
    if ((!rays && (P->width == 0 || P->height == 0 || P->pixel_samples < 1)) || P->max_depth < 0 ||
        P->mc_samples < 0)
        return set_error(TN_ERR_INVALID, "invalid render parameters");
    if (nl > 1) return set_error(TN_ERR_LIMIT, "at most one point light (main.cpp:123)");
    if ((uint32_t)P->max_depth + 2 > kMaxLevels) return set_error(TN_ERR_LIMIT, "max_depth too large");
    if (!rays && (size_t)P->width * P->height >= (1ull << 32))
        return set_error(TN_ERR_LIMIT, "image too large");
    if (rays && nrays >= (1ull << 32)) return set_error(TN_ERR_LIMIT, "too many rays");
    const uint32_t m = (uint32_t)P->mc_samples;
    // queue capacity: a chunk's deepest level holds chunk * m^max_depth segments
    double fan = 1;
    for (int k = 0; k < P->max_depth; ++k) fan *= std::max(1u, m);
    const double cap_target = 1u << 23;
    uint32_t chunk = P->chunk_paths;
    if (chunk == 0) chunk = (uint32_t)std::max(1.0, std::min(cap_target / fan, 1.0 * (1u << 22)));
    if (fan * chunk > 2147483647.0) return set_error(TN_ERR_LIMIT, "queue would exceed 2^31 entries");
    const uint32_t cap = (uint32_t)std::max(1.0, fan * chunk);

    std::lock_guard<std::mutex> lock(g_render_mutex);
    static RenderBuffers rb;
    int rc = alloc_queues(rb, cap);
    if (rc) return rc;

    RenderArgs A{};
    A.sv = sc->view();
    A.normals = sc->d_normals;
    A.albedo = sc->d_albedo;
    A.emission = sc->d_emission;
    A.trirec = sc->d_tris;
    A.emit_ids = sc->d_emit_ids;
    A.emit_cdf = sc->d_emit_cdf;
    A.nemit = sc->nemit;
    A.emit_area = sc->emit_area;
    if (cam) {
        A.cam_pos = make_float4(cam->pos[0], cam->pos[1], cam->pos[2], 0);
        std::memcpy(A.cam_m, cam->trafo, sizeof A.cam_m);
        A.cam_dx = cam->delta_x;
        A.cam_dy = cam->delta_y;
    }
    A.width = P->width;
    A.height = P->height;
    A.spp = rays ? 1u : (uint32_t)P->pixel_samples;
    A.max_depth = P->max_depth;
    A.mc = P->mc_samples;
    A.flags = P->flags;
    A.seed = P->seed;
    A.bg = make_float4(P->bg[0], P->bg[1], P->bg[2], P->bg[3]);
    A.nlights = nl;
    if (nl) {
        A.light_pos = make_float4(lights[0].pos[0], lights[0].pos[1], lights[0].pos[2], 0);
        A.light_col = make_float4(lights[0].color[0], lights[0].color[1], lights[0].color[2],
                                  lights[0].color[3]);
    }
    A.accum = (float4*)accum;
    A.seed_pass = P->seed + (uint64_t)P->sample_offset * 0xD1B54A32D192ED03ull;
    if (guided) {
        A.gv = guide->view();
        A.mis = guide->mis_guided;
    }
    static RecordBuffers recs;
    if (training) {
        A.rstride = (uint32_t)P->max_depth + 1;
        rc = alloc_records(recs, (size_t)chunk * A.rstride);
        if (rc) return rc;
        const size_t ns = recs.slots;
        A.rec_p = recs.base;
        A.rec_n = recs.base + ns;
        A.rec_w = recs.base + 2 * ns;
        A.rec_g = recs.base + 3 * ns;
        A.rec_x = recs.base + 4 * ns;
        // every bounce of every path is a candidate; keep about kTrainSet of them
        constexpr uint32_t kTrainSet = 1u << 21;
        rc = guide_reserve_set(guide, kTrainSet, stream);
        if (rc) return rc;
        const uint64_t total_paths = rays ? nrays : (uint64_t)P->width * P->height * (uint32_t)P->pixel_samples;
        const double candidates = (double)total_paths * std::max(1, P->max_depth);
        A.accept = (float)std::min(1.0, 0.95 * kTrainSet / std::max(1.0, candidates));
        A.ts = guide->ts;
        A.ts_count = guide->ts_count;
        A.ts_cap = guide->ts_cap;
    }
    float* d_jit = nullptr;
    if (!rays && (P->flags & TN_RENDER_REF_JITTER)) {
        std::vector<float> jit = reference_jitter(P->width, A.spp);
        TN_HIP(hipMalloc(&d_jit, jit.size() * sizeof(float)));
        TN_HIP(hipMemcpyAsync(d_jit, jit.data(), jit.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        A.jitter = d_jit;
    }
    k_zero_totals<<<1, 64, 0, stream>>>(rb.q.totals);
    const bool exh = P->flags & TN_RENDER_EXHAUSTIVE;
    const uint64_t total = rays ? nrays : (uint64_t)P->width * P->height * A.spp;
    const int levels = P->max_depth + 1;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, sc->device);
    const dim3 tgrid(cus * 8), tblock(kRBlock);
    auto t0 = std::chrono::steady_clock::now();
    for (uint64_t first = 0; first < total; first += chunk) {
        const uint32_t cnt0 = (uint32_t)std::min<uint64_t>(chunk, total - first);
        k_chunk_init<<<1, 4 * kMaxLevels + 64, 0, stream>>>(rb.q.cnt, cnt0);
        if (training) {
            const uint32_t slots = cnt0 * A.rstride;
            k_rec_clear<<<(slots + 255) / 256, 256, 0, stream>>>(A, slots);
        }
        if (rays)
            k_raygen_rays<<<(cnt0 + kRBlock - 1) / kRBlock, kRBlock, 0, stream>>>(A, rb.q, ray_o, ray_d,
                                                                                first, cnt0);
        else
            k_raygen<<<(cnt0 + kRBlock - 1) / kRBlock, kRBlock, 0, stream>>>(A, rb.q, first, cnt0);
        int buf = 0;
        for (int lv = 0; lv < levels; ++lv) {
            if (exh) k_trace<true><<<tgrid, tblock, 0, stream>>>(A, rb.q, lv, buf);
            else k_trace<false><<<tgrid, tblock, 0, stream>>>(A, rb.q, lv, buf);
            if (guided) k_shade<true><<<tgrid.x * (kRBlock / 64), 64, 0, stream>>>(A, rb.q, lv, buf);
            else k_shade<false><<<tgrid, tblock, 0, stream>>>(A, rb.q, lv, buf);
            if (exh) k_shadow<true><<<tgrid, tblock, 0, stream>>>(A, rb.q, lv);
            else k_shadow<false><<<tgrid, tblock, 0, stream>>>(A, rb.q, lv);
            buf ^= 1;
        }
        k_chunk_totals<<<1, 64, 0, stream>>>(rb.q.cnt, rb.q.totals, levels);
        if (training)
            k_collect<<<(cnt0 + kRBlock - 1) / kRBlock, kRBlock, 0, stream>>>(
                A, cnt0, P->max_depth, A.seed_pass ^ (first * 0x9E3779B97F4A7C15ull));
    }
    TN_HIP(hipGetLastError());
    if (training) {
        rc = guide_train_from_set(guide, train_samples, stream);
        if (rc) return rc;
    }
    unsigned long long tot[2] = {0, 0};
    uint32_t overflow = 0;
    TN_HIP(hipMemcpyAsync(tot, rb.q.totals, 16, hipMemcpyDeviceToHost, stream));
    TN_HIP(hipMemcpyAsync(&overflow, rb.q.cnt + 4 * kMaxLevels, 4, hipMemcpyDeviceToHost, stream));
    TN_HIP(hipStreamSynchronize(stream));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (d_jit) (void)hipFree(d_jit);
    if (overflow) return set_error(TN_ERR_QUEUE, "a ray queue overflowed");
    uint32_t status = 0;
    TN_HIP(hipMemcpy(&status, sc->d_status, 4, hipMemcpyDeviceToHost));
    if (st) {
        st->rays = tot[0];
        st->primary_rays = rays ? 0 : total;
        st->shadow_rays = tot[1];
        st->ms = ms;
    }
    return TN_OK;
}

}  // namespace tn

using namespace tn;

extern "C" {

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
    if (n == 0 || depth > P->max_depth) return TN_OK;  // pathtracer.cpp:17-19
    tn_render_params q = *P;
    q.max_depth = P->max_depth - depth;  // levels depth..max_depth remain
    return render_into(sc, nullptr, lights, nl, &q, rgba, st, (hipStream_t)stream, o3, d3, n);
}

int tn_render(tn_scene* sc, const tn_camera* cam, const tn_light* lights, uint32_t nl,
              const tn_render_params* P, float* linear, float* display, tn_render_stats* st) {
    if (!sc || !cam || !P || (nl && !lights)) return set_error(TN_ERR_INVALID, "NULL argument");
    if (P->width == 0 || P->height == 0 || P->pixel_samples < 1)
        return set_error(TN_ERR_INVALID, "invalid render parameters");
    const size_t np = (size_t)P->width * P->height;
    TN_HIP(hipSetDevice(sc->device));
    float* acc = nullptr;
    TN_HIP(hipMalloc(&acc, np * 16 * 2));
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
    (void)hipFree(acc);
    return rc;
}

}  // extern "C"
