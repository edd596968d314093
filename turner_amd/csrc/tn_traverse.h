// tn_traverse.h -- device kd-tree traversal for gfx950 (wave64).
//
// The traversal restates KDTreeIntersection::intersect (lib/kdtree.cpp:515-607)
// for one ray per lane:
//   * zero direction components become 1e-5 for traversal only (:503-511);
//   * root slab test without a t >= 0 clamp (lib/intersection.h:105-128);
//   * near child = left iff d[axis] > 0, the three-way texit<t / t<tenter /
//     both decision of :540-564;
//   * leaves are tested with the unmodified ray, closest hit with strict <,
//     so the first of equal hits in visit order wins (:566-607).
// All of it is FP32 with the reference's operation order, IEEE division and
// no FMA contraction (the library is compiled with -ffp-contract=off), so
// hit ids and r/s/t are bit-identical to the reference's.
//
// GPU-specific machinery:
//   * nodes are 8 bytes (see DevNode in tn_host.h); a leaf holds an offset
//     and count into a 32-bit triangle-reference array;
//   * triangle records are 64 bytes (4 x float4: v0|uv, n|vv, u|uu, v|denom),
//     one L2 line per triangle for the per-lane gathers;
//   * the stack is a circular short stack of S entries in LDS, 8 bytes per
//     entry (node, texit): the entry's tenter is always the exit of the last
//     leaf processed, so it is not stored.  When entries are lost to
//     overflow, the traversal restarts from the root at the last leaf's exit
//     (kd-restart), which visits the remaining leaves in the same order;
//   * near-first early exit: traversal ends once the best hit lies before the
//     current leaf's exit by a 2^-16 relative margin (no later leaf can then
//     hold a strictly closer hit).  TN_ISECT_EXHAUSTIVE disables it;
//   * the traversal is a per-lane state machine (trav_init / trav_tick: a
//     few node steps or one triangle test per tick) so lanes in different
//     phases share a wave's instructions and a wave can hand finished lanes
//     new rays.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace tn {

constexpr uint32_t kInvalidId = 1u << 30;
// Hit record of a ray a traversal of a refined tree leaves to the reference
// tree (not a triangle id: ids are below 2^30).
constexpr uint32_t kFixId = 0xFFFFFFFFu;
constexpr float kDirEps = 0.00001f;  // lib/types.h:14
constexpr float kExitMargin = 1.0f + 1.0f / 65536.0f;
constexpr int kMaxRestarts = 4096;
// Leaves one ray may visit before its traversal is abandoned as incomplete
// (far above any real ray; bounds every traversal loop).
constexpr uint32_t kMaxLeafVisits = 1u << 22;

// Root of one tree in a packed forest (many scenes in one node / reference /
// triangle allocation): its node index and root box.
struct SceneRoot {
    uint32_t node;
    float bmin[3];
    float bmax[3];
    uint32_t pad;
};

struct SceneView {
    const uint2* nodes;
    const uint32_t* refs;
    const float4* tris;  // 4 per triangle
    // The same records once per triangle reference, in reference order, as
    // two arrays of halves: lra[2r..2r+1] = (v0, uv) (n, vv) and
    // lrb[2r..2r+1] = (u, uu) (v, denom) of triangle refs[r].  The pairs of a
    // leaf are then contiguous and need no reference lookup.
    const float4* lra;
    const float4* lrb;
    float bmin[3];
    float bmax[3];
    const SceneRoot* roots = nullptr;  // forest: per-scene roots (bmin/bmax unused)
};

struct TraversalCounters {
    uint32_t inner = 0, leaves = 0, refs = 0, tests = 0, restarts = 0;
    uint32_t full = 0;  // tests that read the whole 64-byte record
    bool incomplete = false;
};

struct Trav {
    float3 o, d, dinv;
    float tenter, texit, root_exit;
    uint32_t node, root;
    uint32_t leaf_off, leaf_left;  // triangle references still to test in the current leaf
    int head, count, restarts;
    uint32_t visits;  // leaves visited
    bool lost;
    bool tie;  // another triangle was hit at exactly best_r (tie detection only)
    uint32_t best_id;
    float best_r, best_s, best_t;
};

__device__ __forceinline__ float pick(float3 v, uint32_t ax) {
    return ax == 0 ? v.x : (ax == 1 ? v.y : v.z);
}

// True when num / den is certainly < 0: opposite signs and a quotient that
// cannot underflow to -0 (|num / den| >= 2^-100).  Lets a test reject
// without the division while giving exactly the division's verdict.
__device__ __forceinline__ bool certainly_negative(float num, float den) {
    return ((num < 0.f) != (den < 0.f)) && num != 0.f &&
           fabsf(num) >= fabsf(den) * 0x1p-100f;
}

// lib/intersection.h:40-49 + 63-89: plane, then barycentrics; no culling.
// A candidate is also rejected when r is not below `bound` (closest hit: the
// current best, which it must beat strictly) or, with ANY, above it.
// Returns 2 on a hit, 1 for a miss decided after reading the whole record,
// 0 for a miss decided from its first 32 bytes; with TIE (closest hit) 3 for
// a hit at exactly r == bound, which does not replace the best.  pa: the
// record's first half (v0, uv) (n, vv); pc: its second half (u, uu) (v, denom).
template <bool ANY, bool TIE = false>
__device__ __forceinline__ int leaf_test_ref(const float4* __restrict__ pa, const float4* __restrict__ pc,
                                             float3 o, float3 d, float bound, float& r, float& s, float& t) {
    const float4 a = pa[0];  // v0, uv
    const float4 b = pa[1];  // n, vv
    const float den = b.x * d.x + b.y * d.y + b.z * d.z;
    if (den == 0.f) return 0;  // r = lowest() < 0
    const float nom = b.x * (a.x - o.x) + b.y * (a.y - o.y) + b.z * (a.z - o.z);
    if (certainly_negative(nom, den)) return 0;
    if (!ANY) {
        // r = nom / den certainly >= bound (it cannot beat the current best):
        // compare against bound * den widened by 2^-20, far beyond the
        // rounding of the division and of the product.
        const float q = bound * den * (1.0f + 0x1p-20f);
        if (fabsf(q) >= 0x1p-100f && (den > 0.f ? nom >= q : nom <= q)) return 0;
    }
    const float rr = nom / den;
    if (rr < 0) return 0;
    if ((ANY || TIE) ? !(rr <= bound) : !(rr < bound)) return 0;
    const float4 c = pc[0];  // u, uu
    const float4 e = pc[1];  // v, denom
    const float px = o.x + rr * d.x, py = o.y + rr * d.y, pz = o.z + rr * d.z;
    const float wx = px - a.x, wy = py - a.y, wz = pz - a.z;
    const float wv = wx * e.x + wy * e.y + wz * e.z;
    const float wu = wx * c.x + wy * c.y + wz * c.z;
    const float snum = a.w * wv - b.w * wu;
    const float tnum = a.w * wu - c.w * wv;
    // The reference rejects on s < 0, t < 0 or 1 < s + t; the verdict does
    // not depend on the order of the checks, so decide what the numerators
    // decide before dividing: a negative quotient, or s + t above 1 by a
    // 2^-18 margin (s and t then have the sign of the denominator, so the
    // sum of the numerators carries only rounding error).
    if (certainly_negative(snum, e.w) || certainly_negative(tnum, e.w)) return 1;
    {
        const float sum = snum + tnum;
        const float q = e.w * (1.0f + 0x1p-18f);
        if (fabsf(q) >= 0x1p-100f && (e.w < 0.f ? sum < q : sum > q)) return 1;
    }
    const float ss = snum / e.w;
    if (ss < 0) return 1;
    const float tt = tnum / e.w;
    if (tt < 0 || 1 < ss + tt) return 1;
    r = rr;
    s = ss;
    t = tt;
    return (TIE && !ANY && rr == bound) ? 3 : 2;
}

// Moeller & Trumbore 1997 in FP32 on the record's v0, e1 = u, e2 = v.
// Returns 2 on a hit, 1 on a miss (the test always reads v0, u and v).
template <bool ANY>
__device__ __forceinline__ int leaf_test_mt(const float4* __restrict__ pa, const float4* __restrict__ pc,
                                            float3 o, float3 d, float bound, float& r, float& s, float& t) {
    const float4 a = pa[0];
    const float4 c = pc[0];
    const float4 e = pc[1];
    const float px = d.y * e.z - d.z * e.y, py = d.z * e.x - d.x * e.z, pz = d.x * e.y - d.y * e.x;
    const float det = c.x * px + c.y * py + c.z * pz;
    if (det == 0.f) return 1;
    const float inv = 1.0f / det;
    const float tx = o.x - a.x, ty = o.y - a.y, tz = o.z - a.z;
    const float u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.f || u > 1.f) return 1;
    const float qx = ty * c.z - tz * c.y, qy = tz * c.x - tx * c.z, qz = tx * c.y - ty * c.x;
    const float v = (d.x * qx + d.y * qy + d.z * qz) * inv;
    if (v < 0.f || u + v > 1.f) return 1;
    const float tt = (e.x * qx + e.y * qy + e.z * qz) * inv;
    if (tt < 0.f) return 1;
    if (ANY ? !(tt <= bound) : !(tt < bound)) return 1;
    r = tt;
    s = u;
    t = v;
    return 2;
}

// A direction with a zero component is traversed as if that component were
// 1e-5 (kdtree.cpp:503-511): the traversal path then bends away from the ray
// itself, and only the reference tree's own leaves reproduce the reference's
// result for it.
__device__ __forceinline__ bool needs_reference(float3 d) { return d.x == 0.f || d.y == 0.f || d.z == 0.f; }

// Fixes the direction and clips the ray to the root box (kdtree.cpp:519-535)
// of its scene (the tree of `scene` in a forest).  Returns false when the root
// box is missed.
__device__ __forceinline__ bool trav_init(Trav& T, const SceneView& sc, float3 o, float3 d,
                                          uint32_t scene = 0) {
    T.o = o;
    T.d = d;
    float3 fd = d;
    if (fd.x == 0.f) fd.x = kDirEps;
    if (fd.y == 0.f) fd.y = kDirEps;
    if (fd.z == 0.f) fd.z = kDirEps;
    T.dinv = make_float3(1 / fd.x, 1 / fd.y, 1 / fd.z);
    const float3 di = T.dinv;
    float b0[3], b1[3];
    if (sc.roots) {
        const SceneRoot& R = sc.roots[scene];
        T.root = R.node;
        for (int k = 0; k < 3; ++k) {
            b0[k] = R.bmin[k];
            b1[k] = R.bmax[k];
        }
    } else {
        T.root = 0;
        for (int k = 0; k < 3; ++k) {
            b0[k] = sc.bmin[k];
            b1[k] = sc.bmax[k];
        }
    }
    const float tx1 = (b0[0] - o.x) * di.x, tx2 = (b1[0] - o.x) * di.x;
    float lo = fminf(tx1, tx2), hi = fmaxf(tx1, tx2);
    const float ty1 = (b0[1] - o.y) * di.y, ty2 = (b1[1] - o.y) * di.y;
    lo = fmaxf(lo, fminf(ty1, ty2));
    hi = fminf(hi, fmaxf(ty1, ty2));
    const float tz1 = (b0[2] - o.z) * di.z, tz2 = (b1[2] - o.z) * di.z;
    lo = fmaxf(lo, fminf(tz1, tz2));
    hi = fminf(hi, fmaxf(tz1, tz2));
    T.tenter = lo;
    T.texit = hi;
    T.root_exit = hi;
    T.node = T.root;
    T.leaf_off = 0;
    T.leaf_left = 0;
    T.head = 0;
    T.count = 0;
    T.restarts = 0;
    T.visits = 0;
    T.lost = false;
    T.tie = false;
    T.best_id = kInvalidId;
    T.best_r = FLT_MAX;
    T.best_s = 0.f;
    T.best_t = 0.f;
    return !(hi < lo);
}

// Descends from T.node to the next leaf (T.texit becomes the leaf's exit).
// Returns the leaf node.
template <int S, bool COUNT>
__device__ __forceinline__ uint2 trav_descend(Trav& T, const SceneView& sc, uint2* stk, int stride,
                                              TraversalCounters& cnt) {
    static_assert(S > 0 && (S & (S - 1)) == 0, "the circular stack index wraps with & (S - 1)");
    uint32_t node = T.node;
    const float tenter = T.tenter;
    float texit = T.texit;
    uint2 nd = sc.nodes[node];
    while ((nd.y & 3u) != 3u) {
        if (COUNT) cnt.inner++;
        const uint32_t ax = nd.y & 3u;
        const float tt = (__uint_as_float(nd.x) - pick(T.o, ax)) * pick(T.dinv, ax);
        uint32_t nearn = node + 1, farn = nd.y >> 2;
        if (pick(T.d, ax) < 0.f) {  // fixed direction <= 0 (zeros become +1e-5)
            const uint32_t tmp = nearn;
            nearn = farn;
            farn = tmp;
        }
        if (texit < tt) {
            node = nearn;
        } else if (tt < tenter) {
            node = farn;
        } else {
            stk[T.head * stride] = make_uint2(farn, __float_as_uint(texit));
            T.head = (T.head + 1) & (S - 1);
            if (T.count < S) ++T.count;
            else T.lost = true;
            node = nearn;
            texit = tt;
        }
        nd = sc.nodes[node];
    }
    if (COUNT) cnt.leaves++;
    T.node = node;
    T.texit = texit;
    return nd;
}

// After a leaf: early exit, next stack entry or restart.  Returns true when
// the traversal is complete.
template <int S, bool EXH, bool ANY>
__device__ __forceinline__ bool trav_after_leaf(Trav& T, float tmax_any, const uint2* stk,
                                                int stride, TraversalCounters& cnt) {
    static_assert(S > 0 && (S & (S - 1)) == 0, "the circular stack index wraps with & (S - 1)");
    const float texit = T.texit;
    if (!EXH) {
        if (!ANY && T.best_id != kInvalidId && T.best_r * kExitMargin < texit) return true;
        if (ANY && tmax_any * kExitMargin < texit) return true;
    }
    if (T.count == 0) {
        if (!T.lost || !(texit < T.root_exit)) return true;
        if (++T.restarts > kMaxRestarts) {
            cnt.incomplete = true;
            return true;
        }
        T.node = T.root;
        T.tenter = texit;
        T.texit = T.root_exit;
        T.lost = false;
        return false;
    }
    T.head = (T.head - 1) & (S - 1);
    --T.count;
    const uint2 e = stk[T.head * stride];
    T.node = e.x;
    T.tenter = texit;
    T.texit = __uint_as_float(e.y);
    return false;
}

// One leaf per step: descend to the next leaf, test all of its triangles,
// then the early-exit check and the move to the next stack entry.  Returns
// true when the traversal is complete.
template <int S, bool EXH, bool MT, bool ANY, bool COUNT>
__device__ __forceinline__ bool trav_step(Trav& T, const SceneView& sc, float tmax_any, uint2* stk,
                                          int stride, TraversalCounters& cnt) {
    static_assert((S & (S - 1)) == 0, "stack size must be a power of two");
    uint32_t node = T.node;
    float tenter = T.tenter, texit = T.texit;
    uint2 nd = sc.nodes[node];
    while ((nd.y & 3u) != 3u) {
        if (COUNT) cnt.inner++;
        const uint32_t ax = nd.y & 3u;
        const float tt = (__uint_as_float(nd.x) - pick(T.o, ax)) * pick(T.dinv, ax);
        uint32_t nearn = node + 1, farn = nd.y >> 2;
        if (pick(T.d, ax) < 0.f) {  // fixed direction <= 0 (zeros become +1e-5)
            const uint32_t tmp = nearn;
            nearn = farn;
            farn = tmp;
        }
        if (texit < tt) {
            node = nearn;
        } else if (tt < tenter) {
            node = farn;
        } else {
            stk[T.head * stride] = make_uint2(farn, __float_as_uint(texit));
            T.head = (T.head + 1) & (S - 1);
            if (T.count < S) ++T.count;
            else T.lost = true;
            node = nearn;
            texit = tt;
        }
        nd = sc.nodes[node];
    }
    if (COUNT) cnt.leaves++;
    if (++T.visits > kMaxLeafVisits) {
        cnt.incomplete = true;
        return true;
    }
    const uint32_t off = nd.x, nref = nd.y >> 2;
    for (uint32_t k = 0; k < nref; ++k) {
        const uint32_t id = sc.refs[off + k];
        float rr, ss, tt;
        const float bound = ANY ? tmax_any : T.best_r;
        const float4* pa = sc.tris + 4 * (size_t)id;
        const int code = MT ? leaf_test_mt<ANY>(pa, pa + 2, T.o, T.d, bound, rr, ss, tt)
                            : leaf_test_ref<ANY>(pa, pa + 2, T.o, T.d, bound, rr, ss, tt);
        if (COUNT) {
            cnt.refs++;
            cnt.tests++;
            cnt.full += code >= 1 ? 1u : 0u;
        }
        if (code == 2) {
            T.best_r = rr;
            T.best_s = ss;
            T.best_t = tt;
            T.best_id = id;
            if (ANY) return true;
        }
    }
    if (!EXH) {
        if (!ANY && T.best_id != kInvalidId && T.best_r * kExitMargin < texit) return true;
        if (ANY && tmax_any * kExitMargin < texit) return true;
    }
    if (T.count == 0) {
        if (!T.lost || !(texit < T.root_exit)) return true;
        if (++T.restarts > kMaxRestarts) {
            cnt.incomplete = true;
            return true;
        }
        T.node = T.root;
        T.tenter = texit;
        T.texit = T.root_exit;
        T.lost = false;
        return false;
    }
    T.head = (T.head - 1) & (S - 1);
    --T.count;
    const uint2 e = stk[T.head * stride];
    T.node = e.x;
    T.tenter = texit;
    T.texit = __uint_as_float(e.y);
    return false;
}

// Traversal schedules: one leaf per step (a wave runs its largest leaf) or
// ticks of a few node steps / one triangle test.
enum Schedule { kSchedLeaf = 0, kSchedTick = 1, kSchedCoop = 2 };

template <int SCHED, int S, bool EXH, bool MT, bool ANY, bool COUNT>
__device__ __forceinline__ bool trav_advance(Trav& T, const SceneView& sc, float tmax_any,
                                             uint2* stk, int stride, TraversalCounters& cnt);

constexpr int kDescendPerTick = 8;

// One tick of the traversal state machine: at most kDescendPerTick node
// steps towards the next leaf, then at most one triangle test, then -- when
// the leaf is exhausted -- the early-exit check and the move to the next
// stack entry.  Returns true when the traversal is complete.  Lanes of a
// wave advance through different phases in the same tick, so a wave never
// waits for the largest leaf among its lanes.
template <int S, bool EXH, bool MT, bool ANY, bool COUNT>
__device__ __forceinline__ bool trav_tick(Trav& T, const SceneView& sc, float tmax_any, uint2* stk,
                                          int stride, TraversalCounters& cnt) {
    static_assert((S & (S - 1)) == 0, "stack size must be a power of two");
    if (T.leaf_left == 0) {
        uint32_t node = T.node;
        float tenter = T.tenter, texit = T.texit;
        uint2 nd = sc.nodes[node];
        int steps = 0;
        while ((nd.y & 3u) != 3u && steps < kDescendPerTick) {
            if (COUNT) cnt.inner++;
            const uint32_t ax = nd.y & 3u;
            const float tt = (__uint_as_float(nd.x) - pick(T.o, ax)) * pick(T.dinv, ax);
            uint32_t nearn = node + 1, farn = nd.y >> 2;
            if (pick(T.d, ax) < 0.f) {  // fixed direction <= 0 (zeros become +1e-5)
                const uint32_t tmp = nearn;
                nearn = farn;
                farn = tmp;
            }
            if (texit < tt) {
                node = nearn;
            } else if (tt < tenter) {
                node = farn;
            } else {
                stk[T.head * stride] = make_uint2(farn, __float_as_uint(texit));
                T.head = (T.head + 1) & (S - 1);
                if (T.count < S) ++T.count;
                else T.lost = true;
                node = nearn;
                texit = tt;
            }
            nd = sc.nodes[node];
            ++steps;
        }
        T.node = node;
        T.tenter = tenter;
        T.texit = texit;
        if ((nd.y & 3u) != 3u) return false;  // still descending
        if (COUNT) cnt.leaves++;
        if (++T.visits > kMaxLeafVisits) {
            cnt.incomplete = true;
            return true;
        }
        T.leaf_off = nd.x;
        T.leaf_left = nd.y >> 2;
        if (T.leaf_left == 0) goto leaf_done;
    }
    {
        const uint32_t id = sc.refs[T.leaf_off];
        ++T.leaf_off;
        --T.leaf_left;
        float rr, ss, tt;
        const float bound = ANY ? tmax_any : T.best_r;
        const float4* pa = sc.tris + 4 * (size_t)id;
        const int code = MT ? leaf_test_mt<ANY>(pa, pa + 2, T.o, T.d, bound, rr, ss, tt)
                            : leaf_test_ref<ANY>(pa, pa + 2, T.o, T.d, bound, rr, ss, tt);
        if (COUNT) {
            cnt.refs++;
            cnt.tests++;
            cnt.full += code >= 1 ? 1u : 0u;
        }
        if (code == 2) {
            T.best_r = rr;
            T.best_s = ss;
            T.best_t = tt;
            T.best_id = id;
            if (ANY) return true;
        }
        if (T.leaf_left != 0) return false;
    }
leaf_done : {
    const float texit = T.texit;
    if (!EXH) {
        if (!ANY && T.best_id != kInvalidId && T.best_r * kExitMargin < texit) return true;
        if (ANY && tmax_any * kExitMargin < texit) return true;
    }
    if (T.count == 0) {
        if (!T.lost || !(texit < T.root_exit)) return true;
        if (++T.restarts > kMaxRestarts) {
            cnt.incomplete = true;
            return true;
        }
        T.node = T.root;
        T.tenter = texit;
        T.texit = T.root_exit;
        T.lost = false;
        return false;
    }
    T.head = (T.head - 1) & (S - 1);
    --T.count;
    const uint2 e = stk[T.head * stride];
    T.node = e.x;
    T.tenter = texit;
    T.texit = __uint_as_float(e.y);
    return false;
}
}

template <int SCHED, int S, bool EXH, bool MT, bool ANY, bool COUNT>
__device__ __forceinline__ bool trav_advance(Trav& T, const SceneView& sc, float tmax_any,
                                             uint2* stk, int stride, TraversalCounters& cnt) {
    if (SCHED == kSchedTick) return trav_tick<S, EXH, MT, ANY, COUNT>(T, sc, tmax_any, stk, stride, cnt);
    return trav_step<S, EXH, MT, ANY, COUNT>(T, sc, tmax_any, stk, stride, cnt);
}

// Whole traversal of one ray (closest hit, or visibility up to tmax_any).
template <int S, bool EXH, bool MT, bool ANY, bool COUNT>
__device__ __forceinline__ uint32_t traverse(const SceneView& sc, float3 o, float3 d, float tmax_any,
                                             float& r_out, float& s_out, float& t_out, uint2* stk,
                                             int stride, TraversalCounters& cnt) {
    Trav T;
    if (!trav_init(T, sc, o, d)) return kInvalidId;
    while (!trav_step<S, EXH, MT, ANY, COUNT>(T, sc, tmax_any, stk, stride, cnt)) {
    }
    if (COUNT) cnt.restarts += T.restarts;
    r_out = T.best_r;
    s_out = T.best_s;
    t_out = T.best_t;
    return T.best_id;
}

// Wave64 helpers for lane refill.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Inclusive scan (sum, or maximum with MAX) over the 64 lanes of a wave, all
// lanes active: row shifts by 1, 2, 4, 8 within each 16-lane row (lanes whose
// source falls outside the row add the identity 0), then lane 15 of each row
// into rows 1 and 3 and lane 31 into rows 2 and 3.  Six DPP instructions.
template <bool MAX>
__device__ __forceinline__ uint32_t wave_scan(uint32_t x) {
#define TN_SCAN_STEP(ctrl, rows, bound)                                                          \
    {                                                                                            \
        const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, ctrl, rows, 0xf, bound); \
        x = MAX ? max(x, t) : x + t;                                                             \
    }
    TN_SCAN_STEP(0x111, 0xf, true)   // row_shr:1
    TN_SCAN_STEP(0x112, 0xf, true)   // row_shr:2
    TN_SCAN_STEP(0x114, 0xf, true)   // row_shr:4
    TN_SCAN_STEP(0x118, 0xf, true)   // row_shr:8
    TN_SCAN_STEP(0x142, 0xa, false)  // row_bcast:15
    TN_SCAN_STEP(0x143, 0xc, false)  // row_bcast:31
#undef TN_SCAN_STEP
    return x;
}

// LDS of one wave for the cooperative loop: the closest-hit key of each lane's
// leaf and the start markers of the pooled pairs of one round.
struct CoopLds {
    unsigned long long slot[64];
    uint32_t mark[64];
};

// With tie detection: per lane also the largest key (~r bits, pair index) of
// the candidates, i.e. the last pair among those at the smallest distance,
// and so a second hit at exactly the winner's distance when it differs from
// the winner.
struct CoopLdsTie : CoopLds {
    unsigned long long slot2[64];
};


// Wave-cooperative traversal loop.  Every round, each lane with a ray
// descends to its next leaf; then the wave pools the (ray, triangle
// reference) pairs of all its lanes' leaves and tests them 64 at a time, so a
// round costs ceil(sum of leaf sizes / 64) tests instead of the largest leaf.
// Pair p belongs to the lane whose range [excl, incl) of the prefix sum of
// leaf sizes holds it.  In the round of pairs base..base+63 every lane whose
// range meets the round writes a marker (round stamp, lane) at its first pair
// of the round; a prefix maximum of the markers over the lanes gives each
// pair its owner, whose ray is fetched with ds_bpermute.  A hit is folded
// into the owner's 64-bit LDS slot with atomicMin on (r, pair index): the
// minimum is the first of the leaf's closest hits in leaf order, i.e. the
// reference's strict-< scan (kdtree.cpp:587-605); the owner re-tests that
// triangle to recover r, s, t.
// Finished lanes take the next rays of the wave's chunk.
//   chunk(begin, end): called by all lanes of the wave; returns false when no
//                      work is left (may update begin/end);
//   load(k, o, d, tmax, scene), store(k, T, traversed): per-ray I/O (scene:
//                      the ray's tree in a forest, 0 otherwise).
// TIE (closest hit, reference leaf test, WL = CoopLdsTie): T.tie tells at the
// end whether a second triangle was hit at exactly the final best distance
// (in the same leaf, or in a later leaf at the earlier best); the winner among
// such ties depends on the tree's visit order.
template <int S, bool EXH, bool MT, bool ANY, bool COUNT, bool TIE = false, class WL = CoopLds, class Chunk,
          class Load, class Store>
__device__ __forceinline__ void coop_loop(const SceneView& sc, uint2* stk, int stride,
                                          WL* wl, Chunk chunk, Load load,
                                          Store store, TraversalCounters& cnt) {
    static_assert(!TIE || (!ANY && !MT), "tie detection is for the reference leaf test's closest hit");
    const uint32_t lane = threadIdx.x & 63u;
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t next = 0, end = 0, ray = kNone;
    bool more = true;
    Trav T;
    float tmx = 0.f;
    for (;;) {
        const uint64_t idle = __ballot(ray == kNone);
        if (idle && next >= end && more) more = chunk(next, end);
        if (idle && next < end) {
            const uint32_t k = next + lanes_below(idle);
            next += (uint32_t)__popcll(idle);
            if (ray == kNone && k < end) {
                float3 o, d;
                uint32_t scene = 0;
                load(k, o, d, tmx, scene);
                if (trav_init(T, sc, o, d, scene)) ray = k;
                else store(k, T, false);
            }
        }
        if (__ballot(ray != kNone) == 0) {
            if (!more && next >= end) break;
            continue;
        }
        // phase 1: every active lane descends to its next leaf
        uint32_t loff = 0, lcnt = 0;
        if (ray != kNone) {
            const uint2 nd = trav_descend<S, COUNT>(T, sc, stk, stride, cnt);
            loff = nd.x;
            lcnt = nd.y >> 2;
        }
        // phase 2: pooled leaf tests
        const uint32_t incl = wave_scan<false>(lcnt);
        const uint32_t excl = incl - lcnt;
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        const uint32_t rbase = loff - excl;  // refs index of pair p of this lane: rbase + p
        const float bound_own = ANY ? tmx : T.best_r;
        wl->slot[lane] = ~0ull;
        wl->mark[lane] = 0u;
        if constexpr (TIE) wl->slot2[lane] = 0ull;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // owner of pair base + lane; called with all lanes active
        uint32_t stamp = 0;
        auto owner_of = [&](uint32_t base) {
            stamp += 64u;  // above every marker of the earlier rounds
            if (lcnt != 0 && excl < base + 64u && incl > base)
                wl->mark[excl > base ? excl - base : 0u] = stamp | lane;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            return (int)(wave_scan<true>(wl->mark[lane]) & 63u);
        };
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t item = base + lane;
            const int own = owner_of(base);
            const uint32_t rb = __shfl(rbase, own, 64);
            const float3 ro = make_float3(__shfl(T.o.x, own, 64), __shfl(T.o.y, own, 64),
                                          __shfl(T.o.z, own, 64));
            const float3 rd = make_float3(__shfl(T.d.x, own, 64), __shfl(T.d.y, own, 64),
                                          __shfl(T.d.z, own, 64));
            const float bnd = __shfl(bound_own, own, 64);
            if (item < total) {
                const size_t r = rb + item;  // reference slot of the pair
                float rr, ss, tt;
                int code = MT ? leaf_test_mt<ANY>(sc.lra + 2 * r, sc.lrb + 2 * r, ro, rd, bnd, rr, ss, tt)
                              : leaf_test_ref<ANY, TIE>(sc.lra + 2 * r, sc.lrb + 2 * r, ro, rd, bnd, rr, ss, tt);
                if (COUNT) {
                    cnt.refs++;
                    cnt.tests++;
                    cnt.full += code >= 1 ? 1u : 0u;
                }
                if (code >= 2) {
                    // r >= 0 here; -0 maps to +0 so the key orders like the reference's <
                    const uint32_t rbits = __float_as_uint(rr + 0.0f);
                    atomicMin(&wl->slot[own], ((unsigned long long)rbits << 32) | item);
                    if constexpr (TIE) atomicMax(&wl->slot2[own], ((unsigned long long)(~rbits) << 32) | item);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const unsigned long long key = wl->slot[lane];
        unsigned long long key2 = 0ull;
        if constexpr (TIE) key2 = wl->slot2[lane];
        // every active lane moves on after its leaf, empty leaves included
        if (ray != kNone) {
            // TIE: the leaf's hits lie only at exactly the current best distance
            // (code 3): the best stays; tied iff one of them is another
            // triangle -- two or more of them (a triangle occurs once per
            // leaf), or one that is not the best's (a triangle spanning
            // several leaves is met again in later ones)
            if (TIE && lcnt && key != ~0ull && T.best_id != kInvalidId &&
                (uint32_t)(key >> 32) == __float_as_uint(T.best_r + 0.0f)) {
                if ((uint32_t)key != (uint32_t)key2 || sc.refs[rbase + (uint32_t)key] != T.best_id) T.tie = true;
            } else if (lcnt && key != ~0ull) {
                const size_t r = rbase + (uint32_t)key;
                const uint32_t id = sc.refs[r];
                float rr, ss, tt;
                if (MT) leaf_test_mt<ANY>(sc.lra + 2 * r, sc.lrb + 2 * r, T.o, T.d, bound_own, rr, ss, tt);
                else leaf_test_ref<ANY>(sc.lra + 2 * r, sc.lrb + 2 * r, T.o, T.d, bound_own, rr, ss, tt);
                T.best_r = rr;
                T.best_s = ss;
                T.best_t = tt;
                T.best_id = id;
                // a new best: tied iff the last pair at its distance is another pair
                if constexpr (TIE)
                    T.tie = key2 != 0ull && (uint32_t)~(uint32_t)(key2 >> 32) == (uint32_t)(key >> 32) &&
                            (uint32_t)key2 != (uint32_t)key;
                if (ANY) {
                    store(ray, T, true);
                    ray = kNone;
                }
            }
            if (ray != kNone) {
                bool done = trav_after_leaf<S, EXH, ANY>(T, tmx, stk, stride, cnt);
                if (!done && ++T.visits > kMaxLeafVisits) {
                    cnt.incomplete = true;
                    done = true;
                }
                if (done) {
                    if (COUNT) cnt.restarts += T.restarts;
                    store(ray, T, true);
                    ray = kNone;
                }
            }
        }
    }
}

}  // namespace tn
