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
//   * near-first early exit: the loop ends once the best hit lies before the
//     current leaf's exit by a 2^-16 relative margin (no later leaf can then
//     hold a strictly closer hit).  TN_ISECT_EXHAUSTIVE disables it.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace tn {

constexpr uint32_t kInvalidId = 1u << 30;
constexpr float kDirEps = 0.00001f;             // lib/types.h:14
constexpr float kExitMargin = 1.0f + 1.0f / 65536.0f;
constexpr int kMaxRestarts = 4096;

struct SceneView {
    const uint2* nodes;
    const uint32_t* refs;
    const float4* tris;  // 4 per triangle
    float bmin[3];
    float bmax[3];
};

struct TraversalCounters {
    uint32_t inner = 0, leaves = 0, refs = 0, tests = 0, restarts = 0;
    bool incomplete = false;
};

__device__ __forceinline__ float pick(float3 v, uint32_t ax) {
    return ax == 0 ? v.x : (ax == 1 ? v.y : v.z);
}

// lib/intersection.h:40-49 + 63-89: plane, then barycentrics; no culling.
__device__ __forceinline__ bool leaf_test_ref(const float4* __restrict__ T, uint32_t id, float3 o,
                                              float3 d, float& r, float& s, float& t) {
    const float4 a = T[4 * (size_t)id];      // v0, uv
    const float4 b = T[4 * (size_t)id + 1];  // n, vv
    const float den = b.x * d.x + b.y * d.y + b.z * d.z;
    if (den == 0.f) return false;  // r = lowest() < 0
    const float nom = b.x * (a.x - o.x) + b.y * (a.y - o.y) + b.z * (a.z - o.z);
    const float rr = nom / den;
    if (rr < 0) return false;
    const float4 c = T[4 * (size_t)id + 2];  // u, uu
    const float4 e = T[4 * (size_t)id + 3];  // v, denom
    const float px = o.x + rr * d.x, py = o.y + rr * d.y, pz = o.z + rr * d.z;
    const float wx = px - a.x, wy = py - a.y, wz = pz - a.z;
    const float wv = wx * e.x + wy * e.y + wz * e.z;
    const float wu = wx * c.x + wy * c.y + wz * c.z;
    const float ss = (a.w * wv - b.w * wu) / e.w;
    if (ss < 0) return false;
    const float tt = (a.w * wu - c.w * wv) / e.w;
    if (tt < 0 || 1 < ss + tt) return false;
    r = rr;
    s = ss;
    t = tt;
    return true;
}

// Moeller & Trumbore 1997 in FP32 on the record's v0, e1 = u, e2 = v.
__device__ __forceinline__ bool leaf_test_mt(const float4* __restrict__ T, uint32_t id, float3 o,
                                             float3 d, float& r, float& s, float& t) {
    const float4 a = T[4 * (size_t)id];
    const float4 c = T[4 * (size_t)id + 2];
    const float4 e = T[4 * (size_t)id + 3];
    const float px = d.y * e.z - d.z * e.y, py = d.z * e.x - d.x * e.z, pz = d.x * e.y - d.y * e.x;
    const float det = c.x * px + c.y * py + c.z * pz;
    if (det == 0.f) return false;
    const float inv = 1.0f / det;
    const float tx = o.x - a.x, ty = o.y - a.y, tz = o.z - a.z;
    const float u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.f || u > 1.f) return false;
    const float qx = ty * c.z - tz * c.y, qy = tz * c.x - tx * c.z, qz = tx * c.y - ty * c.x;
    const float v = (d.x * qx + d.y * qy + d.z * qz) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    const float tt = (e.x * qx + e.y * qy + e.z * qz) * inv;
    if (tt < 0.f) return false;
    r = tt;
    s = u;
    t = v;
    return true;
}

// Closest hit (ANY = false) or visibility up to tmax_any (ANY = true, returns
// the id of some triangle hit at 0 <= r <= tmax_any).  stk points at this
// lane's first LDS stack slot; consecutive slots are `stride` entries apart.
template <int S, bool EXH, bool MT, bool ANY, bool COUNT>
__device__ __forceinline__ uint32_t traverse(const SceneView& sc, float3 o, float3 d, float tmax_any,
                                             float& r_out, float& s_out, float& t_out, uint2* stk,
                                             int stride, TraversalCounters& cnt) {
    static_assert((S & (S - 1)) == 0, "stack size must be a power of two");
    float3 fd = d;
    if (fd.x == 0.f) fd.x = kDirEps;
    if (fd.y == 0.f) fd.y = kDirEps;
    if (fd.z == 0.f) fd.z = kDirEps;
    const float3 dinv = make_float3(1 / fd.x, 1 / fd.y, 1 / fd.z);
    float tenter, texit;
    {
        const float tx1 = (sc.bmin[0] - o.x) * dinv.x, tx2 = (sc.bmax[0] - o.x) * dinv.x;
        float lo = fminf(tx1, tx2), hi = fmaxf(tx1, tx2);
        const float ty1 = (sc.bmin[1] - o.y) * dinv.y, ty2 = (sc.bmax[1] - o.y) * dinv.y;
        lo = fmaxf(lo, fminf(ty1, ty2));
        hi = fminf(hi, fmaxf(ty1, ty2));
        const float tz1 = (sc.bmin[2] - o.z) * dinv.z, tz2 = (sc.bmax[2] - o.z) * dinv.z;
        lo = fmaxf(lo, fminf(tz1, tz2));
        hi = fminf(hi, fmaxf(tz1, tz2));
        if (hi < lo) return kInvalidId;
        tenter = lo;
        texit = hi;
    }
    const float root_exit = texit;
    uint32_t node = 0;
    int head = 0, count = 0, restarts = 0;
    bool lost = false;
    uint32_t best_id = kInvalidId;
    float best_r = FLT_MAX, best_s = 0.f, best_t = 0.f;
    for (;;) {
        uint2 nd = sc.nodes[node];
        while ((nd.y & 3u) != 3u) {
            if (COUNT) cnt.inner++;
            const uint32_t ax = nd.y & 3u;
            const float tt = (__uint_as_float(nd.x) - pick(o, ax)) * pick(dinv, ax);
            uint32_t nearn = node + 1, farn = nd.y >> 2;
            if (pick(fd, ax) <= 0.f) {
                const uint32_t tmp = nearn;
                nearn = farn;
                farn = tmp;
            }
            if (texit < tt) {
                node = nearn;
            } else if (tt < tenter) {
                node = farn;
            } else {
                stk[head * stride] = make_uint2(farn, __float_as_uint(texit));
                head = (head + 1) & (S - 1);
                if (count < S) ++count;
                else lost = true;
                node = nearn;
                texit = tt;
            }
            nd = sc.nodes[node];
        }
        if (COUNT) cnt.leaves++;
        const uint32_t off = nd.x, nref = nd.y >> 2;
        for (uint32_t k = 0; k < nref; ++k) {
            const uint32_t id = sc.refs[off + k];
            float rr, ss, tt;
            const bool hit = MT ? leaf_test_mt(sc.tris, id, o, d, rr, ss, tt)
                                : leaf_test_ref(sc.tris, id, o, d, rr, ss, tt);
            if (COUNT) {
                cnt.refs++;
                cnt.tests++;
            }
            if (ANY) {
                if (hit && rr <= tmax_any) {
                    r_out = rr;
                    s_out = ss;
                    t_out = tt;
                    return id;
                }
            } else if (hit && rr < best_r) {
                best_r = rr;
                best_s = ss;
                best_t = tt;
                best_id = id;
            }
        }
        if (!EXH) {
            if (!ANY && best_id != kInvalidId && best_r * kExitMargin < texit) break;
            if (ANY && tmax_any * kExitMargin < texit) break;
        }
        if (count == 0) {
            if (!lost || !(texit < root_exit)) break;
            if (++restarts > kMaxRestarts) {
                cnt.incomplete = true;
                break;
            }
            node = 0;
            tenter = texit;
            texit = root_exit;
            lost = false;
            continue;
        }
        head = (head - 1) & (S - 1);
        --count;
        const uint2 e = stk[head * stride];
        node = e.x;
        tenter = texit;
        texit = __uint_as_float(e.y);
    }
    if (COUNT) cnt.restarts += restarts;
    r_out = best_r;
    s_out = best_s;
    t_out = best_t;
    return best_id;
}

}  // namespace tn
