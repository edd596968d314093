/*
 * oracle.c -- CPU restatement of turner's pathtracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see oracle.h for
 * who may load it and what pins each function).  Built with
 * gcc -O2 -std=c11 -ffp-contract=off (no FMA, IEEE division), which matches
 * the reference build's FP environment (x86-64 SSE, ISO C++14, CMakeLists.txt:3).
 *
 * All citations are to the turner reference tree.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define EPS 0.00001f /* lib/types.h:14, src/turner.h:22 */

/* ---------------------------------------------------------------- vectors */
typedef struct {
    float x, y, z;
} v3;

static inline v3 mk(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static inline float comp(v3 a, int ax) { return ax == 0 ? a.x : (ax == 1 ? a.y : a.z); }
static inline void setc(v3* a, int ax, float v) {
    if (ax == 0) a->x = v;
    else if (ax == 1) a->y = v;
    else a->z = v;
}
/* src/geometry.h:223-225 (component-wise, left operand first) */
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
/* src/geometry.h:272-274: s * v = {s*v.x, s*v.y, s*v.z} */
static inline v3 scale(float s, v3 v) { return mk(s * v.x, s * v.y, s * v.z); }
/* src/geometry.h:280-282 and :778-788 (all dot overloads sum x, y, z in order) */
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* src/geometry.h:255-256 */
static inline float length(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
/* src/geometry.h:243-246,302-304,752-755,802-804: v / len == (1 / len) * v */
static inline v3 normalize(v3 a) {
    float inv = 1.0f / length(a);
    return scale(inv, a);
}
/* src/geometry.h:288-300: cross product in double, cast to float */
static inline v3 cross_d(v3 a, v3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return mk((float)((ay * bz) - (az * by)), (float)((az * bx) - (ax * bz)),
              (float)((ax * by) - (ay * bx)));
}
static inline v3 load3(const float* p) { return mk(p[0], p[1], p[2]); }
/* std::min / std::max as in libstdc++: min(a,b) = b<a ? b : a, max(a,b) = a<b ? b : a */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }

/* ---------------------------------------------------------------- boxes */
typedef struct {
    v3 lo, hi;
} box3;

/* src/geometry.h:998-999: Bbox3(p1, p2) = {min(p1,p2), max(p1,p2)} */
static inline box3 box_from(v3 p1, v3 p2) {
    box3 b;
    b.lo = mk(smin(p1.x, p2.x), smin(p1.y, p2.y), smin(p1.z, p2.z));
    b.hi = mk(smax(p1.x, p2.x), smax(p1.y, p2.y), smax(p1.z, p2.z));
    return b;
}
/* src/geometry.h:991-996: default box is [lowest, max] */
static inline box3 box_default(void) {
    box3 b;
    b.lo = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    b.hi = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    return b;
}
/* src/geometry.h:1021-1023 */
static inline int box_empty(box3 b) {
    return b.hi.x <= b.lo.x && b.hi.y <= b.lo.y && b.hi.z <= b.lo.z;
}
/* src/geometry.h:1025-1027 */
static inline int box_planar(box3 b, int ax) {
    return fabsf(comp(b.hi, ax) - comp(b.lo, ax)) < EPS;
}
/* src/geometry.h:1041-1044 */
static inline float box_area(box3 b) {
    v3 d = sub(b.hi, b.lo);
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
/* src/geometry.h:1062-1074 */
static inline void box_split(box3 b, int ax, float pos, box3* l, box3* r) {
    v3 lmax = b.hi;
    setc(&lmax, ax, pos);
    v3 rmin = b.lo;
    setc(&rmin, ax, pos);
    *l = box_from(b.lo, lmax);
    *r = box_from(rmin, b.hi);
}
/* lib/types.h:65-71 fmin/fmax of three, lib/triangle.h:61-71 bbox */
static inline float fmin3(float x, float y, float z) { return fminf(x, smin(y, z)); }
static inline float fmax3(float x, float y, float z) { return fmaxf(x, smax(y, z)); }
static box3 tri_bbox(const float* t9) {
    v3 lo = mk(fmin3(t9[0], t9[3], t9[6]), fmin3(t9[1], t9[4], t9[7]),
               fmin3(t9[2], t9[5], t9[8]));
    v3 hi = mk(fmax3(t9[0], t9[3], t9[6]), fmax3(t9[1], t9[4], t9[7]),
               fmax3(t9[2], t9[5], t9[8]));
    return box_from(lo, hi);
}

/* ---------------------------------------------------------------- triangle */
/* lib/triangle.h:22-39: u = v1-v0, v = v2-v0, normal = normalize(cross(u,v)),
 * uv, vv, uu, denom = uv*uv - uu*vv */
void or_precompute(const float* verts9, uint32_t n, float* out16) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* t = verts9 + 9 * (size_t)i;
        v3 v0 = load3(t), v1 = load3(t + 3), v2 = load3(t + 6);
        v3 u = sub(v1, v0), v = sub(v2, v0);
        v3 nn = normalize(cross_d(u, v));
        float uv = dot(u, v), vv = dot(v, v), uu = dot(u, u);
        float denom = uv * uv - uu * vv;
        float* o = out16 + 16 * (size_t)i;
        o[0] = v0.x; o[1] = v0.y; o[2] = v0.z;
        o[3] = nn.x; o[4] = nn.y; o[5] = nn.z;
        o[6] = u.x; o[7] = u.y; o[8] = u.z;
        o[9] = v.x; o[10] = v.y; o[11] = v.z;
        o[12] = uv; o[13] = vv; o[14] = uu; o[15] = denom;
    }
}

/* lib/intersection.h:40-49 + :63-89 (plane, then barycentric; no culling) */
static inline int leaf_test(const float* p, v3 o, v3 d, float* r, float* s, float* t) {
    v3 v0 = load3(p), n = load3(p + 3), u = load3(p + 6), v = load3(p + 9);
    float uv = p[12], vv = p[13], uu = p[14], denom = p[15];
    float den = dot(n, d);
    float rr;
    if (den == 0.f) {
        rr = -FLT_MAX;
    } else {
        float nom = dot(n, sub(v0, o));
        rr = nom / den;
    }
    *r = rr;
    if (rr < 0) return 0;
    v3 P = add(o, scale(rr, d));
    v3 w = sub(P, v0);
    float wv = dot(w, v);
    float wu = dot(w, u);
    float ss = (uv * wv - vv * wu) / denom;
    *s = ss;
    if (ss < 0) return 0;
    float tt = (uv * wu - uu * wv) / denom;
    *t = tt;
    if (tt < 0 || 1 < ss + tt) return 0;
    return 1;
}

/* lib/intersection.h:105-128 */
static inline int ray_box(v3 o, v3 d, box3 b, float* tmin, float* tmax) {
    v3 dinv = mk(1 / d.x, 1 / d.y, 1 / d.z);
    float tx1 = (b.lo.x - o.x) * dinv.x;
    float tx2 = (b.hi.x - o.x) * dinv.x;
    float lo = fminf(tx1, tx2);
    float hi = fmaxf(tx1, tx2);
    float ty1 = (b.lo.y - o.y) * dinv.y;
    float ty2 = (b.hi.y - o.y) * dinv.y;
    lo = fmaxf(lo, fminf(ty1, ty2));
    hi = fminf(hi, fmaxf(ty1, ty2));
    float tz1 = (b.lo.z - o.z) * dinv.z;
    float tz2 = (b.hi.z - o.z) * dinv.z;
    lo = fmaxf(lo, fminf(tz1, tz2));
    hi = fminf(hi, fmaxf(tz1, tz2));
    *tmin = lo;
    *tmax = hi;
    return !(hi < lo);
}

int or_ray_box(const float* o, const float* d, const float* box6, float* tmin,
               float* tmax) {
    box3 b;
    b.lo = load3(box6);
    b.hi = load3(box6 + 3);
    return ray_box(load3(o), load3(d), b, tmin, tmax);
}

/* ---------------------------------------------------------------- clipping */
enum { ON_PLANE = 0, BEHIND = 1, IN_FRONT = 2 };
/* lib/clipping.h:121-130 */
static inline int classify(v3 p, v3 n, float d) {
    float dist = dot(n, p) - d;
    if (dist > EPS) return IN_FRONT;
    if (dist < -EPS) return BEHIND;
    return ON_PLANE;
}
/* lib/intersection.h:17-22 */
static inline float seg_plane_t(v3 a, v3 b, v3 n, float d) {
    v3 ab = sub(b, a);
    return (d - dot(n, a)) / dot(n, ab);
}
/* lib/clipping.h:135-187 (Sutherland-Hodgman at a thick plane) */
static int clip_poly(const v3* in, int nin, v3* out, v3 n, float d) {
    int nout = 0;
    v3 a = in[nin - 1];
    int a_side = classify(a, n, d);
    for (int i = 0; i < nin; ++i) {
        v3 b = in[i];
        int b_side = classify(b, n, d);
        if (b_side == IN_FRONT) {
            if (a_side == BEHIND) {
                float t = seg_plane_t(a, b, n, d);
                out[nout++] = add(a, scale(t, sub(b, a)));
            }
            out[nout++] = b;
        } else if (b_side == BEHIND) {
            if (a_side == IN_FRONT) {
                float t = seg_plane_t(a, b, n, d);
                out[nout++] = add(a, scale(t, sub(b, a)));
            }
        } else {
            out[nout++] = b;
        }
        a = b;
        a_side = b_side;
    }
    return nout;
}

/* lib/clipping.h:199-235: returns the bbox of the clipped polygon, or the
 * default box [lowest, max] when fewer than 2 points survive a plane. */
static box3 clip_triangle(const float* t9, box3 box, int* npts_out) {
    v3 bufa[40], bufb[40];
    v3* pts = bufa;
    v3* nxt = bufb;
    int np = 3;
    pts[0] = load3(t9);
    pts[1] = load3(t9 + 3);
    pts[2] = load3(t9 + 6);
    for (int ax = 0; ax < 3; ++ax) {
        for (int side = 0; side < 2; ++side) {
            v3 n = mk(0, 0, 0);
            setc(&n, ax, side == 0 ? 1.f : -1.f);
            float dist = side == 0 ? comp(box.lo, ax) : -comp(box.hi, ax);
            np = clip_poly(pts, np, nxt, n, dist);
            v3* tmp = pts;
            pts = nxt;
            nxt = tmp;
            if (np < 2) {
                *npts_out = np;
                return box_default();
            }
        }
    }
    v3 lo = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    v3 hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int ax = 0; ax < 3; ++ax) {
        for (int i = 0; i < np; ++i) {
            float c = comp(pts[i], ax);
            if (c < comp(lo, ax)) setc(&lo, ax, c);
            if (comp(hi, ax) < c) setc(&hi, ax, c);
        }
    }
    *npts_out = np;
    return box_from(lo, hi);
}

int or_clip_triangle(const float* tri9, const float* box6, float* out6) {
    box3 b;
    b.lo = load3(box6);
    b.hi = load3(box6 + 3);
    int np = 0;
    box3 r = clip_triangle(tri9, b, &np);
    out6[0] = r.lo.x; out6[1] = r.lo.y; out6[2] = r.lo.z;
    out6[3] = r.hi.x; out6[4] = r.hi.y; out6[5] = r.hi.z;
    return np;
}

/* ------------------------------------------------- libstdc++ std::sort */
/* The reference sorts SAH events with std::sort (lib/kdtree.cpp:312-316).
 * Equal (point, type) events keep std::sort's (unstable) order, which fixes
 * the triangle order inside leaves and hence which of two equal-distance
 * hits wins.  This restates libstdc++'s introsort (bits/stl_algo.h and
 * bits/stl_heap.h, unchanged from GCC 4.x to GCC 11.4): median-of-3 pivot
 * moved to first, unguarded Hoare partition, heap sort at depth 2*lg(n),
 * final insertion sort with threshold 16. */
#define DEFINE_STD_SORT(NAME, T, LESS)                                          \
    static void NAME##_push_heap(T* f, ptrdiff_t hole, ptrdiff_t top, T val) {  \
        ptrdiff_t parent = (hole - 1) / 2;                                      \
        while (hole > top && LESS(f[parent], val)) {                            \
            f[hole] = f[parent];                                                \
            hole = parent;                                                      \
            parent = (hole - 1) / 2;                                            \
        }                                                                       \
        f[hole] = val;                                                          \
    }                                                                           \
    static void NAME##_adjust_heap(T* f, ptrdiff_t hole, ptrdiff_t len, T val) { \
        const ptrdiff_t top = hole;                                             \
        ptrdiff_t child = hole;                                                 \
        while (child < (len - 1) / 2) {                                         \
            child = 2 * (child + 1);                                            \
            if (LESS(f[child], f[child - 1])) child--;                          \
            f[hole] = f[child];                                                 \
            hole = child;                                                       \
        }                                                                       \
        if ((len & 1) == 0 && child == (len - 2) / 2) {                         \
            child = 2 * (child + 1);                                            \
            f[hole] = f[child - 1];                                             \
            hole = child - 1;                                                   \
        }                                                                       \
        NAME##_push_heap(f, hole, top, val);                                    \
    }                                                                           \
    static void NAME##_heap_sort(T* first, T* last) {                           \
        ptrdiff_t len = last - first;                                           \
        if (len >= 2) {                                                         \
            ptrdiff_t parent = (len - 2) / 2;                                   \
            for (;;) {                                                          \
                T v = first[parent];                                            \
                NAME##_adjust_heap(first, parent, len, v);                      \
                if (parent == 0) break;                                         \
                parent--;                                                       \
            }                                                                   \
        }                                                                       \
        while (last - first > 1) {                                              \
            --last;                                                             \
            T v = *last;                                                        \
            *last = *first;                                                     \
            NAME##_adjust_heap(first, 0, last - first, v);                      \
        }                                                                       \
    }                                                                           \
    static inline void NAME##_swap(T* a, T* b) {                                \
        T t = *a;                                                               \
        *a = *b;                                                                \
        *b = t;                                                                 \
    }                                                                           \
    static void NAME##_median_to_first(T* res, T* a, T* b, T* c) {              \
        if (LESS(*a, *b)) {                                                     \
            if (LESS(*b, *c)) NAME##_swap(res, b);                              \
            else if (LESS(*a, *c)) NAME##_swap(res, c);                         \
            else NAME##_swap(res, a);                                           \
        } else if (LESS(*a, *c)) NAME##_swap(res, a);                           \
        else if (LESS(*b, *c)) NAME##_swap(res, c);                             \
        else NAME##_swap(res, b);                                               \
    }                                                                           \
    static T* NAME##_partition(T* first, T* last, T* pivot) {                   \
        for (;;) {                                                              \
            while (LESS(*first, *pivot)) ++first;                               \
            --last;                                                             \
            while (LESS(*pivot, *last)) --last;                                 \
            if (!(first < last)) return first;                                  \
            NAME##_swap(first, last);                                           \
            ++first;                                                            \
        }                                                                       \
    }                                                                           \
    static void NAME##_loop(T* first, T* last, ptrdiff_t depth) {               \
        while (last - first > 16) {                                             \
            if (depth == 0) {                                                   \
                NAME##_heap_sort(first, last);                                  \
                return;                                                         \
            }                                                                   \
            --depth;                                                            \
            T* mid = first + (last - first) / 2;                                \
            NAME##_median_to_first(first, first + 1, mid, last - 1);            \
            T* cut = NAME##_partition(first + 1, last, first);                  \
            NAME##_loop(cut, last, depth);                                      \
            last = cut;                                                         \
        }                                                                       \
    }                                                                           \
    static inline void NAME##_linear_insert(T* last) {                          \
        T val = *last;                                                          \
        T* next = last - 1;                                                     \
        while (LESS(val, *next)) {                                              \
            *last = *next;                                                      \
            last = next;                                                        \
            --next;                                                             \
        }                                                                       \
        *last = val;                                                            \
    }                                                                           \
    static void NAME##_insertion(T* first, T* last) {                           \
        if (first == last) return;                                              \
        for (T* i = first + 1; i != last; ++i) {                                \
            if (LESS(*i, *first)) {                                             \
                T val = *i;                                                     \
                memmove(first + 1, first, (size_t)(i - first) * sizeof(T));     \
                *first = val;                                                   \
            } else {                                                            \
                NAME##_linear_insert(i);                                        \
            }                                                                   \
        }                                                                       \
    }                                                                           \
    static void NAME(T* first, T* last) {                                       \
        if (first == last) return;                                              \
        ptrdiff_t n = last - first;                                             \
        ptrdiff_t lg = (ptrdiff_t)(63 - __builtin_clzll((unsigned long long)n)); \
        NAME##_loop(first, last, lg * 2);                                       \
        if (last - first > 16) {                                                \
            NAME##_insertion(first, first + 16);                                \
            for (T* i = first + 16; i != last; ++i) NAME##_linear_insert(i);    \
        } else {                                                                \
            NAME##_insertion(first, last);                                      \
        }                                                                       \
    }

/* lib/kdtree.cpp:252-265: event types and record */
enum { ENDING = 0, PLANAR = 1, STARTING = 2 };
typedef struct {
    uint32_t id;
    float point;
    float point_aux;
    int type;
} event_t;
/* lib/kdtree.cpp:313-315 */
#define EVENT_LESS(a, b) ((a).point < (b).point || ((a).point == (b).point && (a).type < (b).type))
DEFINE_STD_SORT(sort_events, event_t, EVENT_LESS)

#define U64_HI_LESS(a, b) (((a) >> 32) < ((b) >> 32))
DEFINE_STD_SORT(sort_u64_hi, uint64_t, U64_HI_LESS)

void or_introsort_u64(uint64_t* keys, size_t n) { sort_u64_hi(keys, keys + n); }

/* ---------------------------------------------------------------- build */
typedef struct tnode {
    int leaf;
    int axis;
    float split;
    struct tnode *left, *right;
    uint32_t* ids;
    size_t nids;
} tnode;

typedef struct {
    uint32_t* v;
    size_t n, cap;
} idvec;
static void idvec_push(idvec* a, uint32_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 16;
        a->v = (uint32_t*)realloc(a->v, a->cap * sizeof(uint32_t));
    }
    a->v[a->n++] = x;
}

typedef struct {
    const float* verts9;
} builder_t;

/* lib/kdtree.cpp:182-187 */
static inline float lambda_(size_t nl, size_t nr) { return (nl == 0 || nr == 0) ? 0.8f : 1.0f; }
/* lib/kdtree.cpp:197-203 (COST_TRAVERSAL = 15, COST_INTERSECTION = 20) */
static inline float cost_(float lr, float rr, size_t nl, size_t nr) {
    return lambda_(nl, nr) * (15 + 20 * (lr * (float)nl + rr * (float)nr));
}

static tnode* new_leaf(const uint32_t* ids, size_t n) {
    tnode* t = (tnode*)calloc(1, sizeof(tnode));
    t->leaf = 1;
    t->ids = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    memcpy(t->ids, ids, n * sizeof(uint32_t));
    t->nids = n;
    return t;
}

/* lib/kdtree.cpp:245-408 (WH06 Algorithm 4: event sweep + classify) */
static float find_plane(const builder_t* B, const uint32_t* tris, size_t ntris,
                        box3 box, int* out_ax, float* out_pos, idvec* L,
                        idvec* R) {
    event_t* ev[3];
    size_t nev[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) ev[a] = (event_t*)malloc(2 * ntris * sizeof(event_t) + 1);
    size_t num_tris = 0;
    for (size_t i = 0; i < ntris; ++i) {
        uint32_t id = tris[i];
        int np;
        box3 cb = clip_triangle(B->verts9 + 9 * (size_t)id, box, &np);
        if (box_empty(cb)) continue;
        num_tris += 1;
        for (int a = 0; a < 3; ++a) {
            if (box_planar(cb, a)) {
                event_t e = {id, comp(cb.lo, a), comp(cb.lo, a), PLANAR};
                ev[a][nev[a]++] = e;
            } else {
                event_t e1 = {id, comp(cb.lo, a), comp(cb.hi, a), STARTING};
                event_t e2 = {id, comp(cb.hi, a), comp(cb.lo, a), ENDING};
                ev[a][nev[a]++] = e1;
                ev[a][nev[a]++] = e2;
            }
        }
    }
    if (num_tris == 0) {
        for (int a = 0; a < 3; ++a) free(ev[a]);
        *out_ax = 0;
        *out_pos = 0;
        return FLT_MAX;
    }
    float min_cost = FLT_MAX;
    int min_ax = 0;
    float min_pos = 0;
    int min_side_left = 1;
    float area = box_area(box);
    for (int a = 0; a < 3; ++a) {
        event_t* E = ev[a];
        size_t n = nev[a];
        sort_events(E, E + n);
        size_t nl = 0, np = 0, nr = num_tris;
        for (size_t i = 0; i < n;) {
            float p = E[i].point;
            int ps = 0, pe = 0, pp = 0;
            while (i < n && E[i].point == p && E[i].type == ENDING) { pe++; i++; }
            while (i < n && E[i].point == p && E[i].type == PLANAR) { pp++; i++; }
            while (i < n && E[i].point == p && E[i].type == STARTING) { ps++; i++; }
            np = (size_t)pp;
            nr -= (size_t)(pp + pe);
            /* lib/kdtree.cpp:218-238 surface_area_heuristics */
            box3 lb, rb;
            box_split(box, a, p, &lb, &rb);
            float lr = box_area(lb) / area;
            float rr = box_area(rb) / area;
            float lcost = cost_(lr, rr, nl + np, nr);
            float rcost = cost_(lr, rr, nl, np + nr);
            float c;
            int left;
            if (lcost < rcost) { c = lcost; left = 1; }
            else { c = rcost; left = 0; }
            if (c < min_cost) {
                min_cost = c;
                min_ax = a;
                min_pos = p;
                min_side_left = left;
            }
            nl += (size_t)(ps + pp);
            np = 0;
        }
    }
    /* classify: lib/kdtree.cpp:369-399 */
    event_t* E = ev[min_ax];
    size_t n = nev[min_ax];
    for (size_t i = 0; i < n; ++i) {
        const event_t* e = &E[i];
        if (e->point < min_pos) {
            if (e->type == ENDING || e->type == PLANAR) {
                idvec_push(L, e->id);
            } else if (min_pos < e->point_aux) {
                idvec_push(L, e->id);
                idvec_push(R, e->id);
            }
        } else if (e->point == min_pos) {
            if (e->type == ENDING) idvec_push(L, e->id);
            else if (e->type == PLANAR) {
                if (min_side_left) idvec_push(L, e->id);
                else idvec_push(R, e->id);
            } else idvec_push(R, e->id);
        } else if (e->type == STARTING || e->type == PLANAR) {
            idvec_push(R, e->id);
        }
    }
    for (int a = 0; a < 3; ++a) free(ev[a]);
    *out_ax = min_ax;
    *out_pos = min_pos;
    return min_cost;
}

/* lib/kdtree.cpp:128-174 */
static tnode* build_rec(const builder_t* B, uint32_t* tris, size_t ntris, box3 box) {
    if (ntris == 0) return NULL;
    if (ntris <= 3) return new_leaf(tris, ntris);
    if (box_area(box) == 0) return new_leaf(tris, ntris);
    idvec L = {0}, R = {0};
    int ax;
    float pos;
    float min_cost = find_plane(B, tris, ntris, box, &ax, &pos, &L, &R);
    if ((float)(20 * (unsigned long)ntris) * lambda_(L.n, R.n) < min_cost) {
        free(L.v);
        free(R.v);
        return new_leaf(tris, ntris);
    }
    box3 lb, rb;
    box_split(box, ax, pos, &lb, &rb);
    tnode* left = build_rec(B, L.v, L.n, lb);
    free(L.v);
    tnode* right = build_rec(B, R.v, R.n, rb);
    free(R.v);
    if (!left) return right;
    if (!right) return left;
    tnode* t = (tnode*)calloc(1, sizeof(tnode));
    t->axis = ax;
    t->split = pos;
    t->left = left;
    t->right = right;
    return t;
}

static void free_tree(tnode* t) {
    if (!t) return;
    if (t->leaf) free(t->ids);
    else {
        free_tree(t->left);
        free_tree(t->right);
    }
    free(t);
}

/* lib/kdtree.h:62-154 FlatNode encoding */
static inline uint64_t flat_inner(int axis, float split, uint32_t right) {
    uint32_t bits;
    memcpy(&bits, &split, 4);
    return ((uint64_t)bits << 32) | (uint64_t)((right << 2) | (uint32_t)axis);
}
static inline uint64_t flat_leaf(uint32_t a, uint32_t b) {
    return ((uint64_t)a << 32) | (uint64_t)((uint32_t)(b << 2) | 3u);
}

uint64_t or_flat_inner(int axis, float split, uint32_t right) { return flat_inner(axis, split, right); }
uint64_t or_flat_leaf(uint32_t a, uint32_t b) { return flat_leaf(a, b); }

typedef struct {
    uint64_t* v;
    size_t n, cap;
} nodevec;
static void nv_push(nodevec* a, uint64_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 64;
        a->v = (uint64_t*)realloc(a->v, a->cap * sizeof(uint64_t));
    }
    a->v[a->n++] = x;
}

/* lib/kdtree.cpp:420-467: DFS, left child = next node, leaves as id pairs,
 * sentinel (inner X, 0, 0) after an even leaf run. */
static void flatten_rec(const tnode* t, nodevec* out) {
    if (!t->leaf) {
        size_t me = out->n;
        nv_push(out, flat_inner(t->axis, t->split, 0x3FFFFFFFu));
        flatten_rec(t->left, out);
        uint32_t right = (uint32_t)out->n;
        out->v[me] = (out->v[me] & 0xFFFFFFFF00000000ull) | (uint64_t)((right << 2) | (uint32_t)t->axis);
        flatten_rec(t->right, out);
    } else {
        size_t i = 1;
        for (; i < t->nids; i += 2) nv_push(out, flat_leaf(t->ids[i - 1], t->ids[i]));
        if (i - 1 < t->nids) {
            nv_push(out, flat_leaf(t->ids[i - 1], 0x3FFFFFFFu));
        } else {
            nv_push(out, flat_inner(0, 0.f, 0));
        }
    }
}

/* lib/kdtree.cpp:474-490 */
size_t or_kdtree_build(const float* verts9, uint32_t n, uint64_t** nodes, float* box6) {
    uint32_t* ids = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    box3 box = tri_bbox(verts9);
    ids[0] = 0;
    for (uint32_t i = 1; i < n; ++i) {
        box3 b = tri_bbox(verts9 + 9 * (size_t)i);
        box = box_from(mk(smin(box.lo.x, b.lo.x), smin(box.lo.y, b.lo.y), smin(box.lo.z, b.lo.z)),
                       mk(smax(box.hi.x, b.hi.x), smax(box.hi.y, b.hi.y), smax(box.hi.z, b.hi.z)));
        ids[i] = i;
    }
    builder_t B = {verts9};
    tnode* root = build_rec(&B, ids, n, box);
    free(ids);
    nodevec out = {0};
    flatten_rec(root, &out);
    free_tree(root);
    box6[0] = box.lo.x; box6[1] = box.lo.y; box6[2] = box.lo.z;
    box6[3] = box.hi.x; box6[4] = box.hi.y; box6[5] = box.hi.z;
    *nodes = out.v;
    return out.n;
}

/* lib/kdtree.h:197-216 */
uint32_t or_kdtree_height(const uint64_t* nodes, size_t n_nodes) {
    (void)n_nodes;
    typedef struct { uint32_t node, level; } item;
    size_t cap = 1024, sp = 0;
    item* st = (item*)malloc(cap * sizeof(item));
    st[sp].node = 0;
    st[sp].level = 0;
    sp++;
    uint32_t h = 0;
    while (sp) {
        item it = st[--sp];
        if (it.level > h) h = it.level;
        uint64_t d = nodes[it.node];
        if ((d & 3) != 3) {
            if (sp + 2 > cap) {
                cap *= 2;
                st = (item*)realloc(st, cap * sizeof(item));
            }
            st[sp].node = it.node + 1;
            st[sp].level = it.level + 1;
            sp++;
            st[sp].node = ((uint32_t)d) >> 2;
            st[sp].level = it.level + 1;
            sp++;
        }
    }
    free(st);
    return h;
}

void or_free(void* p) { free(p); }

/* ---------------------------------------------------------------- traversal */
typedef struct {
    uint32_t node;
    float tenter, texit;
} stack_item;

typedef struct {
    uint64_t inner, leaves, refs, tests;
} counters_t;

/* lib/kdtree.cpp:580-607: test a leaf run until the first inner node / single-id leaf */
static inline uint32_t leaf_run(const float* pre16, const uint64_t* nodes, uint32_t node,
                                v3 o, v3 d, float* min_r, float* min_s, float* min_t,
                                counters_t* cnt) {
    *min_r = FLT_MAX;
    uint32_t res = OR_INVALID_ID;
    for (;; ++node) {
        uint64_t dd = nodes[node];
        if ((dd & 3) != 3) break;
        uint32_t ids[2] = {(uint32_t)(dd >> 32), ((uint32_t)dd) >> 2};
        int has_second = (uint32_t)dd != 0xFFFFFFFFu;
        for (int k = 0; k < 1 + has_second; ++k) {
            float r, s, t;
            if (cnt) { cnt->refs++; cnt->tests++; }
            if (leaf_test(pre16 + 16 * (size_t)ids[k], o, d, &r, &s, &t) && r < *min_r) {
                *min_r = r;
                *min_s = s;
                *min_t = t;
                res = ids[k];
            }
        }
        if (!has_second) break;
    }
    return res;
}

/* lib/kdtree.cpp:503-511 */
static inline v3 fix_direction(v3 d) {
    if (d.x == 0) d.x = EPS;
    if (d.y == 0) d.y = EPS;
    if (d.z == 0) d.z = EPS;
    return d;
}

/* lib/kdtree.cpp:515-578.  early_exit selects the GPU's near-first early
 * termination (stop once the best hit lies before the current leaf's exit,
 * with a 2^-16 relative margin). */
static uint32_t intersect_one(const float* pre16, const uint64_t* nodes, box3 box, v3 o,
                              v3 d, float* r, float* a, float* b, int early_exit,
                              counters_t* cnt, stack_item** stk, size_t* cap) {
    v3 fd = fix_direction(d);
    float tenter, texit;
    if (!ray_box(o, fd, box, &tenter, &texit)) return OR_INVALID_ID;
    size_t sp = 0;
    (*stk)[sp].node = 0;
    (*stk)[sp].tenter = tenter;
    (*stk)[sp].texit = texit;
    sp++;
    v3 dinv = mk(1 / fd.x, 1 / fd.y, 1 / fd.z);
    uint32_t res = OR_INVALID_ID;
    *r = FLT_MAX;
    while (sp) {
        --sp;
        uint32_t node = (*stk)[sp].node;
        tenter = (*stk)[sp].tenter;
        texit = (*stk)[sp].texit;
        for (;;) {
            uint64_t dd = nodes[node];
            if ((dd & 3) == 3) break;
            if (cnt) cnt->inner++;
            int ax = (int)(dd & 3);
            uint32_t bits = (uint32_t)(dd >> 32);
            float split;
            memcpy(&split, &bits, 4);
            float t = (split - comp(o, ax)) * comp(dinv, ax);
            uint32_t nearn = node + 1, farn = ((uint32_t)dd) >> 2;
            if (comp(fd, ax) <= 0) {
                uint32_t tmp = nearn;
                nearn = farn;
                farn = tmp;
            }
            if (texit < t) node = nearn;
            else if (t < tenter) node = farn;
            else {
                if (sp + 1 >= *cap) {
                    *cap *= 2;
                    *stk = (stack_item*)realloc(*stk, *cap * sizeof(stack_item));
                }
                (*stk)[sp].node = farn;
                (*stk)[sp].tenter = t;
                (*stk)[sp].texit = texit;
                sp++;
                node = nearn;
                texit = t;
            }
        }
        if (cnt) cnt->leaves++;
        float nr, na = 0, nb = 0;
        uint32_t next = leaf_run(pre16, nodes, node, o, d, &nr, &na, &nb, cnt);
        if (next != OR_INVALID_ID && nr < *r) {
            res = next;
            *r = nr;
            *a = na;
            *b = nb;
        }
        if (early_exit && res != OR_INVALID_ID && *r * (1.0f + 1.0f / 65536.0f) < texit) break;
    }
    return res;
}

typedef struct {
    const float* pre16;
    const uint64_t* nodes;
    box3 box;
    const float *o3, *d3;
    uint32_t *id;
    float *r, *s, *t;
    uint32_t begin, end;
    int early;
    counters_t cnt;
    int count;
    uint32_t ntris; /* brute force */
} isect_job;

static void* isect_worker(void* arg) {
    isect_job* J = (isect_job*)arg;
    size_t cap = 256;
    stack_item* stk = (stack_item*)malloc(cap * sizeof(stack_item));
    memset(&J->cnt, 0, sizeof(J->cnt));
    for (uint32_t i = J->begin; i < J->end; ++i) {
        v3 o = load3(J->o3 + 3 * (size_t)i), d = load3(J->d3 + 3 * (size_t)i);
        float r = FLT_MAX, a = 0, b = 0;
        uint32_t id = intersect_one(J->pre16, J->nodes, J->box, o, d, &r, &a, &b, J->early,
                                    J->count ? &J->cnt : NULL, &stk, &cap);
        if (J->id) {
            J->id[i] = id;
            J->r[i] = id == OR_INVALID_ID ? FLT_MAX : r;
            J->s[i] = id == OR_INVALID_ID ? 0.f : a;
            J->t[i] = id == OR_INVALID_ID ? 0.f : b;
        }
    }
    free(stk);
    return NULL;
}

static void* brute_worker(void* arg) {
    isect_job* J = (isect_job*)arg;
    for (uint32_t i = J->begin; i < J->end; ++i) {
        v3 o = load3(J->o3 + 3 * (size_t)i), d = load3(J->d3 + 3 * (size_t)i);
        float best = FLT_MAX, bs = 0, bt = 0;
        uint32_t bid = OR_INVALID_ID;
        for (uint32_t k = 0; k < J->ntris; ++k) {
            float r, s, t;
            if (leaf_test(J->pre16 + 16 * (size_t)k, o, d, &r, &s, &t) && r < best) {
                best = r;
                bs = s;
                bt = t;
                bid = k;
            }
        }
        J->id[i] = bid;
        J->r[i] = best;
        J->s[i] = bid == OR_INVALID_ID ? 0.f : bs;
        J->t[i] = bid == OR_INVALID_ID ? 0.f : bt;
    }
    return NULL;
}

static void run_jobs(isect_job* proto, uint32_t nrays, int nthreads, void* (*fn)(void*),
                     counters_t* total) {
    if (nthreads < 1) nthreads = 1;
    if ((uint32_t)nthreads > nrays) nthreads = nrays ? (int)nrays : 1;
    isect_job* jobs = (isect_job*)calloc((size_t)nthreads, sizeof(isect_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int k = 0; k < nthreads; ++k) {
        jobs[k] = *proto;
        jobs[k].begin = (uint32_t)((uint64_t)nrays * k / nthreads);
        jobs[k].end = (uint32_t)((uint64_t)nrays * (k + 1) / nthreads);
        if (nthreads == 1) fn(&jobs[k]);
        else pthread_create(&th[k], NULL, fn, &jobs[k]);
    }
    if (total) memset(total, 0, sizeof(*total));
    for (int k = 0; k < nthreads; ++k) {
        if (nthreads > 1) pthread_join(th[k], NULL);
        if (total) {
            total->inner += jobs[k].cnt.inner;
            total->leaves += jobs[k].cnt.leaves;
            total->refs += jobs[k].cnt.refs;
            total->tests += jobs[k].cnt.tests;
        }
    }
    free(jobs);
    free(th);
}

void or_intersect(const float* pre16, const uint64_t* nodes, const float* box6,
                  const float* o3, const float* d3, uint32_t nrays, uint32_t* id,
                  float* r, float* s, float* t, int nthreads) {
    isect_job P;
    memset(&P, 0, sizeof(P));
    P.pre16 = pre16;
    P.nodes = nodes;
    P.box.lo = load3(box6);
    P.box.hi = load3(box6 + 3);
    P.o3 = o3;
    P.d3 = d3;
    P.id = id;
    P.r = r;
    P.s = s;
    P.t = t;
    run_jobs(&P, nrays, nthreads, isect_worker, NULL);
}

void or_intersect_count(const float* pre16, const uint64_t* nodes, const float* box6,
                        const float* o3, const float* d3, uint32_t nrays,
                        uint64_t* counters4, int nthreads) {
    isect_job P;
    memset(&P, 0, sizeof(P));
    P.pre16 = pre16;
    P.nodes = nodes;
    P.box.lo = load3(box6);
    P.box.hi = load3(box6 + 3);
    P.o3 = o3;
    P.d3 = d3;
    P.early = 1;
    P.count = 1;
    counters_t tot;
    run_jobs(&P, nrays, nthreads, isect_worker, &tot);
    counters4[0] = tot.inner;
    counters4[1] = tot.leaves;
    counters4[2] = tot.refs;
    counters4[3] = tot.tests;
}

void or_intersect_brute(const float* pre16, uint32_t ntris, const float* o3,
                        const float* d3, uint32_t nrays, uint32_t* id, float* r,
                        float* s, float* t, int nthreads) {
    isect_job P;
    memset(&P, 0, sizeof(P));
    P.pre16 = pre16;
    P.ntris = ntris;
    P.o3 = o3;
    P.d3 = d3;
    P.id = id;
    P.r = r;
    P.s = s;
    P.t = t;
    run_jobs(&P, nrays, nthreads, brute_worker, NULL);
}

/* Moeller & Trumbore 1997, "Fast, minimum storage ray/triangle intersection",
 * evaluated in double for the given triangle (distance in units of |d|). */
void or_mt64(const float* verts9, const uint32_t* id, const float* o3, const float* d3,
             uint32_t nrays, double* r, double* b1, double* b2) {
    for (uint32_t i = 0; i < nrays; ++i) {
        if (id[i] == OR_INVALID_ID) {
            r[i] = b1[i] = b2[i] = NAN;
            continue;
        }
        const float* t = verts9 + 9 * (size_t)id[i];
        double v0[3] = {t[0], t[1], t[2]}, e1[3], e2[3], o[3], d[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = (double)t[3 + k] - v0[k];
            e2[k] = (double)t[6 + k] - v0[k];
            o[k] = o3[3 * (size_t)i + k];
            d[k] = d3[3 * (size_t)i + k];
        }
        double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                       d[0] * e2[1] - d[1] * e2[0]};
        double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
        double inv = 1.0 / det;
        double tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
        double u = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) * inv;
        double q[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2],
                       tv[0] * e1[1] - tv[1] * e1[0]};
        double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
        double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
        r[i] = tt;
        b1[i] = u;
        b2[i] = v;
    }
}

/* ---------------------------------------------------------------- RNG */
/* lib/xorshift.h:36-41 (xorshift64*, Vigna / Marsaglia) */
uint64_t or_xorshift64s_next(uint64_t* s) {
    *s ^= *s >> 12;
    *s ^= *s << 25;
    *s ^= *s >> 27;
    return *s * 2685821657736338717ULL;
}
/* lib/xorshift.h:57-62: low 24 bits * 2^-24 */
float or_xorshift64s_float(uint64_t* s) {
    return ldexpf((float)(or_xorshift64s_next(s) & 0xFFFFFFull), -24);
}

/* splitmix64 (Steele, Lea, Flood 2014) to derive per-row seeds */
static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* ---------------------------------------------------------------- camera */
/* lib/types.h:119-123: d = trafo * (-dx(1 - 2px/w), dy(1 - 2py/h), -1) */
void or_raster2cam(const or_camera* c, float px, float py, float w, float h, float* dir) {
    float x = -c->delta_x * (1 - 2 * px / w);
    float y = c->delta_y * (1 - 2 * py / h);
    float z = -1;
    const float* m = c->trafo;
    dir[0] = m[0] * x + m[1] * y + m[2] * z;
    dir[1] = m[3] * x + m[4] * y + m[5] * z;
    dir[2] = m[6] * x + m[7] * y + m[8] * z;
}

/* ---------------------------------------------------------------- trace */
typedef struct {
    float r, g, b, a;
} col4;
static inline col4 c4(float r, float g, float b, float a) {
    col4 c = {r, g, b, a};
    return c;
}
static inline col4 cadd(col4 x, col4 y) { return c4(x.r + y.r, x.g + y.g, x.b + y.b, x.a + y.a); }
static inline col4 cmul(col4 x, col4 y) { return c4(x.r * y.r, x.g * y.g, x.b * y.b, x.a * y.a); }
static inline col4 cscale(float s, col4 x) { return c4(s * x.r, s * x.g, s * x.b, s * x.a); }

typedef struct {
    const or_scene* sc;
    const or_light* lights;
    uint32_t nlights;
    const or_render_params* prm;
    const float* light_cdf; /* cumulative area over emitters */
    const uint32_t* light_ids;
    uint32_t nemit;
    float light_area;
    stack_item* stk;
    size_t cap;
    uint64_t nrays, nshadow;
    uint64_t rng;
} tracer;

static inline float uni(tracer* T) { return or_xorshift64s_float(&T->rng); }

static uint32_t closest(tracer* T, v3 o, v3 d, float* r, float* s, float* t) {
    const or_scene* sc = T->sc;
    box3 b;
    b.lo = load3(sc->box);
    b.hi = load3(sc->box + 3);
    *r = FLT_MAX;
    return intersect_one(sc->pre16, sc->nodes, b, o, d, r, s, t, 0, NULL, &T->stk, &T->cap);
}

/* Moeller & Hughes 1999, "Efficiently building a matrix to rotate one vector
 * to another" -- the construction assimp's aiMatrix3x3::FromToMatrix names
 * (pathtracer.cpp:68-70).  Only the induced direction distribution matters
 * for the checks that use it (any rotation z -> n gives the same one). */
static void from_to(v3 f, v3 t, float m[9]) {
    const float e = 0.00001f;
    v3 v = mk(f.y * t.z - f.z * t.y, f.z * t.x - f.x * t.z, f.x * t.y - f.y * t.x);
    float c = dot(f, t);
    float fc = fabsf(c);
    if (fc > 1.0f - e) {
        v3 x = mk(fabsf(f.x), fabsf(f.y), fabsf(f.z));
        if (x.x < x.y) {
            if (x.x < x.z) x = mk(1, 0, 0);
            else x = mk(0, 0, 1);
        } else {
            if (x.y < x.z) x = mk(0, 1, 0);
            else x = mk(0, 0, 1);
        }
        v3 u = sub(x, f), w = sub(x, t);
        float c1 = 2.0f / dot(u, u), c2 = 2.0f / dot(w, w), c3 = c1 * c2 * dot(u, w);
        float uu[3] = {u.x, u.y, u.z}, ww[3] = {w.x, w.y, w.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                m[3 * i + j] = -c1 * uu[i] * uu[j] - c2 * ww[i] * ww[j] + c3 * ww[i] * uu[j] +
                               (i == j ? 1.f : 0.f);
        return;
    }
    float h = 1.0f / (1.0f + c);
    float hvx = h * v.x, hvz = h * v.z, hvxy = hvx * v.y, hvxz = hvx * v.z, hvyz = hvz * v.y;
    m[0] = c + hvx * v.x; m[1] = hvxy - v.z; m[2] = hvxz + v.y;
    m[3] = hvxy + v.z; m[4] = c + h * v.y * v.y; m[5] = hvyz - v.x;
    m[6] = hvxz - v.y; m[7] = hvyz + v.x; m[8] = c + hvz * v.z;
}

/* lib/triangle.h:54-56 */
static inline v3 interp_normal(const float* n9, float a, float b, float c) {
    v3 n0 = load3(n9), n1 = load3(n9 + 3), n2 = load3(n9 + 6);
    return normalize(add(add(scale(a, n0), scale(b, n1)), scale(c, n2)));
}

/* Direct light from emissive triangles (no reference counterpart): one
 * emitter picked with probability area/total, uniform point on it. */
static col4 area_light(tracer* T, v3 p, v3 p2, v3 n) {
    const or_scene* sc = T->sc;
    float u0 = uni(T), u1 = uni(T), u2 = uni(T);
    float x = u0 * T->light_area;
    uint32_t lo = 0, hi = T->nemit - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (T->light_cdf[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    uint32_t k = T->light_ids[lo];
    const float* tv = sc->verts9 + 9 * (size_t)k;
    float su = sqrtf(u1);
    float b0 = 1.0f - su, b1 = su * (1.0f - u2), b2 = su * u2;
    v3 q = add(add(scale(b0, load3(tv)), scale(b1, load3(tv + 3))), scale(b2, load3(tv + 6)));
    v3 dl = sub(q, p);
    float dist2 = dot(dl, dl);
    float dist = sqrtf(dist2);
    v3 wi = scale(1.0f / dist, dl);
    float cosx = dot(wi, n);
    v3 nl = load3(sc->pre16 + 16 * (size_t)k + 3);
    float cosl = -dot(wi, nl);
    if (cosx <= 0 || cosl <= 0) return c4(0, 0, 0, 0);
    T->nshadow++;
    float tmax = length(sub(q, p2)) * (1.0f - 1e-4f);
    float r, s, t;
    uint32_t hid = closest(T, p2, wi, &r, &s, &t);
    if (hid != OR_INVALID_ID && r < tmax) return c4(0, 0, 0, 0);
    const float* le = sc->emission4 + 4 * (size_t)k;
    float g = cosx * cosl / dist2 * T->light_area;
    return c4(le[0] * g, le[1] * g, le[2] * g, le[3] * g);
}

/* pathtracer.cpp:14-102 (R8, R9).  With OR_FLAG_COSINE the hemisphere is
 * sampled by cosine and the estimator weight becomes rho instead of
 * 2*cos*rho; with OR_FLAG_AREA_LIGHTS emissive triangles add camera-visible
 * emission at depth 0 and next-event estimation at every hit. */
static col4 trace(tracer* T, v3 o, v3 d, int depth) {
    const or_render_params* P = T->prm;
    if (depth > P->max_depth) return c4(0, 0, 0, 0);
    T->nrays++;
    float dist, s, t;
    uint32_t id = closest(T, o, d, &dist, &s, &t);
    if (id == OR_INVALID_ID) return c4(P->bg[0], P->bg[1], P->bg[2], P->bg[3]);
    const or_scene* sc = T->sc;
    v3 p = add(o, scale(dist, d));
    v3 n = interp_normal(sc->normals9 + 9 * (size_t)id, 1.f - s - t, s, t);
    v3 p2 = add(p, scale(0.0001f, n));
    col4 direct = c4(0, 0, 0, 0);
    for (uint32_t li = 0; li < T->nlights; ++li) {
        const or_light* L = &T->lights[li];
        v3 lp = load3(L->pos);
        v3 ldir = normalize(sub(lp, p));
        float dl = length(sub(lp, p2));
        float dn, ss, tt;
        T->nshadow++;
        uint32_t sh = closest(T, p2, ldir, &dn, &ss, &tt);
        if (sh == OR_INVALID_ID || dn > dl) {
            float c = fmaxf(0.f, dot(ldir, n));
            direct = c4(c * L->color[0], c * L->color[1], c * L->color[2], c * L->color[3]);
        }
    }
    col4 emitted = c4(0, 0, 0, 0);
    if ((P->flags & OR_FLAG_AREA_LIGHTS) && T->nemit) {
        direct = cadd(direct, area_light(T, p, p2, n));
        const float* le = sc->emission4 + 4 * (size_t)id;
        v3 ng = load3(sc->pre16 + 16 * (size_t)id + 3);
        if (depth == 0 && (le[0] > 0 || le[1] > 0 || le[2] > 0) && dot(ng, d) < 0)
            emitted = c4(le[0], le[1], le[2], le[3]);
    }
    col4 ind = c4(0, 0, 0, 0);
    int cosine = (P->flags & OR_FLAG_COSINE) != 0;
    float m[9];
    from_to(mk(0, 0, 1), n, m);
    for (int run = 0; run < P->mc_samples; ++run) {
        float u1 = uni(T), u2 = uni(T);
        float x, y, z, w;
        const float two_pi = (float)(2.0 * M_PI);
        if (cosine) {
            float rr = sqrtf(u1);
            float phi = two_pi * u2;
            x = rr * cosf(phi);
            y = rr * sinf(phi);
            z = sqrtf(fmaxf(0.f, 1.f - u1));
            w = 1.f;
        } else { /* lib/sampling.h:20-32 */
            z = u1;
            float rr = sqrtf(fmaxf(0.f, 1.f - z * z));
            float phi = two_pi * u2;
            x = rr * cosf(phi);
            y = rr * sinf(phi);
            w = u1; /* cos(theta); the factor 2 = 1/(2 pi pdf) * pi is applied below */
        }
        v3 dir = mk(m[0] * x + m[1] * y + m[2] * z, m[3] * x + m[4] * y + m[5] * z,
                    m[6] * x + m[7] * y + m[8] * z);
        col4 Li = trace(T, p2, dir, depth + 1);
        ind = cadd(ind, cscale(w, Li));
    }
    if (P->mc_samples > 0) { /* pathtracer.cpp:88 */
        float fm = (float)P->mc_samples;
        ind = c4(ind.r / fm, ind.g / fm, ind.b / fm, ind.a / fm);
    }
    const float* alb = sc->albedo4 + 4 * (size_t)id;
    col4 rho = c4(alb[0], alb[1], alb[2], alb[3]);
    /* pathtracer.cpp:100-101: diffuse * (direct * 1/pi + indirect * 2) */
    col4 res = cmul(rho, cadd(cscale((float)M_1_PI, direct), cscale(cosine ? 1.f : 2.f, ind)));
    return cadd(res, emitted);
}

/* raycaster.cpp:7-24: the hit triangle's diffuse colour, alpha fading with
 * the ray parameter (clamp of src/turner.h:28-36); num_rays counts hits. */
static col4 raycast(tracer* T, v3 o, v3 d) {
    const or_render_params* P = T->prm;
    float dist, s, t;
    uint32_t id = closest(T, o, d, &dist, &s, &t);
    if (id == OR_INVALID_ID) return c4(P->bg[0], P->bg[1], P->bg[2], P->bg[3]);
    T->nrays++;
    const float* alb = T->sc->albedo4 + 4 * (size_t)id;
    float a = 1.f - dist / P->max_visibility;
    a = a < 0.f ? 0.f : (a > 1.f ? 1.f : a);
    return c4(alb[0], alb[1], alb[2], a);
}

/* raytracer.cpp:6-67: Lambert term of the (first) point light
 * (lib/lambertian.h:15-19), mirror reflection weighted by reflectivity and
 * the reflective colour, and darkening by the shadow intensity when the
 * closest hit toward the light from p2 is nearer than the light. */
static col4 raytrace(tracer* T, v3 o, v3 d, int depth) {
    const or_render_params* P = T->prm;
    T->nrays++;
    if (depth > P->max_depth) return c4(0, 0, 0, 0);
    const or_light* L = &T->lights[0];
    float dist, s, t;
    uint32_t id = closest(T, o, d, &dist, &s, &t);
    if (id == OR_INVALID_ID) return c4(P->bg[0], P->bg[1], P->bg[2], P->bg[3]);
    const or_scene* sc = T->sc;
    v3 p = add(o, scale(dist, d));
    v3 lp = load3(L->pos);
    v3 ldir = normalize(sub(lp, p));
    v3 n = interp_normal(sc->normals9 + 9 * (size_t)id, 1.f - s - t, s, t);
    const float* alb = sc->albedo4 + 4 * (size_t)id;
    float ca = fmaxf(0.f, dot(ldir, n));
    col4 direct = cmul(cscale(ca, c4(alb[0], alb[1], alb[2], alb[3])),
                       c4(L->color[0], L->color[1], L->color[2], L->color[3]));
    v3 p2 = add(p, scale(0.0001f, n));
    col4 color = direct;
    float refl = sc->reflectivity ? sc->reflectivity[id] : 0.f;
    if (refl > 0) {
        v3 rd = sub(d, scale(2.f * dot(n, d), n));
        col4 rc = raytrace(T, p2, rd, depth + 1);
        const float* R = sc->reflective4 + 4 * (size_t)id;
        color = cadd(cscale(1.f - refl, direct), cmul(cscale(refl, c4(R[0], R[1], R[2], R[3])), rc));
    }
    v3 ld2 = normalize(sub(lp, p2));
    float dl = length(sub(lp, p2));
    float dn, ss, tt;
    T->nshadow++;
    uint32_t sh = closest(T, p2, ld2, &dn, &ss, &tt);
    if (sh != OR_INVALID_ID && dn < dl) {
        float si = P->shadow_intensity;
        color = c4(color.r - color.r * si, color.g - color.g * si, color.b - color.b * si, color.a - color.a * si);
    }
    return color;
}

/* the renderer's trace() (trace.h:23-25) for a camera ray */
static col4 trace_plugin(tracer* T, v3 o, v3 d) {
    if (T->prm->renderer == OR_RENDERER_RAYCASTER) return raycast(T, o, d);
    if (T->prm->renderer == OR_RENDERER_RAYTRACER) return raytrace(T, o, d, 0);
    return trace(T, o, d, 0);
}

typedef struct {
    tracer proto;
    const or_camera* cam;
    float* out;
    float* out_sq;
    uint32_t y0, y1, ystep;
    uint64_t nrays, nshadow;
} render_job;

static void render_rows(render_job* J, int ref_single_stream) {
    tracer T = J->proto;
    T.cap = 256;
    T.stk = (stack_item*)malloc(T.cap * sizeof(stack_item));
    T.nrays = T.nshadow = 0;
    const or_render_params* P = T.prm;
    uint32_t w = P->width, h = P->height;
    if (ref_single_stream) T.rng = 4; /* lib/sampling.h:12 */
    for (uint32_t y = J->y0; y < J->y1; y += J->ystep) {
        uint64_t jit = 42; /* main.cpp:201 */
        if (!ref_single_stream) {
            T.rng = splitmix64(P->seed * 0x100000001B3ull + y) | 1;
            jit = splitmix64(P->seed ^ (0xA5A5A5A5ull + y)) | 1;
            if (P->flags & OR_FLAG_REF_RNG) jit = 42;
        }
        for (uint32_t x = 0; x < w; ++x) {
            col4 acc = c4(0, 0, 0, 0), acc2 = c4(0, 0, 0, 0);
            for (int i = 0; i < P->pixel_samples; ++i) {
                float dx = or_xorshift64s_float(&jit);
                float dy = or_xorshift64s_float(&jit);
                float dir[3];
                or_raster2cam(J->cam, (float)x + dx, (float)y + dy, (float)w, (float)h, dir);
                col4 L = trace_plugin(&T, load3(J->cam->pos), load3(dir));
                acc = cadd(acc, L);
                acc2 = cadd(acc2, cmul(L, L));
            }
            float inv = 1.0f / (float)P->pixel_samples;
            float* o = J->out + 4 * ((size_t)y * w + x);
            o[0] = acc.r * inv; o[1] = acc.g * inv; o[2] = acc.b * inv; o[3] = acc.a * inv;
            if (J->out_sq) {
                float* q = J->out_sq + 4 * ((size_t)y * w + x);
                q[0] = acc2.r; q[1] = acc2.g; q[2] = acc2.b; q[3] = acc2.a;
            }
        }
    }
    free(T.stk);
    J->nrays = T.nrays;
    J->nshadow = T.nshadow;
}

static void* render_worker(void* arg) {
    render_rows((render_job*)arg, 0);
    return NULL;
}

/* Emitter table of a scene (triangles with emission > 0): ids and the
 * inclusive prefix of their areas, for area-light sampling.  Fills T and
 * returns 0, or -1 when allocation fails. */
static int tracer_setup(tracer* T, const or_scene* sc, const or_light* lights, uint32_t nlights,
                        const or_render_params* prm) {
    memset(T, 0, sizeof(*T));
    T->sc = sc;
    T->lights = lights;
    T->nlights = nlights;
    T->prm = prm;
    uint32_t ne = 0;
    for (uint32_t i = 0; i < sc->ntris; ++i) {
        const float* le = sc->emission4 ? sc->emission4 + 4 * (size_t)i : NULL;
        if (le && (le[0] > 0 || le[1] > 0 || le[2] > 0)) ne++;
    }
    float* cdf = (float*)malloc((ne ? ne : 1) * sizeof(float));
    uint32_t* lid = (uint32_t*)malloc((ne ? ne : 1) * sizeof(uint32_t));
    if (!cdf || !lid) {
        free(cdf);
        free(lid);
        return -1;
    }
    double acc = 0;
    uint32_t k = 0;
    for (uint32_t i = 0; i < sc->ntris && ne; ++i) {
        const float* le = sc->emission4 + 4 * (size_t)i;
        if (!(le[0] > 0 || le[1] > 0 || le[2] > 0)) continue;
        const float* p = sc->pre16 + 16 * (size_t)i;
        v3 c = cross_d(load3(p + 6), load3(p + 9));
        acc += 0.5 * sqrt((double)c.x * c.x + (double)c.y * c.y + (double)c.z * c.z);
        cdf[k] = (float)acc;
        lid[k] = i;
        k++;
    }
    T->light_cdf = cdf;
    T->light_ids = lid;
    T->nemit = ne;
    T->light_area = (float)acc;
    return 0;
}

static void tracer_release(tracer* T) {
    free((void*)T->light_cdf);
    free((void*)T->light_ids);
}

void or_render(const or_scene* sc, const or_camera* cam, const or_light* lights,
               uint32_t nlights, const or_render_params* prm, float* out, float* out_sq,
               uint64_t* n_rays, uint64_t* n_shadow) {
    tracer T;
    memset(out, 0, sizeof(float) * 4 * (size_t)prm->width * prm->height);
    if (out_sq) memset(out_sq, 0, sizeof(float) * 4 * (size_t)prm->width * prm->height);
    if (n_rays) *n_rays = 0;
    if (n_shadow) *n_shadow = 0;
    if (tracer_setup(&T, sc, lights, nlights, prm)) return;
    uint32_t step = prm->row_step ? prm->row_step : 1;
    int single = (prm->flags & OR_FLAG_REF_RNG) && prm->nthreads <= 1;
    int nt = single ? 1 : (prm->nthreads < 1 ? 1 : prm->nthreads);
    render_job* jobs = (render_job*)calloc((size_t)nt, sizeof(render_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nt, sizeof(pthread_t));
    for (int j = 0; j < nt; ++j) {
        jobs[j].proto = T;
        jobs[j].cam = cam;
        jobs[j].out = out;
        jobs[j].out_sq = out_sq;
        jobs[j].y0 = prm->row_offset + (uint32_t)j * step;
        jobs[j].y1 = prm->height;
        jobs[j].ystep = step * (uint32_t)nt;
        if (single) render_rows(&jobs[j], 1);
        else pthread_create(&th[j], NULL, render_worker, &jobs[j]);
    }
    uint64_t nr = 0, ns = 0;
    for (int j = 0; j < nt; ++j) {
        if (!single) pthread_join(th[j], NULL);
        nr += jobs[j].nrays;
        ns += jobs[j].nshadow;
    }
    if (n_rays) *n_rays = nr;
    if (n_shadow) *n_shadow = ns;
    free(jobs);
    free(th);
    tracer_release(&T);
}

/* trace(ray, ..., depth, conf) (trace.h:23-25, pathtracer.cpp:14-102) for
 * caller rays: ray j draws its own xorshift64* stream seeded from
 * (seed, j).  Rays [i0, i1) of one worker. */
typedef struct {
    tracer proto;
    const float *o3, *d3;
    int32_t depth;
    float* out4;
    uint32_t i0, i1;
    uint64_t nrays, nshadow;
} trace_job;

static void* trace_worker(void* arg) {
    trace_job* J = (trace_job*)arg;
    tracer T = J->proto;
    T.cap = 256;
    T.stk = (stack_item*)malloc(T.cap * sizeof(stack_item));
    T.nrays = T.nshadow = 0;
    for (uint32_t j = J->i0; j < J->i1; ++j) {
        T.rng = splitmix64(T.prm->seed * 0x100000001B3ull + j) | 1;
        v3 o = load3(J->o3 + 3 * (size_t)j), d = load3(J->d3 + 3 * (size_t)j);
        col4 L;
        if (T.prm->renderer == OR_RENDERER_RAYCASTER) L = raycast(&T, o, d);
        else if (T.prm->renderer == OR_RENDERER_RAYTRACER) L = raytrace(&T, o, d, J->depth);
        else L = trace(&T, o, d, J->depth);
        float* q = J->out4 + 4 * (size_t)j;
        q[0] = L.r; q[1] = L.g; q[2] = L.b; q[3] = L.a;
    }
    free(T.stk);
    J->nrays = T.nrays;
    J->nshadow = T.nshadow;
    return NULL;
}

void or_trace_rays(const or_scene* sc, const or_light* lights, uint32_t nlights,
                   const or_render_params* prm, const float* o3, const float* d3, uint32_t n,
                   int32_t depth, float* out4, uint64_t* n_rays, uint64_t* n_shadow) {
    tracer T;
    if (n_rays) *n_rays = 0;
    if (n_shadow) *n_shadow = 0;
    if (n == 0 || tracer_setup(&T, sc, lights, nlights, prm)) return;
    int nt = prm->nthreads < 1 ? 1 : prm->nthreads;
    if ((uint32_t)nt > n) nt = (int)n;
    trace_job* jobs = (trace_job*)calloc((size_t)nt, sizeof(trace_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nt, sizeof(pthread_t));
    for (int j = 0; j < nt; ++j) {
        jobs[j].proto = T;
        jobs[j].o3 = o3;
        jobs[j].d3 = d3;
        jobs[j].depth = depth;
        jobs[j].out4 = out4;
        jobs[j].i0 = (uint32_t)((uint64_t)n * (uint64_t)j / (uint64_t)nt);
        jobs[j].i1 = (uint32_t)((uint64_t)n * (uint64_t)(j + 1) / (uint64_t)nt);
        pthread_create(&th[j], NULL, trace_worker, &jobs[j]);
    }
    uint64_t nr = 0, ns = 0;
    for (int j = 0; j < nt; ++j) {
        pthread_join(th[j], NULL);
        nr += jobs[j].nrays;
        ns += jobs[j].nshadow;
    }
    if (n_rays) *n_rays = nr;
    if (n_shadow) *n_shadow = ns;
    free(jobs);
    free(th);
    tracer_release(&T);
}

/* ---------------------------------------------------------------- output */
/* lib/effects.h:15-17 */
float or_exposure(float light, float value) { return 1 - expf(-light * value); }
/* lib/effects.h:35-37 */
float or_gamma(float light, float inverse_gamma) { return powf(light, inverse_gamma); }

/* lib/types.h:26-34 clamp to [0, 255] */
static float clamp255(float x) {
    if (x < 0) return 0;
    if (255 < x) return 255;
    return x;
}

/* lib/raster.h:79-100: P3, width height, 255, rows of "%3d %3d %3d" of
 * int(clamp(255*c*a)); pixels separated by one space, rows by newline. */
size_t or_ppm(const float* rgba, uint32_t w, uint32_t h, char* buf, size_t cap) {
    size_t len = 0;
#define PUT(...)                                                              \
    do {                                                                      \
        int k_ = snprintf(buf + len, len < cap ? cap - len : 0, __VA_ARGS__); \
        len += (size_t)k_;                                                    \
    } while (0)
    PUT("P3\n%u %u\n255", w, h);
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i) {
        if (i % w == 0) PUT("\n");
        else PUT(" ");
        const float* c = rgba + 4 * i;
        PUT("%3d %3d %3d", (int)clamp255(255 * c[0] * c[3]), (int)clamp255(255 * c[1] * c[3]),
            (int)clamp255(255 * c[2] * c[3]));
    }
#undef PUT
    return len;
}
