// kdtree_build.cpp -- host SAH kd-tree builder of libturner_amd.
//
// Builds the same tree as the reference's KDTree::KDTree (lib/kdtree.cpp:474-490):
// Wald & Havran 2006 Algorithm 4 with the reference's cost constants, its
// clipping of triangles at thick box planes, its event order and tie rules,
// and its flattening.  The arithmetic follows the reference operation by
// operation (compiled without FMA contraction) so split positions, leaf
// contents and leaf order are identical; std::sort with the reference's
// comparator keeps even the order of equal events.  Differences from the
// reference are in the machinery only: clipping uses fixed-size arrays
// instead of a heap vector per plane, and independent subtrees are built on
// parallel host threads (the result does not depend on the thread count).
#include <algorithm>
#include <functional>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <future>
#include <limits>
#include <memory>
#include <thread>

#include "tn_host.h"

namespace tn {
namespace {

constexpr float kEps = 0.00001f;  // lib/types.h:14

struct V3 {
    float x, y, z;
};
inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
inline float at(const V3& a, int ax) { return ax == 0 ? a.x : (ax == 1 ? a.y : a.z); }
inline void set(V3& a, int ax, float v) { (ax == 0 ? a.x : (ax == 1 ? a.y : a.z)) = v; }
inline V3 operator-(const V3& a, const V3& b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator+(const V3& a, const V3& b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator*(float s, const V3& v) { return mk(s * v.x, s * v.y, s * v.z); }
inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 load(const float* p) { return mk(p[0], p[1], p[2]); }

struct Box {
    V3 lo, hi;
};
// src/geometry.h:998-999 (std::min / std::max per component)
inline Box box_of(const V3& p1, const V3& p2) {
    return Box{mk(std::min(p1.x, p2.x), std::min(p1.y, p2.y), std::min(p1.z, p2.z)),
               mk(std::max(p1.x, p2.x), std::max(p1.y, p2.y), std::max(p1.z, p2.z))};
}
inline Box box_default() {
    const float lo = std::numeric_limits<float>::lowest(), hi = std::numeric_limits<float>::max();
    return Box{mk(lo, lo, lo), mk(hi, hi, hi)};
}
inline bool empty(const Box& b) { return b.hi.x <= b.lo.x && b.hi.y <= b.lo.y && b.hi.z <= b.lo.z; }
inline bool planar(const Box& b, int ax) { return std::abs(at(b.hi, ax) - at(b.lo, ax)) < kEps; }
inline float area(const Box& b) {
    V3 d = b.hi - b.lo;
    return 2 * (d.x * d.y + d.x * d.z + d.y * d.z);
}
inline void split(const Box& b, int ax, float pos, Box& l, Box& r) {
    V3 lmax = b.hi;
    set(lmax, ax, pos);
    V3 rmin = b.lo;
    set(rmin, ax, pos);
    l = box_of(b.lo, lmax);
    r = box_of(rmin, b.hi);
}
// b overlaps box enlarged by tol on every side
inline bool near_box(const Box& b, const Box& box, float tol = kEps) {
    return !(b.hi.x < box.lo.x - tol || b.hi.y < box.lo.y - tol || b.hi.z < box.lo.z - tol ||
             box.hi.x + tol < b.lo.x || box.hi.y + tol < b.lo.y || box.hi.z + tol < b.lo.z);
}
// b clamped into box; an axis where b lies outside collapses onto box's nearer face
inline Box clamp_to(const Box& b, const Box& box) {
    Box r;
    for (int ax = 0; ax < 3; ++ax) {
        float lo = std::max(at(b.lo, ax), at(box.lo, ax)), hi = std::min(at(b.hi, ax), at(box.hi, ax));
        if (hi < lo) {
            const float face = at(b.hi, ax) < at(box.lo, ax) ? at(box.lo, ax) : at(box.hi, ax);
            lo = hi = face;
        }
        set(r.lo, ax, lo);
        set(r.hi, ax, hi);
    }
    return r;
}
inline float fmin3(float x, float y, float z) { return std::fmin(x, std::min(y, z)); }
inline float fmax3(float x, float y, float z) { return std::fmax(x, std::max(y, z)); }
inline Box tri_box(const float* t) {
    return box_of(mk(fmin3(t[0], t[3], t[6]), fmin3(t[1], t[4], t[7]), fmin3(t[2], t[5], t[8])),
                  mk(fmax3(t[0], t[3], t[6]), fmax3(t[1], t[4], t[7]), fmax3(t[2], t[5], t[8])));
}

// Triangle t meets box enlarged by tol on every side: the separating-axis
// test of Akenine-Moeller, "Fast 3D triangle-box overlap testing" (2001) --
// the box's three face normals, the triangle's normal and the nine cross
// products of edges with the coordinate axes -- in FP64 with <= on touching,
// so a degenerate triangle (zero edges or normal) reduces to its bounding box.
inline bool tri_meets_box(const float* t, const Box& box, double tol) {
    double c[3], h[3], v[3][3];
    for (int ax = 0; ax < 3; ++ax) {
        const double lo = at(box.lo, ax), hi = at(box.hi, ax);
        c[ax] = 0.5 * (lo + hi);
        h[ax] = 0.5 * (hi - lo) + tol;
        for (int k = 0; k < 3; ++k) v[k][ax] = (double)t[3 * k + ax] - c[ax];
    }
    for (int ax = 0; ax < 3; ++ax) {
        const double lo = std::min(v[0][ax], std::min(v[1][ax], v[2][ax]));
        const double hi = std::max(v[0][ax], std::max(v[1][ax], v[2][ax]));
        if (lo > h[ax] || hi < -h[ax]) return false;
    }
    auto sep = [&](const double a[3]) {
        double p0 = a[0] * v[0][0] + a[1] * v[0][1] + a[2] * v[0][2];
        double p1 = a[0] * v[1][0] + a[1] * v[1][1] + a[2] * v[1][2];
        double p2 = a[0] * v[2][0] + a[1] * v[2][1] + a[2] * v[2][2];
        const double rad = h[0] * std::abs(a[0]) + h[1] * std::abs(a[1]) + h[2] * std::abs(a[2]);
        return std::min(p0, std::min(p1, p2)) > rad || std::max(p0, std::max(p1, p2)) < -rad;
    };
    double e[3][3];
    for (int k = 0; k < 3; ++k)
        for (int ax = 0; ax < 3; ++ax) e[k][ax] = v[(k + 1) % 3][ax] - v[k][ax];
    for (int k = 0; k < 3; ++k) {
        const double a0[3] = {0.0, -e[k][2], e[k][1]};  // x cross e
        const double a1[3] = {e[k][2], 0.0, -e[k][0]};  // y cross e
        const double a2[3] = {-e[k][1], e[k][0], 0.0};  // z cross e
        if (sep(a0) || sep(a1) || sep(a2)) return false;
    }
    const double n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                         e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    return !sep(n);
}

// Thick-plane Sutherland-Hodgman against the 6 box planes (clipping.h:121-235,
// intersection.h:17-22), with the reference's full three-term dot products.
inline int clip_plane(const V3* in, int nin, V3* out, const V3& n, float d) {
    auto side = [&](const V3& p) {
        float dist = dot(n, p) - d;
        return dist > kEps ? 2 : (dist < -kEps ? 1 : 0);
    };
    int nout = 0;
    V3 a = in[nin - 1];
    int as = side(a);
    for (int i = 0; i < nin; ++i) {
        const V3 b = in[i];
        const int bs = side(b);
        if (bs == 2) {
            if (as == 1) {
                V3 ab = b - a;
                float t = (d - dot(n, a)) / dot(n, ab);
                out[nout++] = a + t * (b - a);
            }
            out[nout++] = b;
        } else if (bs == 1) {
            if (as == 2) {
                V3 ab = b - a;
                float t = (d - dot(n, a)) / dot(n, ab);
                out[nout++] = a + t * (b - a);
            }
        } else {
            out[nout++] = b;
        }
        a = b;
        as = bs;
    }
    return nout;
}

// clipped (optional): set to true when fewer than two polygon points remain
// -- the triangle does not reach into the box -- and the reference's default
// box (lib/clipping.h:210-212) is returned.
inline Box clip_triangle(const float* t, const Box& box, bool* clipped = nullptr) {
    if (clipped) *clipped = false;
    V3 buf0[40], buf1[40];
    V3* pts = buf0;
    V3* nxt = buf1;
    int np = 3;
    pts[0] = load(t);
    pts[1] = load(t + 3);
    pts[2] = load(t + 6);
    for (int ax = 0; ax < 3; ++ax) {
        for (int s = 0; s < 2; ++s) {
            V3 n = mk(0, 0, 0);
            set(n, ax, s == 0 ? 1.f : -1.f);
            float d = s == 0 ? at(box.lo, ax) : -at(box.hi, ax);
            np = clip_plane(pts, np, nxt, n, d);
            std::swap(pts, nxt);
            if (np < 2) {
                if (clipped) *clipped = true;
                return box_default();
            }
        }
    }
    V3 lo = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    V3 hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int ax = 0; ax < 3; ++ax) {
        for (int i = 0; i < np; ++i) {
            float c = at(pts[i], ax);
            if (c < at(lo, ax)) set(lo, ax, c);
            if (at(hi, ax) < c) set(hi, ax, c);
        }
    }
    return box_of(lo, hi);
}

enum : int { kEnding = 0, kPlanar = 1, kStarting = 2 };
struct Event {  // lib/kdtree.cpp:256-265
    uint32_t id;
    float point;
    float point_aux;
    int type;
};

struct Node {
    bool leaf = false;
    int axis = 0;
    float split = 0;
    std::unique_ptr<Node> left, right;
    std::vector<uint32_t> ids;
};

inline float lambda_(size_t nl, size_t nr) { return (nl == 0 || nr == 0) ? 0.8f : 1.f; }
constexpr int kIsect = 20;      // lib/kdtree.cpp:178-179
constexpr int kRefTrav = 15;    // lib/kdtree.cpp:178-179
// Traversal cost of the rendering tree: its leaves are tested by a whole
// wave at once, so a node step costs more relative to a triangle test than
// on the reference's CPU; 60 gave the fastest renders of 15, 30, 60, 120 and
// 240 on the 1.3 M-triangle scenes (0.83 M nodes instead of 2.7 M).
constexpr int kRenderTrav = 60;
// The same for the sub-trees of a refined reference tree (whose leaf tests
// also track exact ties): 120 gave the fastest renders of 30, 60, 120 and
// 240 on the 1.3 M-triangle scene.
constexpr int kRefinedTrav = 120;
inline float cost_(int trav, float lr, float rr, size_t nl, size_t nr) {
    return lambda_(nl, nr) * (trav + kIsect * (lr * nl + rr * nr));
}

// depth limit of the rendering variant (its trees are about 35 deep)
constexpr int kMaxRenderDepth = 80;

class Builder {
public:
    // keep_tol: how far beyond a node's box a triangle's bounding box may end
    // and the triangle still be kept when clipping removes it (rendering variant)
    Builder(const float* verts, bool parallel, bool keep_empty = false, float keep_tol = kEps,
            int trav = kRenderTrav)
        : v_(verts), parallel_(parallel), keep_empty_(keep_empty), trav_(trav), keep_tol_(keep_tol) {
        unsigned hc = std::thread::hardware_concurrency();
        budget_.store(parallel ? (int)(hc ? hc : 4) - 1 : 0);
    }

    std::unique_ptr<Node> build(std::vector<uint32_t> tris, const Box& box, int depth = 0) {
        if (tris.empty()) return keep_empty_ ? leaf({}) : nullptr;
        if (tris.size() <= 3 || area(box) == 0) return leaf(std::move(tris));
        // in the rendering variant, triangles kept by their clamped bounding
        // box can straddle every split: a depth cap, and a split that gives
        // both children every triangle separates nothing
        if (keep_empty_ && depth >= kMaxRenderDepth) return leaf(std::move(tris));
        std::vector<uint32_t> L, R;
        int ax;
        float pos;
        float min_cost = find_plane(tris, box, ax, pos, L, R);
        if (kIsect * tris.size() * lambda_(L.size(), R.size()) < min_cost) return leaf(std::move(tris));
        if (keep_empty_ && L.size() == tris.size() && R.size() == tris.size()) return leaf(std::move(tris));
        Box lb, rb;
        split(box, ax, pos, lb, rb);
        tris.clear();
        tris.shrink_to_fit();
        std::unique_ptr<Node> left, right;
        const bool fork = parallel_ && L.size() > 4096 && R.size() > 4096 && take_thread();
        if (fork) {
            auto fut = std::async(std::launch::async, [this, &L, lb, depth]() {
                auto r = build(std::move(L), lb, depth + 1);
                budget_.fetch_add(1);
                return r;
            });
            right = build(std::move(R), rb, depth + 1);
            left = fut.get();
        } else {
            left = build(std::move(L), lb, depth + 1);
            right = build(std::move(R), rb, depth + 1);
        }
        if (!left) return right;
        if (!right) return left;
        auto n = std::make_unique<Node>();
        n->axis = ax;
        n->split = pos;
        n->left = std::move(left);
        n->right = std::move(right);
        return n;
    }

private:
    bool take_thread() {
        int b = budget_.load();
        while (b > 0) {
            if (budget_.compare_exchange_weak(b, b - 1)) return true;
        }
        return false;
    }
    static std::unique_ptr<Node> leaf(std::vector<uint32_t> ids) {
        auto n = std::make_unique<Node>();
        n->leaf = true;
        n->ids = std::move(ids);
        return n;
    }

    // lib/kdtree.cpp:245-408
    float find_plane(const std::vector<uint32_t>& tris, const Box& box, int& out_ax,
                     float& out_pos, std::vector<uint32_t>& L, std::vector<uint32_t>& R) const {
        std::vector<Event> ev[3];
        for (auto& e : ev) e.reserve(2 * tris.size());
        size_t num = 0;
        // Rendering variant: in a node thinner than 4 kEps along some axis the
        // thick-plane clip is unreliable (the tolerance zones of its two planes
        // overlap, and the polygon can lose the part inside the node), so the
        // triangle's bounding box clamped to the node stands in for it; such a
        // node is not split again along its thin axes.
        bool thin[3] = {false, false, false};
        bool any_thin = false;
        if (keep_empty_)
            for (int a = 0; a < 3; ++a) any_thin |= (thin[a] = at(box.hi, a) - at(box.lo, a) < 4 * kEps);
        for (uint32_t id : tris) {
            bool gone = false;
            Box cb;
            if (any_thin) {
                const Box tb = tri_box(v_ + 9 * (size_t)id);
                if (!near_box(tb, box, keep_tol_)) continue;
                cb = clamp_to(tb, box);
            } else {
                cb = clip_triangle(v_ + 9 * (size_t)id, box, keep_empty_ ? &gone : nullptr);
            }
            // The reference keeps a fully clipped triangle with an unbounded box,
            // so it follows every later split.  The rendering variant drops it
            // only when its bounding box lies beyond the clipping tolerance: the
            // thick-plane clip also removes triangles that overlap the box in a
            // sliver thinner than kEps, which a ray can still hit inside the box;
            // those keep their bounding box clamped to this box.
            if (gone) {
                const Box tb = tri_box(v_ + 9 * (size_t)id);
                if (!near_box(tb, box, keep_tol_)) continue;
                cb = clamp_to(tb, box);
            }
            // the reference skips a triangle whose clipped box is a single point;
            // the rendering variant keeps it: a ray aimed at that point (a vertex
            // on the node's face) hits it inside this node
            if (!keep_empty_ && empty(cb)) continue;
            ++num;
            for (int a = 0; a < 3; ++a) {
                if (planar(cb, a)) {
                    ev[a].push_back(Event{id, at(cb.lo, a), at(cb.lo, a), kPlanar});
                } else {
                    ev[a].push_back(Event{id, at(cb.lo, a), at(cb.hi, a), kStarting});
                    ev[a].push_back(Event{id, at(cb.hi, a), at(cb.lo, a), kEnding});
                }
            }
        }
        if (num == 0) {
            out_ax = 0;
            out_pos = 0;
            return std::numeric_limits<float>::max();
        }
        float min_cost = std::numeric_limits<float>::max();
        int min_ax = 0;
        float min_pos = 0;
        bool min_left = true;
        const float A = area(box);
        for (int a = 0; a < 3; ++a) {
            if (thin[a]) continue;
            auto& E = ev[a];
            std::sort(E.begin(), E.end(), [](const Event& e1, const Event& e2) {
                return e1.point < e2.point || (e1.point == e2.point && e1.type < e2.type);
            });
            size_t nl = 0, np = 0, nr = num;
            const size_t n = E.size();
            for (size_t i = 0; i < n;) {
                const float p = E[i].point;
                int ps = 0, pe = 0, pp = 0;
                while (i < n && E[i].point == p && E[i].type == kEnding) { ++pe; ++i; }
                while (i < n && E[i].point == p && E[i].type == kPlanar) { ++pp; ++i; }
                while (i < n && E[i].point == p && E[i].type == kStarting) { ++ps; ++i; }
                np = pp;
                nr -= pp + pe;
                Box lb, rb;  // lib/kdtree.cpp:218-238
                split(box, a, p, lb, rb);
                float lr = area(lb) / A, rr = area(rb) / A;
                const int trav = keep_empty_ ? trav_ : kRefTrav;
                float lc = cost_(trav, lr, rr, nl + np, nr);
                float rc = cost_(trav, lr, rr, nl, np + nr);
                float c = lc < rc ? lc : rc;
                bool left = lc < rc;
                if (c < min_cost) {
                    min_cost = c;
                    min_ax = a;
                    min_pos = p;
                    min_left = left;
                }
                nl += ps + pp;
                np = 0;
            }
        }
        // classify: lib/kdtree.cpp:369-399
        for (const Event& e : ev[min_ax]) {
            if (e.point < min_pos) {
                if (e.type == kEnding || e.type == kPlanar) {
                    L.push_back(e.id);
                } else if (min_pos < e.point_aux) {
                    L.push_back(e.id);
                    R.push_back(e.id);
                }
            } else if (e.point == min_pos) {
                if (e.type == kEnding) L.push_back(e.id);
                else if (e.type == kPlanar) (min_left ? L : R).push_back(e.id);
                else R.push_back(e.id);
            } else if (e.type == kStarting || e.type == kPlanar) {
                R.push_back(e.id);
            }
        }
        out_ax = min_ax;
        out_pos = min_pos;
        return min_cost;
    }

    const float* v_;
    bool parallel_;
    bool keep_empty_;  // empty children become empty leaves instead of being dropped
    int trav_;         // traversal cost of the rendering variant
    float keep_tol_;
    std::atomic<int> budget_{0};
};

inline uint64_t flat_inner(int axis, float split, uint32_t right) {
    uint32_t bits;
    std::memcpy(&bits, &split, 4);
    return (uint64_t)bits << 32 | (uint64_t)(right << 2 | (uint32_t)axis);
}
inline uint64_t flat_leaf(uint32_t a, uint32_t b) {
    return (uint64_t)a << 32 | (uint64_t)(uint32_t)(b << 2 | 3u);
}

// lib/kdtree.cpp:420-467 (depth-first, explicit stack, right pushed first)
void flatten(const Node* root, std::vector<uint64_t>& out) {
    constexpr uint32_t kInvalid = 0xFFFFFFFFu >> 2;
    std::vector<std::pair<const Node*, uint32_t>> st;
    st.emplace_back(root, kInvalid);
    while (!st.empty()) {
        const Node* n = st.back().first;
        uint32_t parent = st.back().second;
        st.pop_back();
        uint32_t me = (uint32_t)out.size();
        if (parent != kInvalid) {
            out[parent] = (out[parent] & 0xFFFFFFFF00000000ull) |
                          (uint64_t)(me << 2 | (uint32_t)(out[parent] & 3));
        }
        if (!n->leaf) {
            out.push_back(flat_inner(n->axis, n->split, kInvalid));
            st.emplace_back(n->right.get(), me);
            st.emplace_back(n->left.get(), kInvalid);
        } else {
            size_t i = 1;
            for (; i < n->ids.size(); i += 2) out.push_back(flat_leaf(n->ids[i - 1], n->ids[i]));
            if (i - 1 < n->ids.size()) out.push_back(flat_leaf(n->ids[i - 1], kInvalid));
            else out.push_back(flat_inner(0, 0.f, 0));
        }
    }
}

}  // namespace

void precompute_triangles(const float* verts9, uint32_t n, float* out16) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* t = verts9 + 9 * (size_t)i;
        V3 v0 = load(t), v1 = load(t + 3), v2 = load(t + 6);
        V3 u = v1 - v0, v = v2 - v0;
        // src/geometry.h:288-300: the cross product is evaluated in double
        double ux = u.x, uy = u.y, uz = u.z, vx = v.x, vy = v.y, vz = v.z;
        V3 c = mk((float)((uy * vz) - (uz * vy)), (float)((uz * vx) - (ux * vz)),
                  (float)((ux * vy) - (uy * vx)));
        float len = std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
        V3 nn = (1 / len) * c;  // normalize: (1 / len) * v, src/geometry.h:243-246
        float uv = dot(u, v), vv = dot(v, v), uu = dot(u, u);
        float* o = out16 + 16 * (size_t)i;
        o[0] = v0.x; o[1] = v0.y; o[2] = v0.z;
        o[3] = nn.x; o[4] = nn.y; o[5] = nn.z;
        o[6] = u.x; o[7] = u.y; o[8] = u.z;
        o[9] = v.x; o[10] = v.y; o[11] = v.z;
        o[12] = uv; o[13] = vv; o[14] = uu; o[15] = uv * uv - uu * vv;
    }
}

// Every leaf below n gets every candidate triangle that meets the leaf's cell
// enlarged by tol (separating-axis test), besides the builder's own choice,
// whose thick-plane clipping can drop a sliver beyond a split plane, a
// triangle whose clipped box is thinner than the clipping tolerance, or one
// whose vertex lies on a split plane: a hit point inside a cell is then found
// in that cell.  Leaf order: ascending triangle id.
void cover_tree(const float* verts9, Node* n, const Box& box, const std::vector<uint32_t>& cand, float tol) {
    std::vector<uint32_t> mine;
    mine.reserve(cand.size());
    for (uint32_t id : cand) {
        const float* t = verts9 + 9 * (size_t)id;
        if (near_box(tri_box(t), box, tol) && tri_meets_box(t, box, tol)) mine.push_back(id);
    }
    if (!n->leaf) {
        Box l, r;
        split(box, n->axis, n->split, l, r);
        cover_tree(verts9, n->left.get(), l, mine, tol);
        cover_tree(verts9, n->right.get(), r, mine, tol);
        return;
    }
    n->ids.insert(n->ids.end(), mine.begin(), mine.end());
    std::sort(n->ids.begin(), n->ids.end());
    n->ids.erase(std::unique(n->ids.begin(), n->ids.end()), n->ids.end());
}

void build_kdtree(const float* verts9, uint32_t n, bool parallel, std::vector<uint64_t>& flat,
                  float box6[6]) {
    std::vector<uint32_t> ids(n, 0);
    Box box = tri_box(verts9);
    for (uint32_t i = 1; i < n; ++i) {  // lib/kdtree.cpp:481-485
        Box b = tri_box(verts9 + 9 * (size_t)i);
        box = box_of(mk(std::min(box.lo.x, b.lo.x), std::min(box.lo.y, b.lo.y),
                        std::min(box.lo.z, b.lo.z)),
                     mk(std::max(box.hi.x, b.hi.x), std::max(box.hi.y, b.hi.y),
                        std::max(box.hi.z, b.hi.z)));
        ids[i] = i;
    }
    Builder B(verts9, parallel);
    std::unique_ptr<Node> root = B.build(std::move(ids), box);
    flat.clear();
    flatten(root.get(), flat);
    // tear the tree down iteratively (deep trees would recurse a lot)
    std::vector<std::unique_ptr<Node>> kill;
    kill.push_back(std::move(root));
    while (!kill.empty()) {
        auto n = std::move(kill.back());
        kill.pop_back();
        if (n->left) kill.push_back(std::move(n->left));
        if (n->right) kill.push_back(std::move(n->right));
    }
    box6[0] = box.lo.x; box6[1] = box.lo.y; box6[2] = box.lo.z;
    box6[3] = box.hi.x; box6[4] = box.hi.y; box6[5] = box.hi.z;
}

// The same SAH build with two changes for rendering: a triangle that clipping
// removes from a node entirely is dropped unless its bounding box reaches
// within the clipping tolerance of the node (the reference gives every such
// triangle an unbounded box, lib/clipping.h:210-212, so it sinks into every
// descendant and fills leaves with triangles that do not touch them), and a
// split whose one side is empty keeps that side as an empty leaf (the
// reference drops it, lib/kdtree.cpp:166-171).  Flattened straight into the
// device layout.
void build_device_tree(const float* verts9, uint32_t n, bool parallel, std::vector<DevNode>& nodes,
                       std::vector<uint32_t>& refs, float box6[6], uint32_t& height) {
    std::vector<uint32_t> ids(n, 0);
    Box box = tri_box(verts9);
    for (uint32_t i = 1; i < n; ++i) {
        Box b = tri_box(verts9 + 9 * (size_t)i);
        box = box_of(mk(std::min(box.lo.x, b.lo.x), std::min(box.lo.y, b.lo.y),
                        std::min(box.lo.z, b.lo.z)),
                     mk(std::max(box.hi.x, b.hi.x), std::max(box.hi.y, b.hi.y),
                        std::max(box.hi.z, b.hi.z)));
        ids[i] = i;
    }
    Builder B(verts9, parallel, true);
    std::unique_ptr<Node> root = B.build(std::move(ids), box);
    nodes.clear();
    refs.clear();
    height = 0;
    // depth-first: left child = next node, right child patched when reached
    struct Item {
        const Node* n;
        uint32_t parent;  // node whose right index this one is, or ~0
        uint32_t level;
    };
    std::vector<Item> st{{root.get(), 0xFFFFFFFFu, 0}};
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        const uint32_t me = (uint32_t)nodes.size();
        if (it.parent != 0xFFFFFFFFu) nodes[it.parent].y |= me << 2;
        height = std::max(height, it.level);
        if (!it.n->leaf) {
            uint32_t split;
            std::memcpy(&split, &it.n->split, 4);
            nodes.push_back(DevNode{split, (uint32_t)it.n->axis});
            st.push_back({it.n->right.get(), me, it.level + 1});
            st.push_back({it.n->left.get(), 0xFFFFFFFFu, it.level + 1});
        } else {
            nodes.push_back(DevNode{(uint32_t)refs.size(), (uint32_t)it.n->ids.size() << 2 | 3u});
            refs.insert(refs.end(), it.n->ids.begin(), it.n->ids.end());
        }
    }
    std::vector<std::unique_ptr<Node>> kill;
    kill.push_back(std::move(root));
    while (!kill.empty()) {
        auto k = std::move(kill.back());
        kill.pop_back();
        if (k->left) kill.push_back(std::move(k->left));
        if (k->right) kill.push_back(std::move(k->right));
    }
    box6[0] = box.lo.x; box6[1] = box.lo.y; box6[2] = box.lo.z;
    box6[3] = box.hi.x; box6[4] = box.hi.y; box6[5] = box.hi.z;
}

// The reference tree with refined leaves.  Every inner node of the reference
// device tree is kept, so leaves are reached in the reference's order; each
// leaf keeps only the references whose triangle's bounding box reaches within
// tol of the leaf's box (tol = kEps + 2^-16 of the box's largest coordinate
// magnitude, far above the rounding of any hit point or split crossing: the
// others are triangles the reference's unbounded-box quirk sank into the
// leaf from an ancestor), and a leaf keeping more than four is replaced by a
// rendering-variant sub-tree over its box (SAH node step kRefinedTrav,
// clipped-away triangles kept within tol) whose every leaf then also gets
// each of those triangles that meets its cell (cover_tree, tolerance 2^-26 of
// the magnitude, above the rounding of the FP64 overlap test): a hit point
// inside a cell is found in that cell.  Closest hits then equal the reference tree's
// except where two triangles are hit at exactly the same distance (the
// winner depends on the order inside a leaf) or a direction component is
// zero (its 1e-5 substitute bends the traversal path away from the ray):
// the caller re-traces those rays on the reference tree.
void refine_reference_tree(const float* verts9, const std::vector<DevNode>& rnodes,
                           const std::vector<uint32_t>& rrefs, const float box6[6], bool parallel,
                           std::vector<DevNode>& nodes, std::vector<uint32_t>& refs, uint32_t& height) {
    struct LeafJob {
        uint32_t node;
        Box box;
        std::vector<uint32_t> ids;
        std::unique_ptr<Node> sub;
    };
    // leaves of the reference tree with their boxes
    std::vector<LeafJob> jobs;
    std::vector<uint32_t> job_of(rnodes.size(), 0xFFFFFFFFu);
    {
        struct It {
            uint32_t node;
            Box box;
        };
        std::vector<It> st{{0u, Box{mk(box6[0], box6[1], box6[2]), mk(box6[3], box6[4], box6[5])}}};
        while (!st.empty()) {
            It it = st.back();
            st.pop_back();
            const DevNode d = rnodes[it.node];
            if ((d.y & 3u) == 3u) {
                job_of[it.node] = (uint32_t)jobs.size();
                LeafJob j;
                j.node = it.node;
                j.box = it.box;
                j.ids.assign(rrefs.begin() + d.x, rrefs.begin() + d.x + (d.y >> 2));
                jobs.push_back(std::move(j));
                continue;
            }
            const int ax = (int)(d.y & 3u);
            float pos;
            std::memcpy(&pos, &d.x, 4);
            Box l, r;
            split(it.box, ax, pos, l, r);
            st.push_back({d.y >> 2, r});
            st.push_back({it.node + 1, l});
        }
    }
    auto refine = [&](LeafJob& j) {
        float mag = 0.f;
        for (int ax = 0; ax < 3; ++ax)
            mag = std::max(mag, std::max(std::abs(at(j.box.lo, ax)), std::abs(at(j.box.hi, ax))));
        const float tol = kEps + std::ldexp(mag, -16);
        std::vector<uint32_t> keep;
        keep.reserve(j.ids.size());
        for (uint32_t id : j.ids)
            if (near_box(tri_box(verts9 + 9 * (size_t)id), j.box, tol)) keep.push_back(id);
        if (keep.size() > 4) {
            Builder B(verts9, false, true, tol, kRefinedTrav);
            j.sub = B.build(keep, j.box);
            cover_tree(verts9, j.sub.get(), j.box, keep, std::ldexp(mag, -26));
            j.ids.clear();
        } else {
            j.ids = std::move(keep);
        }
    };
    const unsigned hc = parallel ? std::max(1u, std::thread::hardware_concurrency()) : 1u;
    if (hc > 1 && jobs.size() > 64) {
        std::atomic<size_t> next{0};
        std::vector<std::thread> th;
        for (unsigned t = 0; t < hc; ++t)
            th.emplace_back([&]() {
                for (size_t k; (k = next.fetch_add(1)) < jobs.size();) refine(jobs[k]);
            });
        for (auto& t : th) t.join();
    } else {
        for (auto& j : jobs) refine(j);
    }
    // depth-first emission: reference inner nodes, then each leaf's sub-tree
    nodes.clear();
    refs.clear();
    height = 0;
    struct Item {
        const Node* sub;    // a sub-tree node, or null for a reference node
        uint32_t rnode;     // the reference node (when sub is null)
        uint32_t parent;    // node whose right index this one is, or ~0
        uint32_t level;
    };
    std::vector<Item> st{{nullptr, 0u, 0xFFFFFFFFu, 0u}};
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        const uint32_t me = (uint32_t)nodes.size();
        if (it.parent != 0xFFFFFFFFu) nodes[it.parent].y |= me << 2;
        height = std::max(height, it.level);
        const Node* n = it.sub;
        if (!n) {
            const DevNode d = rnodes[it.rnode];
            if ((d.y & 3u) != 3u) {
                nodes.push_back(DevNode{d.x, d.y & 3u});
                st.push_back({nullptr, d.y >> 2, me, it.level + 1});
                st.push_back({nullptr, it.rnode + 1, 0xFFFFFFFFu, it.level + 1});
                continue;
            }
            LeafJob& j = jobs[job_of[it.rnode]];
            if (!j.sub) {
                nodes.push_back(DevNode{(uint32_t)refs.size(), (uint32_t)j.ids.size() << 2 | 3u});
                refs.insert(refs.end(), j.ids.begin(), j.ids.end());
                continue;
            }
            n = j.sub.get();
        }
        if (!n->leaf) {
            uint32_t sp;
            std::memcpy(&sp, &n->split, 4);
            nodes.push_back(DevNode{sp, (uint32_t)n->axis});
            st.push_back({n->right.get(), 0u, me, it.level + 1});
            st.push_back({n->left.get(), 0u, 0xFFFFFFFFu, it.level + 1});
        } else {
            nodes.push_back(DevNode{(uint32_t)refs.size(), (uint32_t)n->ids.size() << 2 | 3u});
            refs.insert(refs.end(), n->ids.begin(), n->ids.end());
        }
    }
}

uint32_t tree_height(const std::vector<uint64_t>& flat) {  // lib/kdtree.h:197-216
    std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 0u}};
    uint32_t h = 0;
    while (!st.empty()) {
        auto [node, level] = st.back();
        st.pop_back();
        h = std::max(h, level);
        uint64_t d = flat[node];
        if ((d & 3) != 3) {
            st.emplace_back(node + 1, level + 1);
            st.emplace_back((uint32_t)d >> 2, level + 1);
        }
    }
    return h;
}

bool flat_to_device(const std::vector<uint64_t>& flat, uint32_t ntris, std::vector<DevNode>& nodes,
                    std::vector<uint32_t>& refs, std::string& err) {
    nodes.clear();
    refs.clear();
    if (flat.empty()) {
        err = "empty tree";
        return false;
    }
    const uint32_t nf = (uint32_t)flat.size();
    // Pass 1: reachable node starts (inner nodes and leaf-run heads) in DFS order.
    std::vector<uint32_t> map(nf, 0xFFFFFFFFu);
    std::vector<uint32_t> order;
    std::vector<uint32_t> st{0};
    while (!st.empty()) {
        uint32_t i = st.back();
        st.pop_back();
        if (i >= nf) {
            err = "node index out of range";
            return false;
        }
        if (map[i] != 0xFFFFFFFFu) {
            err = "node reached twice";
            return false;
        }
        map[i] = (uint32_t)order.size();
        order.push_back(i);
        uint64_t d = flat[i];
        if ((d & 3) != 3) {
            uint32_t right = (uint32_t)d >> 2;
            if (right <= i || right >= nf || i + 1 >= nf) {
                err = "inner node with bad children";
                return false;
            }
            st.push_back(right);
            st.push_back(i + 1);
        }
    }
    nodes.resize(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
        uint32_t i = order[k];
        uint64_t d = flat[i];
        if ((d & 3) != 3) {
            uint32_t right = map[(uint32_t)d >> 2];
            nodes[k] = DevNode{(uint32_t)(d >> 32), right << 2 | (uint32_t)(d & 3)};
        } else {
            uint32_t off = (uint32_t)refs.size();
            // lib/kdtree.cpp:599-605: pairs until an inner node or a single-id leaf
            for (uint32_t j = i;; ++j) {
                if (j >= nf) {
                    err = "unterminated leaf run";
                    return false;
                }
                uint64_t e = flat[j];
                if ((e & 3) != 3) break;
                uint32_t a = (uint32_t)(e >> 32);
                if (a >= ntris) {
                    err = "triangle id out of range";
                    return false;
                }
                refs.push_back(a);
                if ((uint32_t)e == 0xFFFFFFFFu) break;
                uint32_t b = (uint32_t)e >> 2;
                if (b >= ntris) {
                    err = "triangle id out of range";
                    return false;
                }
                refs.push_back(b);
            }
            uint32_t cnt = (uint32_t)refs.size() - off;
            if (cnt >= (1u << 30)) {
                err = "leaf too large";
                return false;
            }
            nodes[k] = DevNode{off, cnt << 2 | 3u};
        }
    }
    return true;
}

}  // namespace tn
