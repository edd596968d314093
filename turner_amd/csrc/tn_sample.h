// tn_sample.h -- next-event estimation toward triangle emitters: solid-angle
// sampling of one triangle as seen from a shading point.
//
// No reference counterpart: turner's pathtracer has point lights only
// (pathtracer.cpp:44-58).  A point drawn uniformly by area on an emitter
// near the shading point carries the factor cos_l / d^2, which varies by
// orders of magnitude over a light a few of its own widths away; drawing the
// direction uniformly in the triangle's solid angle instead leaves only the
// cosine at the shading point in the estimate.
//
// The sampler is Arvo's (J. Arvo, "Stratified Sampling of Spherical
// Triangles", SIGGRAPH 1995): the triangle is projected onto the unit sphere
// around the shading point (vertices A, B, C, interior angles alpha, beta,
// gamma, area Omega = alpha + beta + gamma - pi); u1 picks the area of a
// sub-triangle A B C' with C' on the arc AC, u2 a point on the arc from B to
// C'.  The density is 1 / Omega per steradian.  Omega comes from the
// interior angles and loses relative precision as it shrinks, so the caller
// falls back to area sampling below kMinSolidAngle (where cos_l / d^2 is
// nearly constant over the triangle anyway).
//
// Host and device: the same arithmetic runs in the shade kernel and in the
// host test hook tn_sample_triangle_host.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace tn {

constexpr float kMinSolidAngle = 1e-3f;  // sr; below this the emitter is sampled by area

__host__ __device__ __forceinline__ float3 v3sub(float3 a, float3 b) {
    return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__host__ __device__ __forceinline__ float v3dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ float3 v3cross(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ float3 v3scale(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }
__host__ __device__ __forceinline__ float3 v3axpy(float s, float3 a, float3 b) {  // s a + b
    return make_float3(s * a.x + b.x, s * a.y + b.y, s * a.z + b.z);
}
// a / |a|, or zero when |a| underflows
__host__ __device__ __forceinline__ float3 v3unit(float3 a) {
    const float l2 = v3dot(a, a);
    return l2 > 1e-30f ? v3scale(1.0f / sqrtf(l2), a) : make_float3(0.f, 0.f, 0.f);
}
__host__ __device__ __forceinline__ float clamp1(float x) { return fminf(1.0f, fmaxf(-1.0f, x)); }

// Solid angle of the triangle with vertices a, b, c (relative to the shading
// point), from the triple product (Van Oosterom and Strackee 1983):
// tan(Omega / 2) = |A . (B x C)| / (1 + A.B + B.C + C.A) for unit A, B, C.
// Used as a selection weight among emitters.
__host__ __device__ inline float tri_solid_angle(float3 a, float3 b, float3 c) {
    const float3 A = v3unit(a), B = v3unit(b), C = v3unit(c);
    const float num = fabsf(v3dot(A, v3cross(B, C)));
    const float den = 1.0f + v3dot(A, B) + v3dot(B, C) + v3dot(C, A);
    return 2.0f * atan2f(num, den);
}

struct SphTri {
    float3 A, B, C;  // unit vectors toward the vertices
    float cos_alpha, sin_alpha, alpha;
    float cos_c;  // cosine of the arc AB
    float omega;  // solid angle (sr)
};

// Projects the triangle (vertices a, b, c relative to the shading point).
// Returns false for a degenerate projection (a vertex at the point, or the
// point in the triangle's plane).
__host__ __device__ inline bool sph_tri_setup(float3 a, float3 b, float3 c, SphTri& T) {
    T.A = v3unit(a);
    T.B = v3unit(b);
    T.C = v3unit(c);
    const float3 nab = v3unit(v3cross(T.A, T.B)), nac = v3unit(v3cross(T.A, T.C)),
                 nbc = v3unit(v3cross(T.B, T.C));
    if (v3dot(nab, nab) == 0.f || v3dot(nac, nac) == 0.f || v3dot(nbc, nbc) == 0.f) return false;
    // interior angle at a vertex = angle between the planes of its two arcs
    T.cos_alpha = clamp1(v3dot(nab, nac));
    const float cos_beta = clamp1(-v3dot(nab, nbc));
    const float cos_gamma = clamp1(v3dot(nac, nbc));
    T.alpha = acosf(T.cos_alpha);
    T.sin_alpha = sqrtf(fmaxf(0.f, 1.0f - T.cos_alpha * T.cos_alpha));
    T.omega = T.alpha + acosf(cos_beta) + acosf(cos_gamma) - 3.14159265358979323846f;
    T.cos_c = v3dot(T.A, T.B);
    return T.omega > 0.f && T.sin_alpha > 0.f;
}

// Unit direction uniform in the triangle's solid angle, from u1, u2 in [0, 1).
__host__ __device__ inline float3 sph_tri_sample(const SphTri& T, float u1, float u2) {
    const float area = u1 * T.omega;  // area of the sub-triangle A B C'
    const float s = sinf(area - T.alpha), t = cosf(area - T.alpha);
    const float u = t - T.cos_alpha;
    const float v = s + T.sin_alpha * T.cos_c;
    // cosine of the arc A C'
    const float q = clamp1(((v * t - u * s) * T.cos_alpha - v) / ((v * s + u * t) * T.sin_alpha));
    const float3 e = v3unit(v3axpy(-v3dot(T.C, T.A), T.A, T.C));
    const float3 c2 = v3axpy(q, T.A, v3scale(sqrtf(fmaxf(0.f, 1.0f - q * q)), e));
    // point on the arc B C' with the cosine of its arc from B uniform in [cos(BC'), 1]
    const float z = 1.0f - u2 * (1.0f - v3dot(c2, T.B));
    const float3 f = v3unit(v3axpy(-v3dot(c2, T.B), T.B, c2));
    return v3unit(v3axpy(z, T.B, v3scale(sqrtf(fmaxf(0.f, 1.0f - z * z)), f)));
}

}  // namespace tn
