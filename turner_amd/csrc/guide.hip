// guide.hip -- training and inference kernels of the learned guiding
// proposal (tn_guide.h) and the tn_guide_* entry points.  No reference
// counterpart: the reference draws indirect directions uniformly
// (lib/sampling.h:20-32).
//
// Every matrix product runs on mfma_f32_16x16x16bf16_1k with BF16 operands
// and FP32 accumulation; one wave (64 samples) per block:
//
//   forward   H_l^T  = W_l H_{l-1}^T          A = W_l (BF16 copy), B = LDS rows
//   backward  D_{l-1}^T = W_l^T D_l^T (.) [H_{l-1} > 0]   A = W_l^T (BF16 copy)
//   gradient  dW_l = D_l^T H_{l-1}            K = the block's 64 samples
//
// Each block writes its gradient slice (and the sums of c and c * -log q)
// to its own row of a partial buffer; a two-stage reduction sums the rows in
// a fixed order, so a step is deterministic.  The per-sample loss gradient
// with respect to the 40 outputs is evaluated in FP32 in log space.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <random>
#include <vector>

#include "tn_guide.h"
#include "tn_scene.h"

namespace tn {
namespace {

constexpr int kGPartial = kGParams + 8;  // gradient slice, sum c, sum c * -log q (+ padding)

__device__ __forceinline__ uint64_t g_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// BF16 operand copies from the FP32 master parameters: W_l padded
// [out_pad][in_pad] and W_l^T padded [in_pad][out_pad], zero padding.
__global__ void k_guide_refresh(const float* __restrict__ theta, uint16_t* __restrict__ wbf) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kGBf16) return;
    float v = 0.f;
    if (e < kGW2) {
        const int o = e / kGInPad, i = e % kGInPad;
        if (i < kGIn) v = theta[kGOffW1 + o * kGIn + i];
    } else if (e < kGW3) {
        v = theta[kGOffW2 + (e - kGW2)];
    } else if (e < kGW4) {
        v = theta[kGOffW3 + (e - kGW3)];
    } else if (e < kGWT2) {
        const int k = e - kGW4, o = k / kGHid, i = k % kGHid;
        if (o < kGOut) v = theta[kGOffW4 + o * kGHid + i];
    } else if (e < kGWT3) {
        const int k = e - kGWT2, i = k / kGHid, o = k % kGHid;
        v = theta[kGOffW2 + o * kGHid + i];
    } else if (e < kGWT4) {
        const int k = e - kGWT3, i = k / kGHid, o = k % kGHid;
        v = theta[kGOffW3 + o * kGHid + i];
    } else {
        const int k = e - kGWT4, i = k / kGOutPad, o = k % kGOutPad;
        if (o < kGOut) v = theta[kGOffW4 + o * kGHid + i];
    }
    wbf[e] = to_bf16(v);
}

__device__ __forceinline__ float3 load3(const float* a, uint32_t i) {
    return make_float3(a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2]);
}

// mode 0: raw outputs (40 per point) into out
// mode 1: q(w) into pdf
// mode 2: w ~ q from u3 into w3out, q(w) into pdf
__global__ __launch_bounds__(64) void k_guide_eval(GuideView G, const float* __restrict__ p3,
                                                   const float* __restrict__ n3,
                                                   const float* __restrict__ w3,
                                                   const float* __restrict__ u3, uint32_t n, int mode,
                                                   float* __restrict__ out, float* __restrict__ pdf,
                                                   float* __restrict__ w3out) {
    __shared__ __align__(16) uint16_t lds[64 * (kGStrideIn + kGStrideHid)];
    const uint32_t l = threadIdx.x;
    const uint32_t base = blockIdx.x * 64u;
    const uint32_t i = base + l;
    const bool act = i < n;
    float3 p = make_float3(0.f, 0.f, 0.f), nn = p;
    if (act) {
        p = load3(p3, i);
        nn = load3(n3, i);
    }
    float* o = reinterpret_cast<float*>(lds);
    guide_forward_wave(G, p, nn, act, o, lds);
    if (mode == 0) {
        const uint32_t rows = min(64u, n - base);
        for (uint32_t k = l; k < rows * kGOut; k += 64) out[(size_t)base * kGOut + k] = o[k];
        return;
    }
    if (!act) return;
    Mixture M;
    mixture_from(o + l * kGOut, M);
    float3 w;
    if (mode == 2) {
        w = mixture_sample(M, u3[3 * (size_t)i], u3[3 * (size_t)i + 1], u3[3 * (size_t)i + 2]);
        w3out[3 * (size_t)i] = w.x;
        w3out[3 * (size_t)i + 1] = w.y;
        w3out[3 * (size_t)i + 2] = w.z;
    } else {
        w = load3(w3, i);
    }
    pdf[i] = mixture_pdf(M, w);
}

// MLP outputs for n points with persistent waves: each wave (block) loads
// every BF16 weight fragment into registers once and then runs groups of 64
// points, so weights are not re-read per group; the products use gfx950's
// K = 32 BF16 MFMA (half the instructions of the K = 16 form).
__global__ __launch_bounds__(64) void k_guide_forward(GuideView G, const float* __restrict__ p3,
                                                      const float* __restrict__ n3, uint32_t n,
                                                      float* __restrict__ out) {
    __shared__ __align__(16) uint16_t lds[64 * (kGStrideIn + kGStrideHid)];
    GuideFrags F;  // register-resident for every group this wave takes
    load_guide_frags(G, F);
    const uint32_t l = threadIdx.x;
    float* o = reinterpret_cast<float*>(lds);
    const uint32_t groups = (n + 63) / 64;
    for (uint32_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const uint32_t base = grp * 64u, i = base + l;
        const bool act = i < n;
        float3 p = make_float3(0.f, 0.f, 0.f), nn = p;
        if (act) {
            p = load3(p3, i);
            nn = load3(n3, i);
        }
        guide_forward_wave32(G, F, p, nn, act, o, lds);
        const uint32_t rows = min(64u, n - base);
        for (uint32_t k = l; k < rows * kGOut; k += 64) out[(size_t)base * kGOut + k] = o[k];
        __syncthreads();
    }
}

// d(c * -log q(w)) / d(out) for one sample, FP32 in log space; returns -log q.
__device__ float guide_dout(const float* o, float3 w, float c, float* d) {
    const float kLog2Pi = 1.83787706640934548356f;
    float mx = o[0];
#pragma unroll
    for (int j = 1; j < kGLobes; ++j) mx = fmaxf(mx, o[5 * j]);
    float se = 0.f;
#pragma unroll
    for (int j = 0; j < kGLobes; ++j) se += expf(o[5 * j] - mx);
    const float lse = mx + logf(se);
    float a[kGLobes], mu[kGLobes][3], inv[kGLobes], kap[kGLobes], cs[kGLobes];
    bool live[kGLobes];
    float amax = -INFINITY;
#pragma unroll
    for (int j = 0; j < kGLobes; ++j) {
        const float mx3 = o[5 * j + 1], my = o[5 * j + 2], mz = o[5 * j + 3];
        inv[j] = 1.0f / fmaxf(sqrtf(mx3 * mx3 + my * my + mz * mz), 1e-12f);
        mu[j][0] = mx3 * inv[j];
        mu[j][1] = my * inv[j];
        mu[j][2] = mz * inv[j];
        const float sp = softplusf(o[5 * j + 4]) + kKappaMin;
        live[j] = sp < kKappaMax;
        kap[j] = fminf(sp, kKappaMax);
        cs[j] = mu[j][0] * w.x + mu[j][1] * w.y + mu[j][2] * w.z;
        const float logv = logf(kap[j]) - kLog2Pi - logf(-expm1f(-2.0f * kap[j])) + kap[j] * (cs[j] - 1.0f);
        a[j] = o[5 * j] - lse + logv;
        amax = fmaxf(amax, a[j]);
    }
    float sa = 0.f;
#pragma unroll
    for (int j = 0; j < kGLobes; ++j) sa += expf(a[j] - amax);
    const float logq = amax + logf(sa);
#pragma unroll
    for (int j = 0; j < kGLobes; ++j) {
        const float gam = expf(a[j] - logq), pi = expf(o[5 * j] - lse);
        d[5 * j] = -(gam - pi) * c;
        const float s = -gam * kap[j] * c;
        const float3 dmu = make_float3(s * w.x, s * w.y, s * w.z);
        const float dot = dmu.x * mu[j][0] + dmu.y * mu[j][1] + dmu.z * mu[j][2];
        d[5 * j + 1] = (dmu.x - dot * mu[j][0]) * inv[j];
        d[5 * j + 2] = (dmu.y - dot * mu[j][1]) * inv[j];
        d[5 * j + 3] = (dmu.z - dot * mu[j][2]) * inv[j];
        // d log vMF / d kappa = 1/k - 2/(e^{2k} - 1) + (cos - 1); series below 0.05
        const float k = kap[j];
        const float dk = (k < 0.05f ? 1.0f - k * (1.0f / 3.0f) + k * k * k * (1.0f / 45.0f)
                                    : 1.0f / k - 2.0f / expm1f(2.0f * k)) +
                         (cs[j] - 1.0f);
        const float sig = 1.0f / (1.0f + expf(-o[5 * j + 4]));
        d[5 * j + 4] = live[j] ? -gam * dk * c * sig : 0.f;
    }
    return -logq;
}

// dW[o][i] = sum over the 64 samples of D[s][o] H[s][i] (A = D^T tile, B =
// H tile, K = samples), stored for o < out_real, i < in_real.
template <int M_TILES, int N_TILES>
__device__ __forceinline__ void weight_grad(const uint16_t* D, int ds, const uint16_t* H, int hs,
                                            float* __restrict__ out, int out_real, int in_real) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, grp = l >> 4;
#pragma unroll
    for (int mt = 0; mt < M_TILES; ++mt) {
#pragma unroll
        for (int nt = 0; nt < N_TILES; ++nt) {
            g_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                g_s4 a, b;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int s = 16 * kt + 4 * grp + r;
                    a[r] = (short)D[s * ds + 16 * mt + row];
                    b[r] = (short)H[s * hs + 16 * nt + row];
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * mt + 4 * grp + r, i = 16 * nt + row;
                if (o < out_real && i < in_real) out[o * in_real + i] = acc[r];
            }
        }
    }
}

// db[o] = sum over the 64 samples of D[s][o].
__device__ __forceinline__ void bias_grad(const uint16_t* D, int ds, int out_real, float* __restrict__ out) {
    const int l = threadIdx.x & 63;
    if (l < out_real) {
        float s = 0.f;
        for (int k = 0; k < 64; ++k) s += from_bf16(D[k * ds + l]);
        out[l] = s;
    }
}

// D_{l-1} = (D_l W_l) (.) [H_{l-1} > 0], BF16 rows of 64 into dst.
template <int K_TILES>
__device__ __forceinline__ void delta_back(const uint16_t* WT, const uint16_t* D, const uint16_t* H,
                                           uint16_t* dst) {
    mfma_wx<K_TILES, 4>(WT, D, kGStrideHid, [&](int o, int smp, g_f4 v) {
        const uint2 h = *reinterpret_cast<const uint2*>(H + smp * kGStrideHid + o);
        const float m0 = from_bf16((uint16_t)h.x) > 0.f ? v[0] : 0.f;
        const float m1 = from_bf16((uint16_t)(h.x >> 16)) > 0.f ? v[1] : 0.f;
        const float m2 = from_bf16((uint16_t)h.y) > 0.f ? v[2] : 0.f;
        const float m3 = from_bf16((uint16_t)(h.y >> 16)) > 0.f ? v[3] : 0.f;
        *reinterpret_cast<uint2*>(dst + smp * kGStrideHid + o) = make_uint2(pack_bf16x2(m0, m1), pack_bf16x2(m2, m3));
    });
}

struct TrainSrc {
    // explicit batch (device arrays) ...
    const float* p3;
    const float* n3;
    const float* w3;
    const float* c;
    uint32_t n;
    // ... or n samples drawn from the training set
    const float4* ts;
    const uint32_t* ts_count;
    uint32_t ts_cap;
    uint64_t seed;
    uint32_t subset;       // distinct samples this pass trains on
    uint64_t subset_seed;  // fixes which ones
};

__global__ __launch_bounds__(64) void k_guide_train(GuideView G, TrainSrc S, float* __restrict__ partial) {
    // 38 KB: four blocks per CU.  The outputs and the output deltas reuse
    // sX (the features are computed again for the layer-1 gradient), and
    // each layer's delta overwrites that layer's activations in place.
    __shared__ __align__(16) uint16_t sX[64 * kGStrideIn];
    __shared__ __align__(16) uint16_t sH[3][64 * kGStrideHid];
    static_assert(64 * kGOut * sizeof(float) <= sizeof(sX), "outputs fit the feature rows");
    const int l = threadIdx.x;
    const uint32_t i = blockIdx.x * 64u + (uint32_t)l;
    float3 p = make_float3(0.f, 0.f, 0.f), nn = p, w = p;
    float c = 0.f;
    if (i < S.n) {
        if (S.ts) {
            const uint32_t cnt = min(*S.ts_count, S.ts_cap);
            if (cnt) {
                // position in this pass's subset of `subset` samples, then the
                // subset's fixed uniform choice from the training set
                const uint32_t j = (uint32_t)(g_splitmix64(S.seed ^ ((uint64_t)i << 20)) % S.subset);
                const uint32_t k = (uint32_t)(g_splitmix64(S.subset_seed ^ ((uint64_t)j << 24)) % cnt);
                const float4 a = S.ts[3 * (size_t)k], b = S.ts[3 * (size_t)k + 1], e = S.ts[3 * (size_t)k + 2];
                p = make_float3(a.x, a.y, a.z);
                c = a.w;
                nn = make_float3(b.x, b.y, b.z);
                w = make_float3(e.x, e.y, e.z);
            }
        } else {
            p = load3(S.p3, i);
            nn = load3(S.n3, i);
            w = load3(S.w3, i);
            c = S.c[i];
        }
    }
    // forward, keeping every layer's BF16 activations
    guide_features(G, p, nn, true, sX + l * kGStrideIn);
    __syncthreads();
    guide_hidden<5>(G.wbf + kGW1, G.theta + kGOffB1, sX, kGStrideIn, sH[0], kGStrideHid);
    __syncthreads();
    guide_hidden<4>(G.wbf + kGW2, G.theta + kGOffB2, sH[0], kGStrideHid, sH[1], kGStrideHid);
    __syncthreads();
    guide_hidden<4>(G.wbf + kGW3, G.theta + kGOffB3, sH[1], kGStrideHid, sH[2], kGStrideHid);
    __syncthreads();
    float* sOut = reinterpret_cast<float*>(sX);
    guide_output(G.wbf + kGW4, G.theta + kGOffB4, sH[2], kGStrideHid, sOut);
    __syncthreads();
    float o[kGOut], d[kGOut];
#pragma unroll
    for (int k = 0; k < kGOut; ++k) o[k] = sOut[l * kGOut + k];
    float nll = 0.f;
    if (c != 0.f) {
        nll = guide_dout(o, w, c, d);
    } else {
#pragma unroll
        for (int k = 0; k < kGOut; ++k) d[k] = 0.f;
    }
    float cs = c, cl = c != 0.f ? c * nll : 0.f;
    for (int off = 32; off > 0; off >>= 1) {
        cs += __shfl_xor(cs, off, 64);
        cl += __shfl_xor(cl, off, 64);
    }
    __syncthreads();  // sOut (aliasing sX) fully read
    uint16_t* dO = sX;
#pragma unroll
    for (int k = 0; k < kGOut; ++k) dO[l * kGStrideHid + k] = to_bf16(d[k]);
#pragma unroll
    for (int k = kGOut; k < kGOutPad; ++k) dO[l * kGStrideHid + k] = 0;
    float* P = partial + (size_t)blockIdx.x * kGPartial;
    if (l == 0) {
        P[kGParams] = cs;
        P[kGParams + 1] = cl;
    }
    __syncthreads();
    // layer 4 (64 -> 40)
    weight_grad<3, 4>(dO, kGStrideHid, sH[2], kGStrideHid, P + kGOffW4, kGOut, kGHid);
    bias_grad(dO, kGStrideHid, kGOut, P + kGOffB4);
    // D3 replaces H3 (each element's mask is read by the lane that writes it)
    delta_back<3>(G.wbf + kGWT4, dO, sH[2], sH[2]);
    __syncthreads();
    // layer 3
    weight_grad<4, 4>(sH[2], kGStrideHid, sH[1], kGStrideHid, P + kGOffW3, kGHid, kGHid);
    bias_grad(sH[2], kGStrideHid, kGHid, P + kGOffB3);
    delta_back<4>(G.wbf + kGWT3, sH[2], sH[1], sH[1]);
    __syncthreads();
    // layer 2
    weight_grad<4, 4>(sH[1], kGStrideHid, sH[0], kGStrideHid, P + kGOffW2, kGHid, kGHid);
    bias_grad(sH[1], kGStrideHid, kGHid, P + kGOffB2);
    delta_back<4>(G.wbf + kGWT2, sH[1], sH[0], sH[0]);
    // the features again (dO is no longer read)
    guide_features(G, p, nn, true, sX + l * kGStrideIn);
    __syncthreads();
    // layer 1 (78 -> 64)
    weight_grad<4, 5>(sH[0], kGStrideHid, sX, kGStrideIn, P + kGOffW1, kGHid, kGIn);
    bias_grad(sH[0], kGStrideHid, kGHid, P + kGOffB1);
}

// Two-stage deterministic reduction of the per-block gradient slices:
// stage 1 sums fixed groups of kGroupRows block rows into kGroups rows (and
// reduces the per-block sums of c and c * nll into totals[0..1]); stage 2
// (k_guide_reduce / k_guide_adam) sums the kGroups rows in order.
constexpr uint32_t kGroups = 32;

__global__ void k_guide_stage1(const float* __restrict__ partial, uint32_t nb, float* __restrict__ groups,
                               float* __restrict__ totals) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = blockIdx.y;
    const uint32_t per = (nb + kGroups - 1) / kGroups;
    const uint32_t b0 = g * per, b1 = min(nb, b0 + per);
    if (p < kGParams) {
        float acc = 0.f;
        for (uint32_t b = b0; b < b1; ++b) acc += partial[(size_t)b * kGPartial + p];
        groups[(size_t)g * kGParams + p] = acc;
    }
    if (g == 0 && blockIdx.x == 0) {  // totals: one wave, lane-strided then a fixed shuffle tree
        __shared__ float red[2][256];
        float sc = 0.f, sl = 0.f;
        for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
            sc += partial[(size_t)b * kGPartial + kGParams];
            sl += partial[(size_t)b * kGPartial + kGParams + 1];
        }
        red[0][threadIdx.x] = sc;
        red[1][threadIdx.x] = sl;
        __syncthreads();
        for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                red[0][threadIdx.x] += red[0][threadIdx.x + o];
                red[1][threadIdx.x] += red[1][threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            totals[0] = red[0][0];
            totals[1] = red[1][0];
        }
    }
}

__device__ __forceinline__ float group_sum(const float* groups, uint32_t p) {
    float g = 0.f;
#pragma unroll 8
    for (uint32_t k = 0; k < kGroups; ++k) g += groups[(size_t)k * kGParams + p];
    return g;
}

// grad = (sum of the block slices) / (sum of c); loss = sum(c nll) / sum(c).
__global__ void k_guide_reduce(const float* __restrict__ groups, const float* __restrict__ totals,
                               float* __restrict__ grad, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const float sc = totals[0], sl = totals[1];
    if (p == 0) {
        loss[0] = sc > 0.f ? sl / sc : 0.f;
        loss[1] = sc;
    }
    if (p >= kGParams) return;
    grad[p] = sc > 0.f ? group_sum(groups, p) / sc : 0.f;
}

// Adam (Kingma & Ba 2015) with bias corrections bc1 = 1 - b1^t, bc2 = 1 - b2^t.
// A batch whose weights sum to zero leaves every parameter unchanged.
__global__ void k_guide_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                             const float* __restrict__ groups, const float* __restrict__ totals, float lr,
                             float bc1, float bc2, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const float sc = totals[0], sl = totals[1];
    if (p == 0) {
        loss[0] = sc > 0.f ? sl / sc : 0.f;
        loss[1] = sc;
    }
    if (p >= kGParams || !(sc > 0.f)) return;
    const float g = group_sum(groups, p) / sc;
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    theta[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
}

__global__ void k_zero_u32(uint32_t* p) {
    if (threadIdx.x == 0) *p = 0u;
}

uint64_t g_host_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

int check_guide(tn_guide* g) {
    if (!g) return set_error(TN_ERR_INVALID, "NULL guide");
    TN_HIP(hipSetDevice(g->device));
    return TN_OK;
}

// Per-block slices, then kGroups group rows and two totals.
int ensure_partial(tn_guide* g, uint32_t blocks) {
    if (g->partial_blocks >= blocks) return TN_OK;
    if (g->partial) (void)hipFree(g->partial);
    g->partial = nullptr;
    g->partial_blocks = 0;
    TN_HIP(hipMalloc(&g->partial, ((size_t)blocks * kGPartial + (size_t)kGroups * kGParams + 64) * sizeof(float)));
    g->partial_blocks = blocks;
    return TN_OK;
}

float* group_rows(tn_guide* g) { return g->partial + (size_t)g->partial_blocks * kGPartial; }
float* group_totals(tn_guide* g) { return group_rows(g) + (size_t)kGroups * kGParams; }

int reduce_stage1(tn_guide* g, uint32_t nb, hipStream_t stream) {
    k_guide_stage1<<<dim3((kGParams + 255) / 256, kGroups), 256, 0, stream>>>(g->partial, nb, group_rows(g),
                                                                             group_totals(g));
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int refresh(tn_guide* g, hipStream_t stream) {
    k_guide_refresh<<<(kGBf16 + 255) / 256, 256, 0, stream>>>(g->theta, g->wbf);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int adam_step(tn_guide* g, uint32_t nb, hipStream_t stream) {
    const uint32_t t = ++g->steps;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)t));
    const float bc2 = (float)(1.0 - std::pow(0.999, (double)t));
    int rc = reduce_stage1(g, nb, stream);
    if (rc) return rc;
    k_guide_adam<<<(kGParams + 255) / 256, 256, 0, stream>>>(g->theta, g->adam_m, g->adam_v, group_rows(g),
                                                             group_totals(g), g->lr, bc1, bc2, g->dloss);
    TN_HIP(hipGetLastError());
    return refresh(g, stream);
}

int check_batch(const float* p3, const float* n3, const float* w3, const float* c, uint32_t n) {
    if (n && (!p3 || !n3 || !w3 || !c)) return set_error(TN_ERR_INVALID, "NULL batch array");
    if (n > (1u << 26)) return set_error(TN_ERR_LIMIT, "batch of %u samples exceeds 2^26", n);
    return TN_OK;
}

}  // namespace

int guide_reserve_set(tn_guide* g, uint32_t cap, hipStream_t stream) {
    if (cap > g->ts_cap) {
        if (g->ts) (void)hipFree(g->ts);
        g->ts = nullptr;
        g->ts_cap = 0;
        TN_HIP(hipMalloc(&g->ts, (size_t)cap * 3 * sizeof(float4)));
        g->ts_cap = cap;
    }
    k_zero_u32<<<1, 64, 0, stream>>>(g->ts_count);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int guide_train_pending(tn_guide* g, hipStream_t stream) {
    if (!g->pending) return TN_OK;
    const uint32_t n = g->pending;
    g->pending = 0;
    return guide_train_from_set(g, n, stream);
}

int guide_defer_begin(tn_guide* g, hipStream_t stream, bool recording, bool* started) {
    *started = false;
    if (!g->pending) return TN_OK;
    if (!g->tstream) {
        TN_HIP(hipStreamCreateWithFlags(&g->tstream, hipStreamNonBlocking));
        TN_HIP(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
        TN_HIP(hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming));
    }
    TN_HIP(hipMemcpyAsync(g->wbf_snap, g->wbf, kGBf16 * sizeof(uint16_t), hipMemcpyDeviceToDevice, stream));
    TN_HIP(hipMemcpyAsync(g->theta_snap, g->theta, kGParams * sizeof(float), hipMemcpyDeviceToDevice, stream));
    TN_HIP(hipEventRecord(g->ev_fork, stream));
    TN_HIP(hipStreamWaitEvent(g->tstream, g->ev_fork, 0));
    // the kernels take the set's pointers by value, so the swap below
    // leaves this training on the pending set
    const int rc = guide_train_pending(g, g->tstream);
    if (rc) {
        (void)hipStreamSynchronize(g->tstream);
        return rc;
    }
    TN_HIP(hipEventRecord(g->ev_join, g->tstream));
    if (recording) {
        std::swap(g->ts, g->ts_alt);
        std::swap(g->ts_cap, g->ts_alt_cap);
        std::swap(g->ts_count, g->ts_count_alt);
    }
    *started = true;
    return TN_OK;
}

int guide_train_from_set(tn_guide* g, uint32_t train_samples, hipStream_t stream) {
    const uint32_t nb = g->batch / 64;
    int rc = ensure_partial(g, nb);
    if (rc) return rc;
    // a uniform subset of train_samples recorded samples, `epochs` passes of
    // mini-batches over it
    const uint64_t draws = (uint64_t)train_samples * std::max(1u, g->epochs);
    const uint32_t steps = (uint32_t)std::max<uint64_t>(1, draws / g->batch);
    const uint64_t subset_seed = g_host_mix(g->seed ^ (0xA5A5A5A5ull + g->passes++));
    for (uint32_t s = 0; s < steps; ++s) {
        TrainSrc S{};
        S.n = g->batch;
        S.ts = g->ts;
        S.ts_count = g->ts_count;
        S.ts_cap = g->ts_cap;
        S.subset = std::max(1u, train_samples);
        S.subset_seed = subset_seed;
        S.seed = g->seed * 0x9E3779B97F4A7C15ull + ((uint64_t)g->steps << 32) + s;
        k_guide_train<<<nb, 64, 0, stream>>>(g->view(), S, g->partial);
        TN_HIP(hipGetLastError());
        rc = adam_step(g, nb, stream);
        if (rc) return rc;
    }
    return TN_OK;
}

}  // namespace tn

using namespace tn;

extern "C" {

int tn_guide_create(const float bmin[3], const float bmax[3], uint64_t seed, tn_guide** out) {
    if (!bmin || !bmax || !out) return set_error(TN_ERR_INVALID, "NULL argument");
    *out = nullptr;
    for (int k = 0; k < 3; ++k)
        if (!(bmax[k] >= bmin[k]) || !std::isfinite(bmin[k]) || !std::isfinite(bmax[k]))
            return set_error(TN_ERR_INVALID, "invalid guide box");
    int dev = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(TN_ERR_NODEVICE, "no HIP device available");
    TN_HIP(hipGetDevice(&dev));
    tn_guide* g = new tn_guide();
    g->device = dev;
    g->seed = seed;
    for (int k = 0; k < 3; ++k) {
        g->bmin[k] = bmin[k];
        g->bmax[k] = bmax[k];
    }
    // Uniform(-1/sqrt(in), 1/sqrt(in)) weights, zero biases.
    std::vector<float> theta(kGParams, 0.f);
    std::mt19937_64 rng(seed);
    const int ins[4] = {kGIn, kGHid, kGHid, kGHid}, outs[4] = {kGHid, kGHid, kGHid, kGOut};
    size_t off = 0;
    for (int L = 0; L < 4; ++L) {
        std::uniform_real_distribution<float> U(-1.0f / std::sqrt((float)ins[L]), 1.0f / std::sqrt((float)ins[L]));
        for (int k = 0; k < ins[L] * outs[L]; ++k) theta[off + k] = U(rng);
        off += (size_t)ins[L] * outs[L] + outs[L];
    }
    auto fail = [&](int rc) {
        tn_guide_destroy(g);
        return rc;
    };
    hipError_t e;
    if ((e = hipMalloc(&g->theta, kGParams * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&g->adam_m, kGParams * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&g->adam_v, kGParams * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&g->wbf, kGBf16 * sizeof(uint16_t))) != hipSuccess ||
        (e = hipMalloc(&g->dloss, 4 * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&g->ts_count, 4 * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&g->wbf_snap, kGBf16 * sizeof(uint16_t))) != hipSuccess ||
        (e = hipMalloc(&g->theta_snap, kGParams * sizeof(float))) != hipSuccess)
        return fail(hip_error(e, "hipMalloc"));
    g->ts_count_alt = g->ts_count + 2;  // words 0 and 2 count the two sets
    if ((e = hipMemset(g->dloss, 0, 4 * sizeof(float))) != hipSuccess ||
        (e = hipMemset(g->ts_count, 0, 4 * sizeof(uint32_t))) != hipSuccess)
        return fail(hip_error(e, "hipMemset"));
    int rc = tn_guide_set_params(g, theta.data());
    if (rc) return fail(rc);
    *out = g;
    return TN_OK;
}

int tn_guide_destroy(tn_guide* g) {
    if (!g) return TN_OK;
    (void)hipSetDevice(g->device);
    (void)hipDeviceSynchronize();
    // ts_count and ts_count_alt are two words of one allocation
    uint32_t* counts = std::min(g->ts_count, g->ts_count_alt ? g->ts_count_alt : g->ts_count);
    for (void* p : {(void*)g->theta, (void*)g->adam_m, (void*)g->adam_v, (void*)g->wbf, (void*)g->partial,
                    (void*)g->dloss, (void*)g->ts, (void*)g->ts_alt, (void*)counts, (void*)g->wbf_snap,
                    (void*)g->theta_snap})
        if (p) (void)hipFree(p);
    if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
    if (g->ev_join) (void)hipEventDestroy(g->ev_join);
    if (g->tstream) (void)hipStreamDestroy(g->tstream);
    delete g;
    return TN_OK;
}

int tn_guide_flush(tn_guide* g, void* stream) {
    int rc = check_guide(g);
    if (rc) return rc;
    return guide_train_pending(g, (hipStream_t)stream);
}

int tn_guide_get_params(tn_guide* g, float* theta) {
    int rc = check_guide(g);
    if (rc) return rc;
    if (!theta) return set_error(TN_ERR_INVALID, "NULL argument");
    TN_HIP(hipDeviceSynchronize());
    TN_HIP(hipMemcpy(theta, g->theta, kGParams * sizeof(float), hipMemcpyDeviceToHost));
    return TN_OK;
}

int tn_guide_set_params(tn_guide* g, const float* theta) {
    int rc = check_guide(g);
    if (rc) return rc;
    if (!theta) return set_error(TN_ERR_INVALID, "NULL argument");
    for (int k = 0; k < kGParams; ++k)
        if (!std::isfinite(theta[k])) return set_error(TN_ERR_INVALID, "non-finite parameter %d", k);
    TN_HIP(hipDeviceSynchronize());
    TN_HIP(hipMemcpy(g->theta, theta, kGParams * sizeof(float), hipMemcpyHostToDevice));
    TN_HIP(hipMemset(g->adam_m, 0, kGParams * sizeof(float)));
    TN_HIP(hipMemset(g->adam_v, 0, kGParams * sizeof(float)));
    TN_HIP(hipMemset(g->wbf, 0, kGBf16 * sizeof(uint16_t)));
    g->steps = 0;
    rc = refresh(g, nullptr);
    if (rc) return rc;
    TN_HIP(hipDeviceSynchronize());
    return TN_OK;
}

int tn_guide_set_options(tn_guide* g, float lr, uint32_t batch, float mis_guided, uint32_t epochs) {
    if (!g) return set_error(TN_ERR_INVALID, "NULL guide");
    if (!(lr > 0.f) || !std::isfinite(lr)) return set_error(TN_ERR_INVALID, "learning rate must be > 0");
    if (batch == 0 || batch % 64 || batch > (1u << 24))
        return set_error(TN_ERR_INVALID, "batch must be a positive multiple of 64 up to 2^24");
    if (!(mis_guided >= 0.f && mis_guided <= 1.f))
        return set_error(TN_ERR_INVALID, "mis_guided must be in [0, 1]");
    if (epochs == 0 || epochs > 1024) return set_error(TN_ERR_INVALID, "epochs must be in 1..1024");
    g->lr = lr;
    g->batch = batch;
    g->mis_guided = mis_guided;
    g->epochs = epochs;
    return TN_OK;
}

int tn_guide_forward(tn_guide* g, const float* p3, const float* n3, uint32_t n, float* out, void* stream) {
    int rc = check_guide(g);
    if (rc) return rc;
    if (n && (!p3 || !n3 || !out)) return set_error(TN_ERR_INVALID, "NULL argument");
    if (n > (1u << 30)) return set_error(TN_ERR_LIMIT, "too many points");
    if (!n) return TN_OK;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device);
    const uint32_t groups = (n + 63) / 64;
    const uint32_t grid = std::min<uint32_t>(groups, (uint32_t)cus * 8);  // resident: 8 one-wave blocks per CU
    k_guide_forward<<<grid, 64, 0, (hipStream_t)stream>>>(g->view(), p3, n3, n, out);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int tn_guide_pdf(tn_guide* g, const float* p3, const float* n3, const float* w3, uint32_t n, float* pdf,
                 void* stream) {
    int rc = check_guide(g);
    if (rc) return rc;
    if (n && (!p3 || !n3 || !w3 || !pdf)) return set_error(TN_ERR_INVALID, "NULL argument");
    if (n > (1u << 30)) return set_error(TN_ERR_LIMIT, "too many points");
    if (!n) return TN_OK;
    k_guide_eval<<<(n + 63) / 64, 64, 0, (hipStream_t)stream>>>(g->view(), p3, n3, w3, nullptr, n, 1, nullptr,
                                                                pdf, nullptr);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int tn_guide_sample(tn_guide* g, const float* p3, const float* n3, const float* u3, uint32_t n, float* w3,
                    float* pdf, void* stream) {
    int rc = check_guide(g);
    if (rc) return rc;
    if (n && (!p3 || !n3 || !u3 || !w3 || !pdf)) return set_error(TN_ERR_INVALID, "NULL argument");
    if (n > (1u << 30)) return set_error(TN_ERR_LIMIT, "too many points");
    if (!n) return TN_OK;
    k_guide_eval<<<(n + 63) / 64, 64, 0, (hipStream_t)stream>>>(g->view(), p3, n3, nullptr, u3, n, 2, nullptr,
                                                                pdf, w3);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int tn_guide_loss_grad(tn_guide* g, const float* p3, const float* n3, const float* w3, const float* c,
                       uint32_t n, float* grad, float* loss, void* stream) {
    int rc = check_guide(g);
    if (rc) return rc;
    if ((rc = check_batch(p3, n3, w3, c, n))) return rc;
    if (!grad) return set_error(TN_ERR_INVALID, "NULL gradient");
    if (n == 0) return set_error(TN_ERR_INVALID, "empty batch");
    const uint32_t nb = (n + 63) / 64;
    if ((rc = ensure_partial(g, nb))) return rc;
    hipStream_t s = (hipStream_t)stream;
    TrainSrc S{p3, n3, w3, c, n, nullptr, nullptr, 0, 0, 1, 0};
    k_guide_train<<<nb, 64, 0, s>>>(g->view(), S, g->partial);
    TN_HIP(hipGetLastError());
    if ((rc = reduce_stage1(g, nb, s))) return rc;
    k_guide_reduce<<<(kGParams + 255) / 256, 256, 0, s>>>(group_rows(g), group_totals(g), grad, g->dloss);
    TN_HIP(hipGetLastError());
    if (loss) {
        TN_HIP(hipMemcpyAsync(loss, g->dloss, sizeof(float), hipMemcpyDeviceToHost, s));
        TN_HIP(hipStreamSynchronize(s));
    }
    return TN_OK;
}

int tn_guide_train_recorded(tn_guide* g, uint32_t train_samples, void* stream) {
    int rc = check_guide(g);
    if (rc) return rc;
    if (train_samples == 0) return set_error(TN_ERR_INVALID, "train_samples must be positive");
    if (!g->ts) return set_error(TN_ERR_INVALID, "no recorded training set: render with train_samples > 0 first");
    if ((rc = guide_train_pending(g, (hipStream_t)stream))) return rc;  // a deferred set trains first
    return guide_train_from_set(g, train_samples, (hipStream_t)stream);
}

int tn_guide_train_step(tn_guide* g, const float* p3, const float* n3, const float* w3, const float* c,
                        uint32_t n, float* loss, void* stream) {
    int rc = check_guide(g);
    if (rc) return rc;
    if ((rc = check_batch(p3, n3, w3, c, n))) return rc;
    if (n == 0) return set_error(TN_ERR_INVALID, "empty batch");
    const uint32_t nb = (n + 63) / 64;
    if ((rc = ensure_partial(g, nb))) return rc;
    hipStream_t s = (hipStream_t)stream;
    TrainSrc S{p3, n3, w3, c, n, nullptr, nullptr, 0, 0, 1, 0};
    k_guide_train<<<nb, 64, 0, s>>>(g->view(), S, g->partial);
    TN_HIP(hipGetLastError());
    if ((rc = adam_step(g, nb, s))) return rc;
    if (loss) {
        TN_HIP(hipMemcpyAsync(loss, g->dloss, sizeof(float), hipMemcpyDeviceToHost, s));
        TN_HIP(hipStreamSynchronize(s));
    }
    return TN_OK;
}

int tn_guide_info(tn_guide* g, uint32_t* train_samples, uint32_t* steps, float* last_loss) {
    int rc = check_guide(g);
    if (rc) return rc;
    TN_HIP(hipDeviceSynchronize());
    if (train_samples) {
        uint32_t cnt = 0;
        TN_HIP(hipMemcpy(&cnt, g->ts_count, 4, hipMemcpyDeviceToHost));
        *train_samples = std::min(cnt, g->ts_cap);
    }
    if (steps) *steps = g->steps;
    if (last_loss) TN_HIP(hipMemcpy(last_loss, g->dloss, 4, hipMemcpyDeviceToHost));
    return TN_OK;
}

long tn_guide_get_train_set(tn_guide* g, float* out, size_t cap) {
    int rc = check_guide(g);
    if (rc) return rc;
    TN_HIP(hipDeviceSynchronize());
    uint32_t cnt = 0;
    TN_HIP(hipMemcpy(&cnt, g->ts_count, 4, hipMemcpyDeviceToHost));
    const size_t n = std::min<size_t>(std::min(cnt, g->ts_cap), cap);
    if (n && !out) return set_error(TN_ERR_INVALID, "NULL output");
    if (n) TN_HIP(hipMemcpy(out, g->ts, n * 3 * sizeof(float4), hipMemcpyDeviceToHost));
    return (long)n;
}

}  // extern "C"
