// tn_guide.h -- learned path-guiding proposal: device-side model description
// shared by guide.hip (training / inference kernels) and render.hip (guided
// sampling).  No reference counterpart: the reference samples the hemisphere
// uniformly (lib/sampling.h:20-32).
//
// features(p, n) = [u, sin(2^k pi u_c), cos(2^k pi u_c) for c = x,y,z, k = 0..11, n]
//                  with u = (p - bmin) / (bmax - bmin)                         (78)
// MLP 78 -> 64 -> 64 -> 64 -> 40, ReLU hidden layers, BF16 weights and
// activations as MFMA operands, FP32 accumulation, FP32 master weights.
// out[5j .. 5j+4] = (logit_j, m_j, kraw_j): 8-lobe von Mises-Fisher mixture
// with pi = softmax(logit), mu = m / |m|, kappa = softplus(kraw) + 0.01.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tn {

constexpr int kGIn = 78, kGInPad = 80, kGHid = 64, kGOut = 40, kGOutPad = 48, kGLobes = 8;
constexpr int kGFreq = 12;
constexpr float kKappaMin = 0.01f, kKappaMax = 1.0e4f;
// flat FP32 parameter layout (per layer W (out x in, row major) then b)
constexpr int kGOffW1 = 0, kGOffB1 = 64 * 78;
constexpr int kGOffW2 = kGOffB1 + 64, kGOffB2 = kGOffW2 + 64 * 64;
constexpr int kGOffW3 = kGOffB2 + 64, kGOffB3 = kGOffW3 + 64 * 64;
constexpr int kGOffW4 = kGOffB3 + 64, kGOffB4 = kGOffW4 + 40 * 64;
constexpr int kGParams = kGOffB4 + 40;  // 15976
// BF16 copies, padded: W_l as [out_pad][in_pad] (forward A operand) and
// W_l^T as [in_pad][out_pad] (backward A operand)
constexpr int kGW1 = 0, kGW2 = kGW1 + 64 * kGInPad, kGW3 = kGW2 + 64 * 64, kGW4 = kGW3 + 64 * 64;
constexpr int kGWT2 = kGW4 + kGOutPad * 64, kGWT3 = kGWT2 + 64 * 64, kGWT4 = kGWT3 + 64 * 64;
constexpr int kGBf16 = kGWT4 + 64 * kGOutPad;

struct GuideView {
    const uint16_t* wbf;  // BF16 weight copies (layout above)
    const float* theta;   // FP32 master parameters (biases are read from here)
    float bmin[3], inv_ext[3];
};

__host__ __device__ __forceinline__ float softplusf(float x) {
    return x > 20.f ? x : log1pf(expf(x));
}

// vMF density on the sphere, stable for small and large kappa.
__device__ __forceinline__ float vmf_pdf(float cosang, float kappa) {
    return kappa / (6.28318530717958647692f * -expm1f(-2.0f * kappa)) * expf(kappa * (cosang - 1.0f));
}

struct Mixture {
    float pi[kGLobes];
    float mu[kGLobes][3];
    float kappa[kGLobes];
};

__device__ __forceinline__ void mixture_from(const float* out, Mixture& M) {
    float mx = out[0];
    for (int j = 1; j < kGLobes; ++j) mx = fmaxf(mx, out[5 * j]);
    float sum = 0.f;
    for (int j = 0; j < kGLobes; ++j) {
        M.pi[j] = expf(out[5 * j] - mx);
        sum += M.pi[j];
        const float mx3 = out[5 * j + 1], my = out[5 * j + 2], mz = out[5 * j + 3];
        const float inv = 1.0f / fmaxf(sqrtf(mx3 * mx3 + my * my + mz * mz), 1e-12f);
        M.mu[j][0] = mx3 * inv;
        M.mu[j][1] = my * inv;
        M.mu[j][2] = mz * inv;
        M.kappa[j] = fminf(softplusf(out[5 * j + 4]) + kKappaMin, kKappaMax);
    }
    const float is = 1.0f / sum;
    for (int j = 0; j < kGLobes; ++j) M.pi[j] *= is;
}

__device__ __forceinline__ float mixture_pdf(const Mixture& M, float3 w) {
    float q = 0.f;
    for (int j = 0; j < kGLobes; ++j)
        q += M.pi[j] * vmf_pdf(M.mu[j][0] * w.x + M.mu[j][1] * w.y + M.mu[j][2] * w.z, M.kappa[j]);
    return q;
}

// Lobe by the CDF of pi (u0: the first lobe whose cumulative weight exceeds
// u0, else the last), vMF cosine by inverse CDF (u1), azimuth (u2).  The
// lobe is selected without dynamic indexing so M stays in registers.
__device__ __forceinline__ float3 mixture_sample(const Mixture& M, float u0, float u1, float u2) {
    float c = 0.f, k = M.kappa[kGLobes - 1];
    float mx = M.mu[kGLobes - 1][0], my = M.mu[kGLobes - 1][1], mz = M.mu[kGLobes - 1][2];
    bool found = false;
#pragma unroll
    for (int q = 0; q < kGLobes - 1; ++q) {
        c += M.pi[q];
        const bool take = !found && u0 < c;
        k = take ? M.kappa[q] : k;
        mx = take ? M.mu[q][0] : mx;
        my = take ? M.mu[q][1] : my;
        mz = take ? M.mu[q][2] : mz;
        found = found || take;
    }
    const float cw = fmaxf(-1.0f, fminf(1.0f, 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * k)) / k));
    const float sw = sqrtf(fmaxf(0.f, 1.f - cw * cw));
    float sp, cp;
    __sincosf(6.28318530717958647692f * u2, &sp, &cp);
    const float sign = copysignf(1.0f, mz);
    const float a = -1.0f / (sign + mz);
    const float b = mx * my * a;
    const float3 b1 = make_float3(1.0f + sign * mx * mx * a, sign * b, -sign * mx);
    const float3 b2 = make_float3(b, sign + my * my * a, -my);
    return make_float3(cw * mx + sw * (cp * b1.x + sp * b2.x), cw * my + sw * (cp * b1.y + sp * b2.y),
                       cw * mz + sw * (cp * b1.z + sp * b2.z));
}

typedef __attribute__((ext_vector_type(4))) float g_f4;
typedef __attribute__((ext_vector_type(4))) short g_s4;

// Two FP32 values -> packed BF16 (a low, b high), round to nearest even:
// one v_cvt_pk_bf16_f32 on gfx950.
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    typedef __attribute__((ext_vector_type(2))) float f2;
    typedef __attribute__((ext_vector_type(2))) __bf16 h2;
    const h2 h = __builtin_convertvector((f2){a, b}, h2);
    uint32_t r;
    __builtin_memcpy(&r, &h, 4);
    return r;
}

// FP32 -> BF16, round to nearest even (finite x; the host form rounds the
// bits, the device form converts in hardware, with equal results).
__host__ __device__ __forceinline__ uint16_t to_bf16(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint16_t)pack_bf16x2(x, 0.f);
#else
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
#endif
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kGStrideIn = kGInPad + 8;  // LDS row strides (u16), padded against bank conflicts
constexpr int kGStrideHid = kGHid + 8;

// Features of one sample (lane) into an LDS row of kGInPad BF16 values.
__device__ __forceinline__ void guide_features(const GuideView& G, float3 p, float3 n, bool active,
                                               uint16_t* row) {
    float u[3] = {(p.x - G.bmin[0]) * G.inv_ext[0], (p.y - G.bmin[1]) * G.inv_ext[1],
                  (p.z - G.bmin[2]) * G.inv_ext[2]};
    if (!active) u[0] = u[1] = u[2] = 0.f;
    row[0] = to_bf16(u[0]);
    row[1] = to_bf16(u[1]);
    row[2] = to_bf16(u[2]);
    // sin / cos of 2^q pi u = of 2 pi (2^(q-1) u): the hardware sine and
    // cosine take revolutions; the argument's fraction (exact, as is the
    // power-of-two scaling) keeps them in range.  Absolute error about
    // 1.3e-7 (sincospif: 5e-8), far below the BF16 step.
    int k = 3;
    for (int c = 0; c < 3; ++c) {
        float f = 0.5f;
        for (int q = 0; q < kGFreq; ++q) {
            const float r = __builtin_amdgcn_fractf(f * u[c]);
            const uint32_t sc = pack_bf16x2(__builtin_amdgcn_sinf(r), __builtin_amdgcn_cosf(r));
            row[k++] = (uint16_t)sc;
            row[k++] = (uint16_t)(sc >> 16);
            f *= 2.0f;
        }
    }
    row[75] = to_bf16(active ? n.x : 0.f);
    row[76] = to_bf16(active ? n.y : 0.f);
    row[77] = to_bf16(active ? n.z : 0.f);
    row[78] = 0;
    row[79] = 0;
}

// Y^T = W X^T on the wave's 64 samples (4 tiles of 16) with
// mfma_f32_16x16x16bf16_1k.  A = W tile (lane: row 16mt + l%16, columns
// 16kt + 4(l/16) .. +3 of a row-major [16 OUT_TILES][16 IN_TILES] matrix),
// B = X^T tile read from the [sample][feature] LDS rows (lane: sample
// 16nt + l%16, features 16kt + 4(l/16) .. +3).  The accumulator of lane l
// holds outputs 16mt + 4(l/16) .. +3 of sample 16nt + l%16 -- four
// consecutive features of one row -- handed to epi(first output, sample, v).
template <int IN_TILES, int OUT_TILES, class Epi>
__device__ __forceinline__ void mfma_wx(const uint16_t* __restrict__ W, const uint16_t* in, int in_stride,
                                        Epi epi) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, grp = l >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        g_s4 bfrag[IN_TILES];
#pragma unroll
        for (int kt = 0; kt < IN_TILES; ++kt)
            bfrag[kt] = *reinterpret_cast<const g_s4*>(in + (16 * nt + row) * in_stride + 16 * kt + 4 * grp);
#pragma unroll
        for (int mt = 0; mt < OUT_TILES; ++mt) {
            g_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < IN_TILES; ++kt) {
                const g_s4 a =
                    *reinterpret_cast<const g_s4*>(W + (16 * mt + row) * (16 * IN_TILES) + 16 * kt + 4 * grp);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bfrag[kt], acc, 0, 0, 0);
            }
            epi(16 * mt + 4 * grp, 16 * nt + row, acc);
        }
    }
}

// Same product with the A fragments already in registers (Wf[mt][kt]).
template <int IN_TILES, int OUT_TILES, class Epi>
__device__ __forceinline__ void mfma_wx_reg(const g_s4 (&Wf)[OUT_TILES][IN_TILES], const uint16_t* in,
                                            int in_stride, Epi epi) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, grp = l >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        g_s4 bfrag[IN_TILES];
#pragma unroll
        for (int kt = 0; kt < IN_TILES; ++kt)
            bfrag[kt] = *reinterpret_cast<const g_s4*>(in + (16 * nt + row) * in_stride + 16 * kt + 4 * grp);
#pragma unroll
        for (int mt = 0; mt < OUT_TILES; ++mt) {
            g_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < IN_TILES; ++kt)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Wf[mt][kt], bfrag[kt], acc, 0, 0, 0);
            epi(16 * mt + 4 * grp, 16 * nt + row, acc);
        }
    }
}

// Loads the A fragments of a [16 OUT_TILES][16 IN_TILES] BF16 matrix.
template <int IN_TILES, int OUT_TILES>
__device__ __forceinline__ void load_frags(const uint16_t* __restrict__ W, g_s4 (&Wf)[OUT_TILES][IN_TILES]) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, grp = l >> 4;
#pragma unroll
    for (int mt = 0; mt < OUT_TILES; ++mt)
#pragma unroll
        for (int kt = 0; kt < IN_TILES; ++kt)
            Wf[mt][kt] = *reinterpret_cast<const g_s4*>(W + (16 * mt + row) * (16 * IN_TILES) + 16 * kt + 4 * grp);
}

typedef __attribute__((ext_vector_type(8))) __bf16 g_bf8;

// K = 32 tiles with gfx950's mfma_f32_16x16x32_bf16: A lane holds row
// 16mt + l%16, columns 32kt + 8(l/16) .. +7; B lane holds sample 16nt + l%16,
// features 32kt + 8(l/16) .. +7; the accumulator layout is the 16x16x16 one.
// KT32 tiles of 32 then, if TAIL, one 16x16x16 tile for columns 32 KT32 .. +15.
template <int KT32, bool TAIL, int OUT_TILES>
struct Frags32 {
    g_bf8 a[OUT_TILES][KT32];
    g_s4 t[OUT_TILES][TAIL ? 1 : 1];
};

template <int KT32, bool TAIL, int OUT_TILES>
__device__ __forceinline__ void load_frags32(const uint16_t* __restrict__ W, Frags32<KT32, TAIL, OUT_TILES>& F) {
    constexpr int kCols = 32 * KT32 + (TAIL ? 16 : 0);
    const int l = threadIdx.x & 63;
    const int row = l & 15, grp = l >> 4;
#pragma unroll
    for (int mt = 0; mt < OUT_TILES; ++mt) {
#pragma unroll
        for (int kt = 0; kt < KT32; ++kt)
            F.a[mt][kt] = *reinterpret_cast<const g_bf8*>(W + (16 * mt + row) * kCols + 32 * kt + 8 * grp);
        if (TAIL) F.t[mt][0] = *reinterpret_cast<const g_s4*>(W + (16 * mt + row) * kCols + 32 * KT32 + 4 * grp);
    }
}

template <int KT32, bool TAIL, int OUT_TILES, class Epi>
__device__ __forceinline__ void mfma_wx32(const Frags32<KT32, TAIL, OUT_TILES>& F, const uint16_t* in,
                                          int in_stride, Epi epi) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, grp = l >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const uint16_t* x = in + (16 * nt + row) * in_stride;
        g_bf8 b[KT32];
#pragma unroll
        for (int kt = 0; kt < KT32; ++kt) b[kt] = *reinterpret_cast<const g_bf8*>(x + 32 * kt + 8 * grp);
        g_s4 bt = {0, 0, 0, 0};
        if (TAIL) bt = *reinterpret_cast<const g_s4*>(x + 32 * KT32 + 4 * grp);
#pragma unroll
        for (int mt = 0; mt < OUT_TILES; ++mt) {
            g_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < KT32; ++kt) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F.a[mt][kt], b[kt], acc, 0, 0, 0);
            if (TAIL) {  // the K = 16 tail accumulates separately: no mixed-shape accumulation chain
                g_f4 t = {0.f, 0.f, 0.f, 0.f};
                t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F.t[mt][0], bt, t, 0, 0, 0);
                acc += t;
            }
            epi(16 * mt + 4 * grp, 16 * nt + row, acc);
        }
    }
}

// mfma_wx32 with the A fragments read from W (global, L2-resident) inside
// the loop instead of held in registers: the same MFMA sequence, so the
// same sums, for kernels that cannot spare 128 VGPRs for the weights.
template <int KT32, bool TAIL, int OUT_TILES, class Epi>
__device__ __forceinline__ void mfma_wx32_mem(const uint16_t* __restrict__ W, const uint16_t* in, int in_stride,
                                              Epi epi) {
    constexpr int kCols = 32 * KT32 + (TAIL ? 16 : 0);
    const int l = threadIdx.x & 63;
    const int row = l & 15, grp = l >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const uint16_t* x = in + (16 * nt + row) * in_stride;
        g_bf8 b[KT32];
#pragma unroll
        for (int kt = 0; kt < KT32; ++kt) b[kt] = *reinterpret_cast<const g_bf8*>(x + 32 * kt + 8 * grp);
        g_s4 bt = {0, 0, 0, 0};
        if (TAIL) bt = *reinterpret_cast<const g_s4*>(x + 32 * KT32 + 4 * grp);
#pragma unroll
        for (int mt = 0; mt < OUT_TILES; ++mt) {
            const uint16_t* w = W + (16 * mt + row) * kCols;
            g_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < KT32; ++kt)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const g_bf8*>(w + 32 * kt + 8 * grp),
                                                              b[kt], acc, 0, 0, 0);
            if (TAIL) {
                g_f4 t = {0.f, 0.f, 0.f, 0.f};
                t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(*reinterpret_cast<const g_s4*>(w + 32 * KT32 + 4 * grp),
                                                              bt, t, 0, 0, 0);
                acc += t;
            }
            epi(16 * mt + 4 * grp, 16 * nt + row, acc);
        }
    }
}


// Epilogue of a hidden layer: bias, ReLU, BF16, one 8-byte LDS store of the
// four consecutive features.
struct HiddenEpi {
    const float* bias;
    uint16_t* out;
    int stride;
    __device__ __forceinline__ void operator()(int o, int smp, g_f4 v) const {
        const float4 b = *reinterpret_cast<const float4*>(bias + o);
        uint2 w;
        w.x = pack_bf16x2(fmaxf(v[0] + b.x, 0.f), fmaxf(v[1] + b.y, 0.f));
        w.y = pack_bf16x2(fmaxf(v[2] + b.z, 0.f), fmaxf(v[3] + b.w, 0.f));
        *reinterpret_cast<uint2*>(out + smp * stride + o) = w;
    }
};

// Epilogue of the output layer: 40 FP32 outputs per sample at out[smp * 40].
struct OutputEpi {
    const float* bias;
    float* out;
    __device__ __forceinline__ void operator()(int o, int smp, g_f4 v) const {
        if (o >= kGOut) return;  // padded rows 40..47
        const float4 b = *reinterpret_cast<const float4*>(bias + o);
        *reinterpret_cast<float4*>(out + smp * kGOut + o) =
            make_float4(v[0] + b.x, v[1] + b.y, v[2] + b.z, v[3] + b.w);
    }
};

// Hidden layer: bias, ReLU, BF16 into the next [sample][feature] rows.
template <int IN_TILES>
__device__ __forceinline__ void guide_hidden(const uint16_t* W, const float* bias, const uint16_t* in,
                                             int in_stride, uint16_t* out, int out_stride) {
    mfma_wx<IN_TILES, 4>(W, in, in_stride, HiddenEpi{bias, out, out_stride});
}

// Output layer: 40 FP32 outputs per sample into out[sample * 40 + o].
__device__ __forceinline__ void guide_output(const uint16_t* W, const float* bias, const uint16_t* in,
                                             int in_stride, float* out) {
    mfma_wx<4, 3>(W, in, in_stride, OutputEpi{bias, out});
}

// The A fragments of all four layers (K = 80 as 2 x 32 + 16, then 2 x 32):
// 128 VGPRs, loaded once per wave and reused for every 64 samples.
struct GuideFrags {
    Frags32<2, true, 4> w1;
    Frags32<2, false, 4> w2, w3;
    Frags32<2, false, 3> w4;
};

__device__ __forceinline__ void load_guide_frags(const GuideView& G, GuideFrags& F) {
    load_frags32(G.wbf + kGW1, F.w1);
    load_frags32(G.wbf + kGW2, F.w2);
    load_frags32(G.wbf + kGW3, F.w3);
    load_frags32(G.wbf + kGW4, F.w4);
}

// Runs the MLP for the 64 lanes' samples of one wave (the block must be that
// one wave: the layers are separated by block barriers).  Leaves the 40
// outputs of lane l's sample at out[l * 40 ..] (out may alias the first
// 64 * kGStrideIn uint16 of lds).  `lds` holds 64 * (kGStrideIn +
// kGStrideHid) uint16 values.  Every forward evaluation of the guide (render,
// pdf, sampling, tn_guide_forward) runs this arithmetic, so a density the
// renderer sampled with is the density tn_guide_pdf reports.
__device__ __forceinline__ void guide_forward_wave32(const GuideView& G, const GuideFrags& F, float3 p,
                                                     float3 n, bool active, float* out, uint16_t* lds) {
    const int l = threadIdx.x & 63;
    uint16_t* a = lds;                    // [64][kGStrideIn]
    uint16_t* b = lds + 64 * kGStrideIn;  // [64][kGStrideHid]
    guide_features(G, p, n, active, a + l * kGStrideIn);
    __syncthreads();
    mfma_wx32(F.w1, a, kGStrideIn, HiddenEpi{G.theta + kGOffB1, b, kGStrideHid});
    __syncthreads();
    mfma_wx32(F.w2, b, kGStrideHid, HiddenEpi{G.theta + kGOffB2, a, kGStrideIn});
    __syncthreads();
    mfma_wx32(F.w3, a, kGStrideIn, HiddenEpi{G.theta + kGOffB3, b, kGStrideHid});
    __syncthreads();
    mfma_wx32(F.w4, b, kGStrideHid, OutputEpi{G.theta + kGOffB4, out});
    __syncthreads();
}

// As guide_forward_wave32 with the weights read from memory layer by layer
// (same sums, far fewer registers: for the shade and evaluation kernels).
__device__ __forceinline__ void guide_forward_wave(const GuideView& G, float3 p, float3 n, bool active,
                                                   float* out, uint16_t* lds) {
    const int l = threadIdx.x & 63;
    uint16_t* a = lds;                    // [64][kGStrideIn]
    uint16_t* b = lds + 64 * kGStrideIn;  // [64][kGStrideHid]
    guide_features(G, p, n, active, a + l * kGStrideIn);
    __syncthreads();
    mfma_wx32_mem<2, true, 4>(G.wbf + kGW1, a, kGStrideIn, HiddenEpi{G.theta + kGOffB1, b, kGStrideHid});
    __syncthreads();
    mfma_wx32_mem<2, false, 4>(G.wbf + kGW2, b, kGStrideHid, HiddenEpi{G.theta + kGOffB2, a, kGStrideIn});
    __syncthreads();
    mfma_wx32_mem<2, false, 4>(G.wbf + kGW3, a, kGStrideIn, HiddenEpi{G.theta + kGOffB3, b, kGStrideHid});
    __syncthreads();
    mfma_wx32_mem<2, false, 3>(G.wbf + kGW4, b, kGStrideHid, OutputEpi{G.theta + kGOffB4, out});
    __syncthreads();
}

}  // namespace tn

// Guide state behind tn_guide* (device buffers owned).
struct tn_guide {
    int device = 0;
    float* theta = nullptr;   // FP32 master parameters
    float* adam_m = nullptr;  // Adam moments
    float* adam_v = nullptr;
    uint16_t* wbf = nullptr;  // BF16 operand copies
    float* partial = nullptr; // per-block gradient slices
    uint32_t partial_blocks = 0;
    float* dloss = nullptr;   // [0] loss of the last batch, [1] its sum of weights
    float4* ts = nullptr;     // training set: 3 float4 per sample (p, c), (n, cos/pdf), (w, tag)
    uint32_t ts_cap = 0;
    uint32_t* ts_count = nullptr;
    // Deferred training (TN_RENDER_DEFER_TRAINING): the records of the last
    // render wait in ts (pending = its train_samples) and train during the
    // next render, on tstream, while that render samples from a snapshot of
    // the weights and records into the other set (ts_alt; the two swap).
    float4* ts_alt = nullptr;
    uint32_t ts_alt_cap = 0;
    uint32_t* ts_count_alt = nullptr;
    uint32_t pending = 0;
    uint16_t* wbf_snap = nullptr;
    float* theta_snap = nullptr;
    hipStream_t tstream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    float bmin[3] = {0, 0, 0}, bmax[3] = {1, 1, 1};
    uint64_t seed = 1;
    uint32_t steps = 0;
    float lr = 1e-3f;
    uint32_t batch = 16384;
    float mis_guided = 0.5f;
    uint32_t epochs = 1;   // passes over each training pass's subset of samples
    uint32_t passes = 0;   // training passes so far
    uint32_t last_samples = 0;

    tn::GuideView view() const {
        tn::GuideView v;
        v.wbf = wbf;
        v.theta = theta;
        for (int k = 0; k < 3; ++k) {
            v.bmin[k] = bmin[k];
            const float ext = bmax[k] - bmin[k];
            v.inv_ext[k] = ext > 0 ? 1.0f / ext : 0.f;
        }
        return v;
    }
};

namespace tn {
// Grows the training set to hold `cap` samples and clears it.
int guide_reserve_set(tn_guide* g, uint32_t cap, hipStream_t stream);
// train_samples / batch Adam steps on mini-batches drawn uniformly (with
// replacement) from the training set.
int guide_train_from_set(tn_guide* g, uint32_t train_samples, hipStream_t stream);
// Trains the pending (deferred) set now, on stream.
int guide_train_pending(tn_guide* g, hipStream_t stream);
// Starts the pending set's training on the guide's stream after stream's
// current point, with a snapshot of the weights taken first on stream for
// the render to sample from; recording: the render's records go to the
// other set.  *started is false when nothing was pending.
int guide_defer_begin(tn_guide* g, hipStream_t stream, bool recording, bool* started);
}  // namespace tn
