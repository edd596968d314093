// capi.hip -- scene management, batched intersection and visibility entry
// points of libturner_amd (C ABI declared in include/turner_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <type_traits>

#include "tn_scene.h"

namespace tn {

static thread_local std::string g_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int hip_error(hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory)
        return set_error(TN_ERR_OOM, "%s: %s", what, hipGetErrorString(e));
    return set_error(TN_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

namespace {

// Batched traversal with lane refill: each wave owns a chunk of kChunk rays
// and, after every leaf step, hands lanes whose ray finished (or missed the
// root box) the chunk's next rays (ballot + mbcnt), so a wave is not held by
// its slowest ray.  ANY = visibility query against tmax (tn_occluded).
template <int SCHED, bool EXH, bool MT, bool ANY, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_rays(SceneView sc, const float* __restrict__ o3,
                                                 const float* __restrict__ d3,
                                                 const float* __restrict__ tmax, uint32_t n,
                                                 uint32_t chunk, uint32_t* __restrict__ out_id,
                                                 float* __restrict__ out_r,
                                                 float* __restrict__ out_s,
                                                 float* __restrict__ out_t,
                                                 uint8_t* __restrict__ occ,
                                                 unsigned long long* counters, uint32_t* status,
                                                 uint32_t* /*work: cooperative kernel only*/) {
    __shared__ uint2 stk[kStack * kBlock];
    __shared__ uint32_t red[8];
    if (COUNT && threadIdx.x < 8) red[threadIdx.x] = 0;
    if (COUNT) __syncthreads();
    uint2* my = stk + threadIdx.x;
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of chunks so rays that are adjacent in the batch share
    // an L2.  The grid is a multiple of 8 blocks, so this is a bijection.
    const uint32_t per_xcd = gridDim.x / kXcds;
    const uint32_t lblock = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    const uint32_t wave = (lblock * kBlock + threadIdx.x) >> 6;
    uint32_t next = wave * chunk;
    const uint32_t end = next < n ? min(next + chunk, n) : next;
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t ray = kNone;
    Trav T;
    float tmx = 0.f;
    TraversalCounters cnt;
    uint32_t nrays = 0;
    for (;;) {
        const uint64_t idle = __ballot(ray == kNone);
        if (idle && next < end) {
            const uint32_t k = next + lanes_below(idle);
            next += (uint32_t)__popcll(idle);
            if (ray == kNone && k < end) {
                const float3 o = make_float3(o3[3 * (size_t)k], o3[3 * (size_t)k + 1], o3[3 * (size_t)k + 2]);
                const float3 d = make_float3(d3[3 * (size_t)k], d3[3 * (size_t)k + 1], d3[3 * (size_t)k + 2]);
                if (ANY) tmx = tmax[k];
                if (COUNT) nrays++;
                if (trav_init(T, sc, o, d)) {
                    ray = k;
                } else if (ANY) {
                    occ[k] = 0;
                } else {
                    out_id[k] = kInvalidId;
                    out_r[k] = FLT_MAX;
                    out_s[k] = 0.f;
                    out_t[k] = 0.f;
                }
            }
        }
        if (__ballot(ray != kNone) == 0) {
            if (next >= end) break;
            continue;
        }
        if (ray != kNone) {
            if (trav_advance<SCHED, kStack, EXH, MT, ANY, COUNT>(T, sc, tmx, my, kBlock, cnt)) {
                const bool hit = T.best_id != kInvalidId;
                if (ANY) {
                    occ[ray] = hit ? 1 : 0;
                } else {
                    out_id[ray] = T.best_id;
                    out_r[ray] = hit ? T.best_r : FLT_MAX;
                    out_s[ray] = hit ? T.best_s : 0.f;
                    out_t[ray] = hit ? T.best_t : 0.f;
                }
                if (COUNT) {
                    cnt.restarts += T.restarts;
                    atomicMax(&red[7], T.visits);  // leaves of the longest traversal
                }
                ray = kNone;
            }
        }
    }
    if (cnt.incomplete) atomicAdd(status, 1u);
    if (COUNT) {
        atomicAdd(&red[0], cnt.inner);
        atomicAdd(&red[1], cnt.leaves);
        atomicAdd(&red[2], cnt.refs);
        atomicAdd(&red[3], cnt.tests);
        atomicAdd(&red[4], nrays);
        atomicAdd(&red[5], cnt.restarts);
        atomicAdd(&red[6], cnt.full);
        __syncthreads();
        if (threadIdx.x < 7) atomicAdd(&counters[threadIdx.x], (unsigned long long)red[threadIdx.x]);
        if (threadIdx.x == 7) atomicMax(&counters[7], (unsigned long long)red[7]);
    }
}

// Rays of one wave of the cooperative kernel.  Without `work`: one range of
// `size` rays per wave (XCD-contiguous numbering as k_rays).  With `work` (a
// zeroed slot: 8 per-XCD counters and a finished-wave counter, 16 words
// apart): chunks of `size` rays, the waves of XCD x taking chunks x, x + 8,
// ... from counter x until they run out; the last wave to finish clears the
// slot for the next launch.  Called by all lanes of the wave.
struct RayChunks {
    uint32_t* work;
    uint32_t n, size, nwaves, first, x;
    bool given;
    __device__ bool operator()(uint32_t& b, uint32_t& e) {
        if (!work) {
            if (given || first >= n) return false;
            given = true;
            b = first;
            e = min(first + size, n);
            return true;
        }
        uint32_t k = 0;
        if ((threadIdx.x & 63u) == 0) k = atomicAdd(work + 16 * x, 1u);
        const uint64_t idx = x + (uint64_t)kXcds * (uint32_t)__shfl((int)k, 0, 64);
        if (idx * size < n) {
            b = (uint32_t)(idx * size);
            e = min(b + size, n);
            return true;
        }
        if ((threadIdx.x & 63u) == 0) {
            __threadfence();  // this wave's last chunk claim precedes its finish count
            if (atomicAdd(work + 16 * kXcds, 1u) == nwaves - 1) {
                for (uint32_t i = 0; i < kXcds; ++i) atomicExch(work + 16 * i, 0u);
                atomicExch(work + 16 * kXcds, 0u);
            }
        }
        return false;
    }
};

// Wave-cooperative variant of k_rays (coop_loop): rays by RayChunks, pooled
// leaf tests.  WPE caps the VGPRs so that WPE waves fit on each SIMD.
// Measured best on 2^24 rays against 1.31 M triangles, 512-ray chunks:
// 5: 1048, 6: 1052, 7: 1143, 8: 1022 Mrays/s.
constexpr int kCoopWavesPerEU = 7;
// REF: traversal of the refined tree of a TN_SCENE_REFINED scene; a ray tied
// at its closest distance (closest hit) or with a zero direction component is
// marked for k_fix_rays (id kFixId, occ 2) instead of being stored.
template <bool EXH, bool MT, bool ANY, bool COUNT, int WPE = kCoopWavesPerEU, bool REF = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_rays_coop(
    SceneView sc, const float* __restrict__ o3, const float* __restrict__ d3,
    const float* __restrict__ tmax, uint32_t n, uint32_t chunk, uint32_t* __restrict__ out_id,
    float* __restrict__ out_r, float* __restrict__ out_s, float* __restrict__ out_t,
    uint8_t* __restrict__ occ, unsigned long long* counters, uint32_t* status, uint32_t* work) {
    constexpr bool TIE = REF && !ANY;
    using WL = std::conditional_t<TIE, CoopLdsTie, CoopLds>;
    __shared__ uint2 stk[kStack * kBlock];
    __shared__ WL wl[kBlock / 64];
    __shared__ uint32_t red[8];  // [0..6] sums, [7] longest traversal in leaves
    if (COUNT && threadIdx.x < 8) red[threadIdx.x] = 0;
    if (COUNT) __syncthreads();
    const uint32_t per_xcd = gridDim.x / kXcds;
    const uint32_t lblock = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    const uint32_t wave = (lblock * kBlock + threadIdx.x) >> 6;
    uint32_t nrays = 0;
    TraversalCounters cnt;
    coop_loop<kStack, EXH, MT, ANY, COUNT, TIE>(
        sc, stk + threadIdx.x, kBlock, wl + threadIdx.x / 64,
        RayChunks{work, n, chunk, gridDim.x * (kBlock / 64), wave * chunk, blockIdx.x % kXcds, false},
        // rays and results stream through once: non-temporal loads and
        // stores keep them from evicting tree nodes and records from L2
        [&](uint32_t k, float3& o, float3& d, float& tm, uint32_t&) {
            const float* po = o3 + 3 * (size_t)k;
            const float* pd = d3 + 3 * (size_t)k;
            o = make_float3(__builtin_nontemporal_load(po), __builtin_nontemporal_load(po + 1),
                            __builtin_nontemporal_load(po + 2));
            d = make_float3(__builtin_nontemporal_load(pd), __builtin_nontemporal_load(pd + 1),
                            __builtin_nontemporal_load(pd + 2));
            if (ANY) tm = __builtin_nontemporal_load(tmax + k);
            if (COUNT) nrays++;
        },
        [&](uint32_t k, const Trav& T, bool traversed) {
            const bool hit = traversed && T.best_id != kInvalidId;
            if (COUNT) atomicMax(&red[7], T.visits);  // leaves of the longest traversal
            if (REF && (needs_reference(T.d) || (TIE && T.tie))) {
                if (ANY) occ[k] = 2;
                else out_id[k] = kFixId;
                return;
            }
            if (ANY) {
                __builtin_nontemporal_store((uint8_t)(hit ? 1 : 0), occ + k);
            } else {
                __builtin_nontemporal_store(hit ? T.best_id : kInvalidId, out_id + k);
                __builtin_nontemporal_store(hit ? T.best_r : FLT_MAX, out_r + k);
                __builtin_nontemporal_store(hit ? T.best_s : 0.f, out_s + k);
                __builtin_nontemporal_store(hit ? T.best_t : 0.f, out_t + k);
            }
        },
        cnt);
    if (cnt.incomplete) atomicAdd(status, 1u);
    if (COUNT) {
        atomicAdd(&red[0], cnt.inner);
        atomicAdd(&red[1], cnt.leaves);
        atomicAdd(&red[2], cnt.refs);
        atomicAdd(&red[3], cnt.tests);
        atomicAdd(&red[4], nrays);
        atomicAdd(&red[5], cnt.restarts);
        atomicAdd(&red[6], cnt.full);
        __syncthreads();
        if (threadIdx.x < 7) atomicAdd(&counters[threadIdx.x], (unsigned long long)red[threadIdx.x]);
        if (threadIdx.x == 7) atomicMax(&counters[7], (unsigned long long)red[7]);
    }
}

// Rays k_rays_coop<..., REF> marked (id kFixId / occ 2), traversed one per
// lane on the reference tree.
template <bool EXH, bool ANY>
__global__ __launch_bounds__(kBlock) void k_fix_rays(SceneView sc, const float* __restrict__ o3,
                                                    const float* __restrict__ d3, const float* __restrict__ tmax,
                                                    uint32_t n, uint32_t* __restrict__ out_id,
                                                    float* __restrict__ out_r, float* __restrict__ out_s,
                                                    float* __restrict__ out_t, uint8_t* __restrict__ occ,
                                                    uint32_t* status) {
    __shared__ uint2 stk[kStack * kBlock];
    TraversalCounters cnt;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (ANY ? occ[k] != 2 : out_id[k] != kFixId) continue;
        const float3 o = make_float3(o3[3 * (size_t)k], o3[3 * (size_t)k + 1], o3[3 * (size_t)k + 2]);
        const float3 d = make_float3(d3[3 * (size_t)k], d3[3 * (size_t)k + 1], d3[3 * (size_t)k + 2]);
        float r = FLT_MAX, s = 0.f, t = 0.f;
        const uint32_t id = traverse<kStack, EXH, false, ANY, false>(sc, o, d, ANY ? tmax[k] : 0.f, r, s, t,
                                                                   stk + threadIdx.x, kBlock, cnt);
        const bool hit = id != kInvalidId;
        if (ANY) {
            occ[k] = hit ? 1 : 0;
        } else {
            out_id[k] = hit ? id : kInvalidId;
            out_r[k] = hit ? r : FLT_MAX;
            out_s[k] = hit ? s : 0.f;
            out_t[k] = hit ? t : 0.f;
        }
    }
    if (cnt.incomplete) atomicAdd(status, 1u);
}

__global__ void k_tonemap(const float4* __restrict__ acc, float4* __restrict__ out, uint32_t n,
                          float inv_spp, float exposure, float inv_gamma, int gamma_on) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = acc[i];
    c.x *= inv_spp;
    c.y *= inv_spp;
    c.z *= inv_spp;
    c.w *= inv_spp;
    // lib/effects.h:15-49: exposure then gamma on RGB, alpha untouched
    c.x = 1 - expf(-c.x * exposure);
    c.y = 1 - expf(-c.y * exposure);
    c.z = 1 - expf(-c.z * exposure);
    if (gamma_on) {
        c.x = powf(c.x, inv_gamma);
        c.y = powf(c.y, inv_gamma);
        c.z = powf(c.z, inv_gamma);
    }
    out[i] = c;
}

// One thread per triangle reference: copies its record into the two
// reference-ordered half arrays.
__global__ __launch_bounds__(256) void k_gather_leaf_records(const float4* __restrict__ tris,
                                                             const uint32_t* __restrict__ refs,
                                                             size_t nrefs, float4* __restrict__ out) {
    const size_t r = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= nrefs) return;
    const float4* src = tris + 4 * (size_t)refs[r];
    out[2 * r] = src[0];
    out[2 * r + 1] = src[1];
    out[2 * nrefs + 2 * r] = src[2];
    out[2 * nrefs + 2 * r + 1] = src[3];
}

int check_device(int* dev) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return set_error(TN_ERR_NODEVICE, "no HIP device available (%s)",
                         e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    TN_HIP(hipGetDevice(dev));
    return TN_OK;
}

template <class T>
int dev_upload(T** dst, const T* src, size_t count) {
    if (count == 0) count = 1;
    TN_HIP(hipMalloc((void**)dst, count * sizeof(T)));
    if (src) TN_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return TN_OK;
}

}  // namespace
}  // namespace tn

int tn_scene::reserve_stage(size_t bytes) {
    if (d_stage && stage_cap >= bytes) return TN_OK;
    if (d_stage) (void)hipFree(d_stage);
    d_stage = nullptr;
    stage_cap = 0;
    TN_HIP(hipMalloc(&d_stage, bytes));
    stage_cap = bytes;
    return TN_OK;
}

namespace tn {
namespace {

void free_device(tn_scene* sc) {
    (void)hipFree(sc->d_nodes);
    (void)hipFree(sc->d_refs);
    (void)hipFree(sc->d_tris);
    (void)hipFree(sc->d_lrec);
    (void)hipFree(sc->d_fnodes);
    (void)hipFree(sc->d_frefs);
    (void)hipFree(sc->d_flrec);
    (void)hipFree(sc->d_normals);
    (void)hipFree(sc->d_albedo);
    (void)hipFree(sc->d_emission);
    (void)hipFree(sc->d_emit_ids);
    (void)hipFree(sc->d_emit_cdf);
    (void)hipFree(sc->d_reflective);
    (void)hipFree(sc->d_reflectivity);
    (void)hipFree(sc->d_counters);
    (void)hipFree(sc->d_status);
    (void)hipFree(sc->d_work);
    (void)hipFree(sc->d_stage);
    for (hipEvent_t& e : sc->work_ev)
        if (e) (void)hipEventDestroy(e);
}

int finish_scene(tn_scene* sc, const float* verts9, const float* normals9, const float* albedo4,
                 const float* emission4) {
    const uint32_t n = sc->ntris;
    std::string err;
    if (!sc->tree.flat.empty()) {  // else the device layout was built directly
        if (!flat_to_device(sc->tree.flat, n, sc->tree.nodes, sc->tree.refs, err))
            return set_error(TN_ERR_INVALID, "invalid kd-tree: %s", err.c_str());
        sc->tree.height = tree_height(sc->tree.flat);
    }
    int rc;
    if ((rc = dev_upload(&sc->d_nodes, reinterpret_cast<const uint2*>(sc->tree.nodes.data()),
                         sc->tree.nodes.size())))
        return rc;
    if ((rc = dev_upload(&sc->d_refs, sc->tree.refs.data(), sc->tree.refs.size()))) return rc;
    // device record: (v0, uv) (n, vv) (u, uu) (v, denom), one 64-byte line per triangle
    std::vector<float4> rec((size_t)n * 4);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = &sc->pre16[16 * (size_t)i];
        rec[4 * (size_t)i] = make_float4(p[0], p[1], p[2], p[12]);
        rec[4 * (size_t)i + 1] = make_float4(p[3], p[4], p[5], p[13]);
        rec[4 * (size_t)i + 2] = make_float4(p[6], p[7], p[8], p[14]);
        rec[4 * (size_t)i + 3] = make_float4(p[9], p[10], p[11], p[15]);
    }
    if ((rc = dev_upload(&sc->d_tris, rec.data(), rec.size()))) return rc;
    sc->nrefs = (uint32_t)sc->tree.refs.size();
    if ((rc = dev_upload<float4>(&sc->d_lrec, nullptr, 4 * (size_t)sc->nrefs))) return rc;
    if ((rc = gather_leaf_records(sc->d_tris, sc->d_refs, sc->nrefs, sc->d_lrec))) return rc;
    if (sc->refined) {
        refine_reference_tree(verts9, sc->tree.nodes, sc->tree.refs, sc->tree.box, sc->parallel_build,
                              sc->fnodes, sc->frefs, sc->fheight);
        if ((rc = dev_upload(&sc->d_fnodes, reinterpret_cast<const uint2*>(sc->fnodes.data()), sc->fnodes.size())))
            return rc;
        if ((rc = dev_upload(&sc->d_frefs, sc->frefs.data(), sc->frefs.size()))) return rc;
        if ((rc = dev_upload<float4>(&sc->d_flrec, nullptr, 4 * sc->frefs.size()))) return rc;
        if ((rc = gather_leaf_records(sc->d_tris, sc->d_frefs, sc->frefs.size(), sc->d_flrec))) return rc;
    }
    std::vector<float> nrm((size_t)n * 9);
    if (normals9) {
        std::memcpy(nrm.data(), normals9, nrm.size() * sizeof(float));
    } else {
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) nrm[9 * (size_t)i + 3 * k + c] = sc->pre16[16 * (size_t)i + 3 + c];
    }
    if ((rc = dev_upload(&sc->d_normals, nrm.data(), nrm.size()))) return rc;
    std::vector<float4> alb(n), emi(n);
    std::vector<uint32_t> eids;
    std::vector<float> ecdf;
    double acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        alb[i] = albedo4 ? make_float4(albedo4[4 * (size_t)i], albedo4[4 * (size_t)i + 1],
                                       albedo4[4 * (size_t)i + 2], albedo4[4 * (size_t)i + 3])
                         : make_float4(0.8f, 0.8f, 0.8f, 1.f);
        emi[i] = emission4 ? make_float4(emission4[4 * (size_t)i], emission4[4 * (size_t)i + 1],
                                         emission4[4 * (size_t)i + 2], emission4[4 * (size_t)i + 3])
                           : make_float4(0.f, 0.f, 0.f, 0.f);
        if (emi[i].x > 0 || emi[i].y > 0 || emi[i].z > 0) {
            const float* p = &sc->pre16[16 * (size_t)i];
            double ux = p[6], uy = p[7], uz = p[8], vx = p[9], vy = p[10], vz = p[11];
            double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
            acc += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            eids.push_back(i);
            ecdf.push_back((float)acc);
        }
    }
    sc->nemit = (uint32_t)eids.size();
    sc->emit_area = (float)acc;
    if ((rc = dev_upload(&sc->d_albedo, alb.data(), alb.size()))) return rc;
    if ((rc = dev_upload(&sc->d_emission, emi.data(), emi.size()))) return rc;
    if ((rc = dev_upload(&sc->d_emit_ids, eids.data(), eids.size()))) return rc;
    if ((rc = dev_upload(&sc->d_emit_cdf, ecdf.data(), ecdf.size()))) return rc;
    if ((rc = dev_upload<unsigned long long>(&sc->d_counters, nullptr, 8))) return rc;
    if ((rc = dev_upload<uint32_t>(&sc->d_status, nullptr, 4))) return rc;
    TN_HIP(hipMemset(sc->d_counters, 0, 8 * sizeof(unsigned long long)));
    TN_HIP(hipMemset(sc->d_status, 0, 4 * sizeof(uint32_t)));
    const size_t work_words = (size_t)tn_scene::kWorkSlots * tn_scene::kWorkWords;
    if ((rc = dev_upload<uint32_t>(&sc->d_work, nullptr, work_words))) return rc;
    TN_HIP(hipMemset(sc->d_work, 0, work_words * sizeof(uint32_t)));
    return TN_OK;
}

int prepare_triangles(tn_scene* sc, const float* verts9, uint32_t n) {
    sc->ntris = n;
    sc->pre16.resize((size_t)n * 16);
    precompute_triangles(verts9, n, sc->pre16.data());
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = &sc->pre16[16 * (size_t)i];
        for (int k = 0; k < 16; ++k)
            if (!std::isfinite(p[k]))
                return set_error(TN_ERR_INVALID,
                                 "triangle %u is degenerate or not finite (the reference asserts)", i);
    }
    return TN_OK;
}

}  // namespace

int gather_leaf_records(const float4* tris, const uint32_t* refs, size_t nrefs, float4* out) {
    if (nrefs == 0) return TN_OK;
    const size_t blocks = (nrefs + 255) / 256;
    if (blocks > 0x7FFFFFFFu) return set_error(TN_ERR_INVALID, "too many triangle references: %zu", nrefs);
    hipLaunchKernelGGL(k_gather_leaf_records, dim3((uint32_t)blocks), dim3(256), 0, 0, tris, refs, nrefs, out);
    TN_HIP(hipGetLastError());
    TN_HIP(hipStreamSynchronize(0));
    return TN_OK;
}

int launch_rays(tn_scene* sc, const float* o3, const float* d3, const float* tmax, uint32_t n,
                uint32_t* id, float* r, float* s, float* t, uint8_t* occ, uint32_t flags,
                hipStream_t stream, uint32_t* fixed_work) {
    if (n == 0) return TN_OK;
    if (n > (1u << 31)) return set_error(TN_ERR_LIMIT, "%u rays exceed 2^31 per call", n);
    const bool exh = flags & TN_ISECT_EXHAUSTIVE, mt = flags & TN_ISECT_MT,
               cnt = flags & TN_ISECT_COUNT, any = occ != nullptr,
               tick = flags & TN_ISECT_TICK, leaf = flags & TN_ISECT_LEAF, refd = flags & TN_ISECT_REFINED;
    if (refd && !sc->refined) return set_error(TN_ERR_INVALID, "TN_ISECT_REFINED needs a TN_SCENE_REFINED scene");
    if (refd && (mt || tick || leaf || ((flags >> 24) & 15u)))
        return set_error(TN_ERR_INVALID, "TN_ISECT_REFINED runs the cooperative reference leaf test only");
    const SceneView v = refd ? sc->fview() : sc->view();
    // Rays per wave.  The cooperative kernel refills its lanes from one range
    // per wave, so every resident wave (7 per SIMD) gets an equal share of the
    // batch and all finish together (fixed 512-ray chunks left the last of
    // several generations of waves running on part of the chip); at least 64.
    // The per-lane kernels take up to kChunk rays per wave.
    const uint32_t slots = (sc->cus ? sc->cus : 256u) * 4u * kResidentWavesPerSimd;
    uint32_t chunk = (tick || leaf) ? rays_per_wave(n, sc->cus)
                                    : std::max(64u, (uint32_t)(((uint64_t)n + slots - 1) / slots));
    uint32_t waves = (uint32_t)(((uint64_t)n + chunk - 1) / chunk);
    // Large cooperative launches hand out chunks of about 1/32 of a wave's
    // share from a work slot of the scene, so waves that draw cheap
    // rays take more of them and all finish together.  Not under stream
    // capture (the slot's reuse is ordered by an event recorded outside).
    uint32_t* work = nullptr;
    int wslot = -1;
    std::unique_lock<std::mutex> work_lock;
    if (!(tick || leaf) && n >= slots * 256u && fixed_work) {
        // the caller's counters (zeroed, not shared with a concurrent launch)
        work = fixed_work;
        const uint64_t per = ((uint64_t)n + 32ull * slots - 1) / (32ull * slots);
        chunk = (uint32_t)std::min<uint64_t>(512, std::max<uint64_t>(64, (per + 63) & ~63ull));
        waves = slots;
    } else if (!(tick || leaf) && n >= slots * 256u) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream, &cs) != hipSuccess) {
            (void)hipGetLastError();  // e.g. the null stream while another stream captures
            cs = hipStreamCaptureStatusActive;
        }
        if (cs == hipStreamCaptureStatusNone) {
            work_lock = std::unique_lock<std::mutex>(sc->work_mu);
            wslot = (int)(sc->work_next++ % tn_scene::kWorkSlots);
            hipEvent_t& ev = sc->work_ev[wslot];
            if (!ev) TN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            if (sc->work_ev_used[wslot]) TN_HIP(hipStreamWaitEvent(stream, ev, 0));
            work = sc->d_work + (size_t)wslot * tn_scene::kWorkWords;
            const uint64_t per = ((uint64_t)n + 32ull * slots - 1) / (32ull * slots);
            chunk = (uint32_t)std::min<uint64_t>(512, std::max<uint64_t>(64, (per + 63) & ~63ull));
            waves = slots;
        }
    }
    const uint32_t blocks = (waves + kBlock / 64 - 1) / (kBlock / 64);
    const dim3 grid((blocks + kXcds - 1) / kXcds * kXcds), block(kBlock);
    auto go = [&](auto kern) {
        kern<<<grid, block, 0, stream>>>(v, o3, d3, tmax, n, chunk, id, r, s, t, occ, sc->d_counters,
                                         sc->d_status, work);
    };
    // after the launch: the slot is free again once this event has passed
    auto finish = [&]() -> int {
        TN_HIP(hipGetLastError());
        if (refd) {  // marked rays on the reference tree
            const SceneView ref = sc->view();
            const uint32_t fb = std::min<uint32_t>((n + kBlock - 1) / kBlock, (uint32_t)sc->cus * 8u);
            if (exh) {
                if (any) k_fix_rays<true, true><<<fb, kBlock, 0, stream>>>(ref, o3, d3, tmax, n, id, r, s, t, occ, sc->d_status);
                else k_fix_rays<true, false><<<fb, kBlock, 0, stream>>>(ref, o3, d3, tmax, n, id, r, s, t, occ, sc->d_status);
            } else {
                if (any) k_fix_rays<false, true><<<fb, kBlock, 0, stream>>>(ref, o3, d3, tmax, n, id, r, s, t, occ, sc->d_status);
                else k_fix_rays<false, false><<<fb, kBlock, 0, stream>>>(ref, o3, d3, tmax, n, id, r, s, t, occ, sc->d_status);
            }
            TN_HIP(hipGetLastError());
        }
        if (wslot >= 0) {
            TN_HIP(hipEventRecord(sc->work_ev[wslot], stream));
            sc->work_ev_used[wslot] = true;
        }
        return TN_OK;
    };
    const uint32_t wpe = (flags >> 24) & 15u;  // occupancy variant (tuning), 0 = default
    if (!any && !cnt && !tick && !(flags & TN_ISECT_LEAF) && !exh && !mt && wpe) {
        if (wpe == 5) go(k_rays_coop<false, false, false, false, 5>);
        else if (wpe == 7) go(k_rays_coop<false, false, false, false, 7>);
        else if (wpe == 8) go(k_rays_coop<false, false, false, false, 8>);
        else go(k_rays_coop<false, false, false, false, 6>);
        return finish();
    }
#define TN_SEL2(SC, A, C)                                                \
    if (exh) {                                                           \
        if (mt) go(k_rays<SC, true, true, A, C>);                        \
        else go(k_rays<SC, true, false, A, C>);                          \
    } else {                                                             \
        if (mt) go(k_rays<SC, false, true, A, C>);                       \
        else go(k_rays<SC, false, false, A, C>);                         \
    }
#define TN_SELC(A, C)                                                    \
    if (refd) {                                                          \
        if (exh) go(k_rays_coop<true, false, A, C, kCoopWavesPerEU, true>);  \
        else go(k_rays_coop<false, false, A, C, kCoopWavesPerEU, true>);     \
    } else if (exh) {                                                    \
        if (mt) go(k_rays_coop<true, true, A, C>);                       \
        else go(k_rays_coop<true, false, A, C>);                         \
    } else {                                                             \
        if (mt) go(k_rays_coop<false, true, A, C>);                      \
        else go(k_rays_coop<false, false, A, C>);                        \
    }
#define TN_SEL(A, C)                                                     \
    if (tick) {                                                          \
        TN_SEL2(kSchedTick, A, C)                                        \
    } else if (leaf) {                                                   \
        TN_SEL2(kSchedLeaf, A, C)                                        \
    } else {                                                             \
        TN_SELC(A, C)                                                    \
    }
    if (any) {
        TN_SEL(true, false)
    } else if (cnt) {
        TN_SEL(false, true)
    } else {
        TN_SEL(false, false)
    }
#undef TN_SEL2
#undef TN_SELC
#undef TN_SEL
    return finish();
}

int launch_intersect(tn_scene* sc, const float* o3, const float* d3, uint32_t n, uint32_t* id,
                     float* r, float* s, float* t, uint32_t flags, hipStream_t stream) {
    return launch_rays(sc, o3, d3, nullptr, n, id, r, s, t, nullptr, flags, stream);
}

}  // namespace tn

// Work counters of tn_intersect_noalloc: one work slot (8 per-XCD chunk
// counters and a finished-wave counter, 16 words apart) in a device array, so
// that the call allocates nothing.  Every device holds its own copy until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns with that device current (the array's name is a
// host-side stand-in, not a device address).
constexpr uint32_t kScratchWords = 16 * 9;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace tn {
namespace {

constexpr int kScratchDevices = 64;
constexpr uint32_t kNoAllocMaxRays = 1u << 31;
uint32_t* g_scratch[kScratchDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on device dev (current), looked up once per device.
int scratch_address(int dev, uint32_t** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        TN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(TN_ERR_HIP, "no device address for the work counters");
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[dev];
    return TN_OK;
}

__global__ void k_clear_words(uint32_t* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

}  // namespace
}  // namespace tn

using namespace tn;

extern "C" {

const char* tn_last_error(void) { return g_error.c_str(); }

int tn_device_info(char* name, size_t cap, int* cus) {
    int dev;
    int rc = check_device(&dev);
    if (rc) return rc;
    hipDeviceProp_t p;
    TN_HIP(hipGetDeviceProperties(&p, dev));
    if (name && cap) {
        std::snprintf(name, cap, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cus) *cus = p.multiProcessorCount;
    return TN_OK;
}

int tn_scene_create(const float* verts9, const float* normals9, const float* albedo4,
                    const float* emission4, uint32_t n, uint32_t flags, tn_scene** out) {
    if (!out) return set_error(TN_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!verts9 || n == 0) return set_error(TN_ERR_INVALID, "no triangles");
    if (n > TN_MAX_TRIANGLES) return set_error(TN_ERR_LIMIT, "%u triangles exceed 2^30-1", n);
    if ((flags & TN_SCENE_REFINED) && (flags & TN_SCENE_EMPTY_LEAVES))
        return set_error(TN_ERR_INVALID, "TN_SCENE_REFINED refines the reference tree (not with TN_SCENE_EMPTY_LEAVES)");
    std::unique_ptr<tn_scene> sc(new (std::nothrow) tn_scene);
    if (!sc) return set_error(TN_ERR_OOM, "host allocation failed");
    sc->refined = (flags & TN_SCENE_REFINED) != 0;
    sc->parallel_build = !(flags & TN_SCENE_SERIAL_BUILD);
    int rc = check_device(&sc->device);
    if (!rc) (void)hipDeviceGetAttribute(&sc->cus, hipDeviceAttributeMultiprocessorCount, sc->device);
    if (rc) return rc;
    try {
        if ((rc = prepare_triangles(sc.get(), verts9, n))) return rc;
        auto t0 = std::chrono::steady_clock::now();
        if (flags & TN_SCENE_EMPTY_LEAVES)
            build_device_tree(verts9, n, !(flags & TN_SCENE_SERIAL_BUILD), sc->tree.nodes, sc->tree.refs,
                              sc->tree.box, sc->tree.height);
        else
            build_kdtree(verts9, n, !(flags & TN_SCENE_SERIAL_BUILD), sc->tree.flat, sc->tree.box);
        sc->tree.build_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = finish_scene(sc.get(), verts9, normals9, albedo4, emission4))) {
            free_device(sc.get());
            return rc;
        }
    } catch (const std::bad_alloc&) {
        free_device(sc.get());
        return set_error(TN_ERR_OOM, "host allocation failed during build");
    }
    *out = sc.release();
    return TN_OK;
}

int tn_scene_create_with_tree(const float* verts9, const float* normals9, const float* albedo4,
                              const float* emission4, uint32_t n, const uint64_t* flat,
                              size_t num_nodes, const float* box6, tn_scene** out) {
    if (!out) return set_error(TN_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!verts9 || n == 0 || !flat || num_nodes == 0 || !box6)
        return set_error(TN_ERR_INVALID, "missing triangles or tree");
    if (n > TN_MAX_TRIANGLES) return set_error(TN_ERR_LIMIT, "%u triangles exceed 2^30-1", n);
    if (num_nodes >= (1u << 30)) return set_error(TN_ERR_LIMIT, "too many nodes");
    std::unique_ptr<tn_scene> sc(new (std::nothrow) tn_scene);
    if (!sc) return set_error(TN_ERR_OOM, "host allocation failed");
    int rc = check_device(&sc->device);
    if (!rc) (void)hipDeviceGetAttribute(&sc->cus, hipDeviceAttributeMultiprocessorCount, sc->device);
    if (rc) return rc;
    try {
        if ((rc = prepare_triangles(sc.get(), verts9, n))) return rc;
        sc->tree.flat.assign(flat, flat + num_nodes);
        std::memcpy(sc->tree.box, box6, 6 * sizeof(float));
        if ((rc = finish_scene(sc.get(), verts9, normals9, albedo4, emission4))) {
            free_device(sc.get());
            return rc;
        }
    } catch (const std::bad_alloc&) {
        free_device(sc.get());
        return set_error(TN_ERR_OOM, "host allocation failed");
    }
    *out = sc.release();
    return TN_OK;
}

int tn_scene_set_reflection(tn_scene* sc, const float* reflective4, const float* reflectivity) {
    if (!sc || !reflective4 || !reflectivity) return set_error(TN_ERR_INVALID, "NULL argument");
    for (uint32_t i = 0; i < sc->ntris; ++i)
        if (!std::isfinite(reflectivity[i])) return set_error(TN_ERR_INVALID, "reflectivity %u is not finite", i);
    TN_HIP(hipSetDevice(sc->device));
    int rc;
    if (!sc->d_reflective && (rc = dev_upload(&sc->d_reflective, (const float4*)nullptr, sc->ntris))) return rc;
    if (!sc->d_reflectivity && (rc = dev_upload(&sc->d_reflectivity, (const float*)nullptr, sc->ntris))) return rc;
    TN_HIP(hipMemcpy(sc->d_reflective, reflective4, (size_t)sc->ntris * 16, hipMemcpyHostToDevice));
    TN_HIP(hipMemcpy(sc->d_reflectivity, reflectivity, (size_t)sc->ntris * 4, hipMemcpyHostToDevice));
    return TN_OK;
}

int tn_scene_destroy(tn_scene* sc) {
    if (!sc) return set_error(TN_ERR_INVALID, "scene is NULL");
    free_device(sc);
    delete sc;
    return TN_OK;
}

int tn_scene_get_info(const tn_scene* sc, tn_scene_info* info) {
    if (!sc || !info) return set_error(TN_ERR_INVALID, "NULL argument");
    info->num_triangles = sc->ntris;
    info->num_nodes = (uint32_t)(sc->tree.flat.empty() ? sc->tree.nodes.size() : sc->tree.flat.size());
    info->height = sc->tree.height;
    info->device_nodes = (uint32_t)sc->tree.nodes.size();
    info->device_refs = (uint32_t)sc->tree.refs.size();
    info->num_emitters = sc->nemit;
    std::memcpy(info->box, sc->tree.box, sizeof(info->box));
    info->build_ms = sc->tree.build_ms;
    info->refined_nodes = (uint32_t)sc->fnodes.size();
    info->refined_refs = (uint32_t)sc->frefs.size();
    return TN_OK;
}

int tn_scene_get_flat_nodes(const tn_scene* sc, uint64_t* out, size_t cap) {
    if (!sc || !out) return set_error(TN_ERR_INVALID, "NULL argument");
    if (sc->tree.flat.empty())
        return set_error(TN_ERR_INVALID, "scene built with TN_SCENE_EMPTY_LEAVES has no reference FlatNode array");
    if (cap < sc->tree.flat.size()) return set_error(TN_ERR_LIMIT, "capacity too small");
    std::memcpy(out, sc->tree.flat.data(), sc->tree.flat.size() * sizeof(uint64_t));
    return TN_OK;
}

int tn_scene_get_triangle_records(const tn_scene* sc, float* out16, size_t cap) {
    if (!sc || !out16) return set_error(TN_ERR_INVALID, "NULL argument");
    if (cap < sc->ntris) return set_error(TN_ERR_LIMIT, "capacity too small");
    std::memcpy(out16, sc->pre16.data(), sc->pre16.size() * sizeof(float));
    return TN_OK;
}

int tn_intersect(tn_scene* sc, const float* o3, const float* d3, uint32_t n, uint32_t* id,
                 float* r, float* s, float* t, uint32_t flags, void* stream) {
    if (!sc) return set_error(TN_ERR_INVALID, "scene is NULL");
    if (n && (!o3 || !d3 || !id || !r || !s || !t)) return set_error(TN_ERR_INVALID, "NULL array");
    return launch_intersect(sc, o3, d3, n, id, r, s, t, flags, (hipStream_t)stream);
}

int tn_intersect_noalloc(tn_scene* sc, const float* o3, const float* d3, uint32_t n, uint32_t* id,
                         float* r, float* s, float* t, uint32_t flags, void* stream) {
    if (!sc) return set_error(TN_ERR_INVALID, "scene is NULL");
    if (n && (!o3 || !d3 || !id || !r || !s || !t)) return set_error(TN_ERR_INVALID, "NULL array");
    if (n == 0) return TN_OK;
    if (n > kNoAllocMaxRays) return set_error(TN_ERR_LIMIT, "%u rays exceed the limit of 2^31 per call", n);
    int dev = -1;
    TN_HIP(hipGetDevice(&dev));
    if (dev != sc->device)
        return set_error(TN_ERR_INVALID, "device %d is current, the scene lives on device %d", dev, sc->device);
    if (dev < 0 || dev >= kScratchDevices) return set_error(TN_ERR_LIMIT, "device index %d", dev);
    uint32_t* w = nullptr;
    if (int rc = scratch_address(dev, &w)) return rc;
    hipStream_t st = (hipStream_t)stream;
    // the counters start at zero (a kernel, not a memset, so that a captured
    // graph replays it)
    k_clear_words<<<1, 256, 0, st>>>(w, kScratchWords);
    TN_HIP(hipGetLastError());
    return launch_rays(sc, o3, d3, nullptr, n, id, r, s, t, nullptr, flags, st, w);
}

int tn_intersect_host(tn_scene* sc, const float* o3, const float* d3, uint32_t n, uint32_t* id,
                      float* r, float* s, float* t, uint32_t flags) {
    if (!sc) return set_error(TN_ERR_INVALID, "scene is NULL");
    if (n == 0) return TN_OK;
    if (!o3 || !d3 || !id || !r || !s || !t) return set_error(TN_ERR_INVALID, "NULL array");
    TN_HIP(hipSetDevice(sc->device));
    std::lock_guard<std::mutex> lock(sc->stage_mu);
    const size_t ray_bytes = (size_t)n * 12;
    if (int rc = sc->reserve_stage(2 * ray_bytes + (size_t)n * 16)) return rc;
    char* buf = sc->d_stage;
    float* d_o = (float*)buf;
    float* d_d = (float*)(buf + ray_bytes);
    uint32_t* d_id = (uint32_t*)(buf + 2 * ray_bytes);
    float* d_r = (float*)(d_id + n);
    float* d_s = d_r + n;
    float* d_t = d_s + n;
    int rc = TN_OK;
    hipError_t e;
    if ((e = hipMemcpy(d_o, o3, ray_bytes, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d_d, d3, ray_bytes, hipMemcpyHostToDevice)) != hipSuccess) {
        rc = hip_error(e, "hipMemcpy(rays)");
    } else if ((rc = launch_intersect(sc, d_o, d_d, n, d_id, d_r, d_s, d_t, flags, nullptr)) == TN_OK) {
        if ((e = hipMemcpy(id, d_id, (size_t)n * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(r, d_r, (size_t)n * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(s, d_s, (size_t)n * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
            (e = hipMemcpy(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_error(e, "hipMemcpy(hits)");
    }
    if (rc == TN_OK) {
        uint32_t st = 0;
        TN_HIP(hipMemcpy(&st, sc->d_status, 4, hipMemcpyDeviceToHost));
        if (st) {
            TN_HIP(hipMemset(sc->d_status, 0, 4));
            return set_error(TN_ERR_INCOMPLETE, "%u traversals hit the restart cap", st);
        }
    }
    return rc;
}

int tn_occluded(tn_scene* sc, const float* o3, const float* d3, const float* tmax, uint32_t n,
                uint8_t* occ, uint32_t flags, void* stream) {
    if (!sc) return set_error(TN_ERR_INVALID, "scene is NULL");
    if (n == 0) return TN_OK;
    if (!o3 || !d3 || !tmax || !occ) return set_error(TN_ERR_INVALID, "NULL array");
    return launch_rays(sc, o3, d3, tmax, n, nullptr, nullptr, nullptr, nullptr, occ,
                       flags & ~TN_ISECT_COUNT, (hipStream_t)stream);
}

int tn_get_counters(tn_scene* sc, uint64_t* out8, int reset) {
    if (!sc || !out8) return set_error(TN_ERR_INVALID, "NULL argument");
    unsigned long long c[8];
    TN_HIP(hipDeviceSynchronize());
    TN_HIP(hipMemcpy(c, sc->d_counters, sizeof c, hipMemcpyDeviceToHost));
    for (int k = 0; k < 8; ++k) out8[k] = c[k];
    if (reset) TN_HIP(hipMemset(sc->d_counters, 0, sizeof c));
    return TN_OK;
}

int tn_tonemap(const float* accum, float* out, uint32_t np, float inv_spp, float exposure,
               float inverse_gamma, int gamma_enabled, void* stream) {
    if (!accum || !out) return set_error(TN_ERR_INVALID, "NULL array");
    if (np == 0) return TN_OK;
    hipLaunchKernelGGL(k_tonemap, dim3((np + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       (const float4*)accum, (float4*)out, np, inv_spp, exposure, inverse_gamma,
                       gamma_enabled);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

long tn_write_ppm(const float* rgba, uint32_t w, uint32_t h, char* buf, size_t cap) {
    if (!rgba || w == 0) return set_error(TN_ERR_INVALID, "bad image");
    // lib/raster.h:79-100: "P3\n<w> <h>\n255", then per pixel ' ' or '\n'
    // before "%3d %3d %3d" of int(clamp(255 * c * a)).
    auto clamp255 = [](float x) { return x < 0 ? 0.f : (255 < x ? 255.f : x); };
    std::string s;
    s.reserve((size_t)w * h * 12 + 32);
    char tmp[64];
    std::snprintf(tmp, sizeof tmp, "P3\n%u %u\n255", w, h);
    s += tmp;
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i) {
        s += (i % w == 0) ? '\n' : ' ';
        const float* c = rgba + 4 * i;
        std::snprintf(tmp, sizeof tmp, "%3d %3d %3d", (int)clamp255(255 * c[0] * c[3]),
                      (int)clamp255(255 * c[1] * c[3]), (int)clamp255(255 * c[2] * c[3]));
        s += tmp;
    }
    if (buf) {
        if (cap < s.size() + 1) return set_error(TN_ERR_LIMIT, "buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    }
    return (long)s.size();
}

}  // extern "C"
