// tn_scene.h -- the scene handle behind tn_scene* and helpers shared by the
// HIP translation units of libturner_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "tn_host.h"
#include "tn_traverse.h"
#include "turner_amd.h"

struct tn_scene {
    uint32_t ntris = 0;
    int device = 0;
    int cus = 256;  // compute units of the device
    tn::HostTree tree;
    std::vector<float> pre16;
    uint32_t nemit = 0;
    float emit_area = 0.f;
    // device buffers (owned)
    uint2* d_nodes = nullptr;
    uint32_t* d_refs = nullptr;
    float4* d_tris = nullptr;
    float4* d_lrec = nullptr;    // records per reference: 2 x nrefs first halves, then second halves
    uint32_t nrefs = 0;
    // TN_SCENE_REFINED: the refined reference tree (same layout, own records per reference)
    bool refined = false;
    bool parallel_build = true;
    std::vector<tn::DevNode> fnodes;
    std::vector<uint32_t> frefs;
    uint32_t fheight = 0;
    uint2* d_fnodes = nullptr;
    uint32_t* d_frefs = nullptr;
    float4* d_flrec = nullptr;
    float* d_normals = nullptr;  // 9 per triangle
    float4* d_albedo = nullptr;
    float4* d_emission = nullptr;
    uint32_t* d_emit_ids = nullptr;
    float* d_emit_cdf = nullptr;  // inclusive prefix of emitter areas
    float4* d_reflective = nullptr;    // raytracer: reflective colour (null: no reflection)
    float* d_reflectivity = nullptr;   // raytracer: reflectivity factor
    unsigned long long* d_counters = nullptr;  // 8 counters
    uint32_t* d_status = nullptr;              // [0] incomplete traversals, [1] queue overflow
    // Work counters of large tn_intersect / tn_occluded launches: a ring of
    // slots, each 8 per-XCD chunk counters and a finished-wave counter (16
    // words apart), zero between launches (the last wave of a launch clears
    // its slot).  A slot is reused only after the event recorded behind its
    // previous launch, so launches on different streams never share one.
    static constexpr int kWorkSlots = 16;
    static constexpr int kWorkWords = 16 * 9;
    uint32_t* d_work = nullptr;
    hipEvent_t work_ev[kWorkSlots] = {};
    bool work_ev_used[kWorkSlots] = {};
    uint32_t work_next = 0;
    std::mutex work_mu;
    // Staging memory of the host-pointer calls (tn_intersect_host, tn_render):
    // grows to the largest call and is reused, held under stage_mu for the
    // whole (synchronous) call.
    char* d_stage = nullptr;
    size_t stage_cap = 0;
    std::mutex stage_mu;

    // Device memory of at least `bytes` in d_stage (caller holds stage_mu).
    int reserve_stage(size_t bytes);

    // the refined tree (TN_SCENE_REFINED scenes only)
    tn::SceneView fview() const {
        tn::SceneView v = view();
        v.nodes = d_fnodes;
        v.refs = d_frefs;
        v.lra = d_flrec;
        v.lrb = d_flrec ? d_flrec + 2 * frefs.size() : nullptr;
        return v;
    }

    tn::SceneView view() const {
        tn::SceneView v;
        v.nodes = d_nodes;
        v.refs = d_refs;
        v.tris = d_tris;
        v.lra = d_lrec;
        v.lrb = d_lrec ? d_lrec + 2 * (size_t)nrefs : nullptr;
        for (int k = 0; k < 3; ++k) {
            v.bmin[k] = tree.box[k];
            v.bmax[k] = tree.box[3 + k];
        }
        return v;
    }
};

namespace tn {

int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what);

#define TN_HIP(call)                                              \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return ::tn::hip_error(e_, #call);  \
    } while (0)

// Writes out[2r], out[2r+1] = tris[4 refs[r]], tris[4 refs[r] + 1] and
// out[2 nrefs + 2r], out[2 nrefs + 2r + 1] = tris[4 refs[r] + 2], tris[4 refs[r] + 3]
// for r < nrefs (every refs[r] indexes tris), on the null stream, and waits.
int gather_leaf_records(const float4* tris, const uint32_t* refs, size_t nrefs, float4* out);

// Fixed launch shape of the traversal kernels: 256 lanes = 4 wave64.
constexpr int kBlock = 256;
constexpr int kStack = 8;
constexpr int kChunk = 512;  // rays per wave in the refill kernels
constexpr uint32_t kXcds = 8;  // MI355X: 8 XCDs of 32 CUs, one L2 each
constexpr uint32_t kResidentWavesPerSimd = 7;

// Rays per wave for a batch of n: kChunk when the batch fills every resident
// wave slot (cus x 4 SIMDs x 7), otherwise fewer (a multiple of 64, at least
// 64) so that the batch still spreads over the whole chip -- a traversal is
// latency-bound, and a wave alone on its SIMD cannot hide it.
__host__ __device__ inline uint32_t rays_per_wave(uint32_t n, uint32_t cus) {
    const uint32_t slots = (cus ? cus : 256u) * 4u * kResidentWavesPerSimd;
    uint32_t c = (uint32_t)(((uint64_t)n + slots - 1) / slots);
    c = (c + 63u) & ~63u;
    return c < 64u ? 64u : (c > (uint32_t)kChunk ? (uint32_t)kChunk : c);
}

// Traversal over n rays (device arrays): closest hit into id/r/s/t, or with
// occ != NULL visibility against tmax into occ.
// fixed_work: zeroed work counters (16 x 9 words) the launch may use instead
// of a work slot of the scene, also under stream capture (null: a slot).
int launch_rays(tn_scene* sc, const float* o3, const float* d3, const float* tmax, uint32_t n,
                uint32_t* id, float* r, float* s, float* t, uint8_t* occ, uint32_t flags,
                hipStream_t stream, uint32_t* fixed_work = nullptr);
// Launches closest-hit traversal over n rays (device arrays) into id/r/s/t.
int launch_intersect(tn_scene* sc, const float* o3, const float* d3, uint32_t n, uint32_t* id,
                     float* r, float* s, float* t, uint32_t flags, hipStream_t stream);

}  // namespace tn
