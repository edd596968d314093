// host_api.cpp -- host-only entry points of libturner_amd: kd-tree build and
// triangle precompute without a device, and the hash-keyed tree cache.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <new>
#include <vector>

#include "tn_host.h"
#include "turner_amd.h"

namespace tn {
int set_error(int code, const char* fmt, ...);
}

namespace {

constexpr char kMagic[4] = {'T', 'N', 'K', 'D'};
constexpr uint32_t kVersion = 1;

// FNV-1a over the triangle coordinates and count.
uint64_t geometry_hash(const float* verts9, uint32_t n) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t len) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < len; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    };
    mix(&n, sizeof n);
    mix(verts9, (size_t)n * 9 * sizeof(float));
    return h;
}

}  // namespace

extern "C" {

int tn_kdtree_build(const float* verts9, uint32_t n, uint32_t flags, uint64_t** nodes,
                    size_t* num_nodes, float* box6, uint32_t* height) {
    if (!verts9 || n == 0 || !nodes || !num_nodes || !box6)
        return tn::set_error(TN_ERR_INVALID, "NULL argument or no triangles");
    if (n > TN_MAX_TRIANGLES) return tn::set_error(TN_ERR_LIMIT, "too many triangles");
    try {
        std::vector<uint64_t> flat;
        tn::build_kdtree(verts9, n, !(flags & TN_SCENE_SERIAL_BUILD), flat, box6);
        uint64_t* out = (uint64_t*)std::malloc(flat.size() * sizeof(uint64_t));
        if (!out) return tn::set_error(TN_ERR_OOM, "host allocation failed");
        std::memcpy(out, flat.data(), flat.size() * sizeof(uint64_t));
        *nodes = out;
        *num_nodes = flat.size();
        if (height) *height = tn::tree_height(flat);
    } catch (const std::bad_alloc&) {
        return tn::set_error(TN_ERR_OOM, "host allocation failed");
    }
    return TN_OK;
}

int tn_kdtree_build_device(const float* verts9, uint32_t n, uint32_t flags, uint32_t** nodes2,
                           size_t* num_nodes, uint32_t** refs, size_t* num_refs, float* box6,
                           uint32_t* height) {
    if (!verts9 || n == 0 || !nodes2 || !num_nodes || !refs || !num_refs || !box6)
        return tn::set_error(TN_ERR_INVALID, "NULL argument or no triangles");
    if (n > TN_MAX_TRIANGLES) return tn::set_error(TN_ERR_LIMIT, "too many triangles");
    try {
        std::vector<tn::DevNode> nodes;
        std::vector<uint32_t> r;
        uint32_t h = 0;
        const bool par = !(flags & TN_SCENE_SERIAL_BUILD);
        if ((flags & TN_SCENE_EMPTY_LEAVES) && (flags & TN_SCENE_REFINED)) {
            return tn::set_error(TN_ERR_INVALID, "TN_SCENE_REFINED refines the reference tree");
        } else if (flags & TN_SCENE_EMPTY_LEAVES) {
            tn::build_device_tree(verts9, n, par, nodes, r, box6, h);
        } else if (flags & TN_SCENE_REFINED) {
            std::vector<uint64_t> flat;
            tn::build_kdtree(verts9, n, par, flat, box6);
            std::string err;
            std::vector<tn::DevNode> rn;
            std::vector<uint32_t> rr;
            if (!tn::flat_to_device(flat, n, rn, rr, err))
                return tn::set_error(TN_ERR_INVALID, "invalid kd-tree: %s", err.c_str());
            tn::refine_reference_tree(verts9, rn, rr, box6, par, nodes, r, h);
        } else {
            std::vector<uint64_t> flat;
            tn::build_kdtree(verts9, n, par, flat, box6);
            std::string err;
            if (!tn::flat_to_device(flat, n, nodes, r, err))
                return tn::set_error(TN_ERR_INVALID, "invalid kd-tree: %s", err.c_str());
            h = tn::tree_height(flat);
        }
        uint32_t* on = (uint32_t*)std::malloc(std::max<size_t>(nodes.size(), 1) * 8);
        uint32_t* orf = (uint32_t*)std::malloc(std::max<size_t>(r.size(), 1) * 4);
        if (!on || !orf) {
            std::free(on);
            std::free(orf);
            return tn::set_error(TN_ERR_OOM, "host allocation failed");
        }
        std::memcpy(on, nodes.data(), nodes.size() * 8);
        std::memcpy(orf, r.data(), r.size() * 4);
        *nodes2 = on;
        *num_nodes = nodes.size();
        *refs = orf;
        *num_refs = r.size();
        if (height) *height = h;
    } catch (const std::bad_alloc&) {
        return tn::set_error(TN_ERR_OOM, "host allocation failed");
    }
    return TN_OK;
}

void tn_buffer_free(void* p) { std::free(p); }

int tn_triangle_records_host(const float* verts9, uint32_t n, float* out16) {
    if (n && (!verts9 || !out16)) return tn::set_error(TN_ERR_INVALID, "NULL argument");
    if (n > TN_MAX_TRIANGLES) return tn::set_error(TN_ERR_LIMIT, "too many triangles");
    tn::precompute_triangles(verts9, n, out16);
    return TN_OK;
}

int tn_tree_cache_save(const char* path, const float* verts9, uint32_t n, const uint64_t* nodes,
                       size_t num_nodes, const float* box6) {
    if (!path || !verts9 || !nodes || !box6 || num_nodes == 0)
        return tn::set_error(TN_ERR_INVALID, "NULL argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return tn::set_error(TN_ERR_INVALID, "cannot open %s for writing", path);
    const uint64_t h = geometry_hash(verts9, n), nn = num_nodes;
    bool ok = std::fwrite(kMagic, 1, 4, f) == 4 && std::fwrite(&kVersion, 4, 1, f) == 1 &&
              std::fwrite(&n, 4, 1, f) == 1 && std::fwrite(&h, 8, 1, f) == 1 &&
              std::fwrite(&nn, 8, 1, f) == 1 && std::fwrite(box6, 4, 6, f) == 6 &&
              std::fwrite(nodes, 8, num_nodes, f) == num_nodes;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? TN_OK : tn::set_error(TN_ERR_INVALID, "short write to %s", path);
}

int tn_tree_cache_load(const char* path, const float* verts9, uint32_t n, uint64_t** nodes,
                       size_t* num_nodes, float* box6) {
    if (!path || !verts9 || !nodes || !num_nodes || !box6)
        return tn::set_error(TN_ERR_INVALID, "NULL argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return tn::set_error(TN_ERR_INVALID, "no cache at %s", path);
    char magic[4];
    uint32_t ver = 0, cn = 0;
    uint64_t h = 0, nn = 0;
    float box[6];
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, kMagic, 4) == 0 &&
              std::fread(&ver, 4, 1, f) == 1 && ver == kVersion && std::fread(&cn, 4, 1, f) == 1 &&
              std::fread(&h, 8, 1, f) == 1 && std::fread(&nn, 8, 1, f) == 1 &&
              std::fread(box, 4, 6, f) == 6;
    if (!ok) {
        std::fclose(f);
        return tn::set_error(TN_ERR_INVALID, "%s is not a tree cache of this version", path);
    }
    if (cn != n || h != geometry_hash(verts9, n)) {
        std::fclose(f);
        return tn::set_error(TN_ERR_INVALID, "%s belongs to other geometry (stale cache)", path);
    }
    if (nn == 0 || nn >= (1ull << 30)) {
        std::fclose(f);
        return tn::set_error(TN_ERR_INVALID, "%s: bad node count", path);
    }
    uint64_t* out = (uint64_t*)std::malloc(nn * 8);
    if (!out) {
        std::fclose(f);
        return tn::set_error(TN_ERR_OOM, "host allocation failed");
    }
    if (std::fread(out, 8, nn, f) != nn) {
        std::free(out);
        std::fclose(f);
        return tn::set_error(TN_ERR_INVALID, "%s is truncated", path);
    }
    std::fclose(f);
    std::string err;
    std::vector<uint64_t> flat(out, out + nn);
    std::vector<tn::DevNode> dn;
    std::vector<uint32_t> refs;
    if (!tn::flat_to_device(flat, n, dn, refs, err)) {
        std::free(out);
        return tn::set_error(TN_ERR_INVALID, "%s: corrupt tree (%s)", path, err.c_str());
    }
    *nodes = out;
    *num_nodes = nn;
    std::memcpy(box6, box, sizeof box);
    return TN_OK;
}

}  // extern "C"
