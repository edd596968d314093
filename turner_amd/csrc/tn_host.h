// tn_host.h -- host-side data structures shared by the C ABI, the SAH
// builder and the device code of libturner_amd.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace tn {

// Device kd-tree layout (8 bytes per node, depth-first, left child = i + 1):
//   inner: x = bits of split position, y = right_index << 2 | axis (0..2)
//   leaf : x = offset into the triangle-reference array, y = count << 2 | 3
struct DevNode {
    uint32_t x, y;
};

struct HostTree {
    std::vector<uint64_t> flat;     // reference FlatNode array (lib/kdtree.h:62-154)
    std::vector<DevNode> nodes;     // device layout
    std::vector<uint32_t> refs;     // leaf triangle references (device layout)
    float box[6];
    uint32_t height = 0;
    double build_ms = 0;
};

// 16 floats per triangle: v0 n u v uv vv uu denom (lib/triangle.h:22-39).
void precompute_triangles(const float* verts9, uint32_t n, float* out16);

// SAH kd-tree (lib/kdtree.cpp:124-490) -> reference FlatNode array + root box.
void build_kdtree(const float* verts9, uint32_t n, bool parallel, std::vector<uint64_t>& flat,
                  float box6[6]);

// Reference FlatNode array -> device layout; validates the array (indices in
// range, leaf runs terminated).  Returns false with a message on a bad tree.
bool flat_to_device(const std::vector<uint64_t>& flat, uint32_t num_triangles,
                    std::vector<DevNode>& nodes, std::vector<uint32_t>& refs, std::string& err);

uint32_t tree_height(const std::vector<uint64_t>& flat);

// The reference device tree (rnodes / rrefs over box6) with every leaf's
// references filtered to triangles near the leaf and larger leaves replaced
// by rendering-variant sub-trees; inner reference nodes and their order stay.
void refine_reference_tree(const float* verts9, const std::vector<DevNode>& rnodes,
                           const std::vector<uint32_t>& rrefs, const float box6[6], bool parallel,
                           std::vector<DevNode>& nodes, std::vector<uint32_t>& refs, uint32_t& height);

// SAH build that keeps empty children as empty leaves, straight into the
// device layout (no reference FlatNode array).
void build_device_tree(const float* verts9, uint32_t n, bool parallel, std::vector<DevNode>& nodes,
                       std::vector<uint32_t>& refs, float box6[6], uint32_t& height);

}  // namespace tn
