// turner.hpp -- host C++ interface of turner's pathtracer hot path, kept with
// the reference's names and argument meaning, running on the MI355X through
// the C ABI of libturner_amd (include/turner_amd.h).
//
//   KDTree(Triangles)                      lib/kdtree.h:187-264
//   KDTreeIntersection::intersect(...)     lib/kdtree.h:270-309
//   trace(ray, intersection, lights, depth, conf)   trace.h:23-25
//   TracerConfig                           config.h:103-153
//
// Errors are reported as turner_amd::Error exceptions carrying the C ABI
// status (the reference asserts); no call falls back to the CPU.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "turner_amd.h"

namespace turner_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void check(int rc);

struct Vec3 {
    float x = 0, y = 0, z = 0;
};
struct Color {  // aiColor4D: RGBA, alpha transported like a channel (types.h:60)
    float r = 0, g = 0, b = 0, a = 0;
};
struct Ray {  // src/geometry.h:809-822
    Vec3 o, d;
};
struct Light {  // lib/types.h:81-84
    Vec3 position;
    Color color;
};
struct Triangle {  // the parts of lib/triangle.h:95-114 the path reads
    std::array<Vec3, 3> vertices;
    std::array<Vec3, 3> normals;
    Color diffuse{0.8f, 0.8f, 0.8f, 1.f};
    Color emissive;
    Color reflective;        // raytracer (lib/triangle.h:100-101)
    float reflectivity = 0;  // raytracer: 0 = no mirror ray
};
using Triangles = std::vector<Triangle>;

class OptionalId {  // lib/kdtree.h:156-181
public:
    OptionalId() = default;
    explicit OptionalId(uint32_t id) : id_(id) {}
    explicit operator bool() const { return id_ < TN_INVALID_ID; }
    uint32_t value() const { return id_; }

private:
    uint32_t id_ = TN_INVALID_ID;
};

struct Camera {  // lib/types.h:89-123
    Vec3 position;
    float trafo[9] = {-1, 0, 0, 0, 1, 0, 0, 0, -1};
    float delta_x = 0.35708f, delta_y = 0.35708f, aspect = 1.f;
};

class KDTree {
public:
    KDTree() = default;
    explicit KDTree(Triangles tris);
    // Uses a tree from the cache file when it belongs to these triangles,
    // builds (and caches) it otherwise.
    KDTree(Triangles tris, const std::string& cache_path);
    // Scene flags TN_SCENE_* (e.g. TN_SCENE_EMPTY_LEAVES: the rendering tree).
    KDTree(Triangles tris, uint32_t scene_flags);
    KDTree(const KDTree&) = delete;
    KDTree& operator=(const KDTree&) = delete;
    KDTree(KDTree&& o) noexcept;
    KDTree& operator=(KDTree&& o) noexcept;
    ~KDTree();

    size_t height() const { return info_.height; }
    size_t num_nodes() const { return info_.num_nodes; }
    size_t num_triangles() const { return tris_.size(); }
    const Triangles& triangles() const { return tris_; }
    const float* box() const { return info_.box; }
    const Triangle& operator[](uint32_t id) const { return tris_[id]; }
    const Triangle& at(uint32_t id) const { return tris_.at(id); }
    tn_scene* handle() const { return scene_; }
    double build_ms() const { return info_.build_ms; }
    // built with TN_SCENE_REFINED: intersect() walks the refined tree
    bool refined() const { return info_.refined_nodes != 0; }

private:
    void create(const uint64_t* nodes, size_t n, const float* box, uint32_t flags = 0);
    Triangles tris_;
    tn_scene* scene_ = nullptr;
    tn_scene_info info_{};
};

class KDTreeIntersection {
public:
    explicit KDTreeIntersection(const KDTree& tree) : tree_(&tree) {}
    const Triangle& operator[](uint32_t id) const { return (*tree_)[id]; }
    const Triangle& at(uint32_t id) const { return tree_->at(id); }
    // One ray (lib/kdtree.cpp:515-578): a and b are written on a hit only; r
    // on a hit, and FLT_MAX on a miss of a ray that enters the root box.
    OptionalId intersect(const Ray& ray, float& r, float& a, float& b);
    OptionalId intersect(const Ray& ray) {
        float u;
        return intersect(ray, u, u, u);
    }
    // Batched: host arrays of n rays; results per ray (TN_INVALID_ID on miss).
    void intersect(const std::vector<Ray>& rays, std::vector<uint32_t>& ids, std::vector<float>& r,
                   std::vector<float>& s, std::vector<float>& t);
    const KDTree& tree() const { return *tree_; }

private:
    const KDTree* tree_;
};

struct TracerConfig {  // config.h:45-60 + 103-153, pathtracer USAGE defaults
    float aspect = 1;
    size_t width = 640;
    bool verbose = false;
    size_t num_threads = 1;
    float inverse_gamma = 0.454545f;
    float exposure = 1;
    Color bg_color{0, 0, 0, 1};
    bool gamma_correction_enabled = true;
    std::string filename;
    int max_recursion_depth = 3;
    int num_pixel_samples = 1;
    int num_monte_carlo_samples = 8;
    // extensions of this implementation
    uint32_t flags = 0;  // TN_RENDER_*
    uint64_t seed = 1;
    std::string tree_cache;
    // which trace() plugin (TN_RENDERER_*) and its options
    uint32_t renderer = TN_RENDERER_PATHTRACER;
    float max_visibility = 2;    // raycaster --max-visibility (config.h:110)
    float shadow_intensity = 0.5f;  // raytracer --shadow (config.h:113)
};

// trace.h:23-25: radiance of one ray from recursion level `depth`.
Color trace(const Ray& ray, KDTreeIntersection& ti, const std::vector<Light>& lights, int depth,
            const TracerConfig& conf);
// Batched twin: radiance of many rays.
std::vector<Color> trace(const std::vector<Ray>& rays, KDTreeIntersection& ti,
                         const std::vector<Light>& lights, int depth, const TracerConfig& conf);

struct RenderResult {
    std::vector<Color> linear;  // mean over pixel samples (main.cpp:216)
    std::vector<Color> image;   // after exposure and gamma (main.cpp:218-223)
    tn_render_stats stats{};
};
// The pixel loop of main.cpp:187-226.
RenderResult render(const KDTree& tree, const Camera& cam, const std::vector<Light>& lights,
                    const TracerConfig& conf, size_t height);

// Learned guiding proposal (include/turner_amd.h, tn_guide_*): no reference
// counterpart.  Owns a tn_guide whose position encoding spans the tree's box.
class Guide {
public:
    explicit Guide(const KDTree& tree, uint64_t seed = 1, float lr = 1e-3f, uint32_t batch = 16384,
                   uint32_t epochs = 5, float mis_guided = 0.7f);
    Guide(const Guide&) = delete;
    Guide& operator=(const Guide&) = delete;
    ~Guide();
    tn_guide* handle() const { return g_; }

private:
    tn_guide* g_ = nullptr;
};

// Progressive guided rendering: train_passes passes of train_spp samples
// per pixel train the guide online (each on a uniform subset of
// train_samples radiance records) and stay out of the image; then passes of
// pass_spp samples render conf.num_pixel_samples samples per pixel with the
// guide (one-sample MIS with the cosine lobe, unbiased).  defer_training:
// each training pass's records train beside the next training pass's render
// (TN_RENDER_DEFER_TRAINING), the last set before the image passes.  The
// defaults (with Guide's epochs 5 and guide share 0.7) are the fastest
// schedule to a relMSE target measured on the occluded-light scene.  Needs
// num_monte_carlo_samples == 1.
struct GuidedSchedule {
    int train_passes = 6;
    int train_spp = 1;
    uint32_t train_samples = 1u << 18;
    int pass_spp = 64;
    bool defer_training = true;
};
RenderResult render_guided(const KDTree& tree, const Camera& cam, const std::vector<Light>& lights,
                           const TracerConfig& conf, size_t height, Guide& guide,
                           const GuidedSchedule& schedule = GuidedSchedule());

// PPM P3 as lib/raster.h:79-100 writes it.
std::string to_ppm(const std::vector<Color>& image, size_t width, size_t height);
// Portable float map (PF, little-endian FP32 RGB, bottom row first) of an
// image, e.g. the linear mean before exposure, for error measurements.
void write_pfm(const std::string& path, const std::vector<Color>& image, size_t width, size_t height);

// Scene file written by turner_amd.scenes.save_tns (Python): triangles,
// vertex normals, RGBA albedo and emission (version 2 adds the reflective
// colour and reflectivity), camera and point lights.
struct Scene {
    Triangles triangles;
    Camera camera;
    std::vector<Light> lights;
};
Scene load_tns(const std::string& path);

}  // namespace turner_amd
