// turner.cpp -- the host C++ interface of turner.hpp over the C ABI.
#include "turner.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace turner_amd {

void check(int rc) {
    if (rc < 0) throw Error(rc, std::string("turner_amd: ") + tn_last_error());
}

namespace {

void flatten(const Triangles& tris, std::vector<float>& v, std::vector<float>& n,
             std::vector<float>& a, std::vector<float>& e) {
    v.resize(tris.size() * 9);
    n.resize(tris.size() * 9);
    a.resize(tris.size() * 4);
    e.resize(tris.size() * 4);
    for (size_t i = 0; i < tris.size(); ++i) {
        const Triangle& t = tris[i];
        for (int k = 0; k < 3; ++k) {
            v[9 * i + 3 * k] = t.vertices[k].x;
            v[9 * i + 3 * k + 1] = t.vertices[k].y;
            v[9 * i + 3 * k + 2] = t.vertices[k].z;
            n[9 * i + 3 * k] = t.normals[k].x;
            n[9 * i + 3 * k + 1] = t.normals[k].y;
            n[9 * i + 3 * k + 2] = t.normals[k].z;
        }
        std::memcpy(&a[4 * i], &t.diffuse, 16);
        std::memcpy(&e[4 * i], &t.emissive, 16);
    }
}

struct DeviceBuffer {  // owns device memory for the batched helpers
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) throw Error(TN_ERR_OOM, "hipMalloc failed");
    }
    ~DeviceBuffer() { (void)hipFree(p); }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

void to_device(void* dst, const void* src, size_t bytes) {
    if (hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
        throw Error(TN_ERR_HIP, "hipMemcpy to device failed");
}
void to_host(void* dst, const void* src, size_t bytes) {
    if (hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        throw Error(TN_ERR_HIP, "hipMemcpy to host failed");
}

tn_render_params params_of(const TracerConfig& conf, size_t width, size_t height) {
    tn_render_params p{};
    p.width = (uint32_t)width;
    p.height = (uint32_t)height;
    p.max_depth = conf.max_recursion_depth;
    p.mc_samples = conf.num_monte_carlo_samples;
    p.pixel_samples = conf.num_pixel_samples;
    p.flags = conf.flags;
    p.seed = conf.seed;
    std::memcpy(p.bg, &conf.bg_color, 16);
    p.exposure = conf.exposure;
    p.inverse_gamma = conf.inverse_gamma;
    p.gamma_enabled = conf.gamma_correction_enabled ? 1 : 0;
    p.renderer = conf.renderer;
    p.max_visibility = conf.max_visibility;
    p.shadow_intensity = conf.shadow_intensity;
    return p;
}

std::vector<tn_light> lights_of(const std::vector<Light>& lights) {
    std::vector<tn_light> out(lights.size());
    for (size_t i = 0; i < lights.size(); ++i) {
        std::memcpy(out[i].pos, &lights[i].position, 12);
        std::memcpy(out[i].color, &lights[i].color, 16);
    }
    return out;
}

}  // namespace

KDTree::KDTree(Triangles tris) : tris_(std::move(tris)) { create(nullptr, 0, nullptr); }

KDTree::KDTree(Triangles tris, const std::string& cache) : tris_(std::move(tris)) {
    std::vector<float> v, n, a, e;
    flatten(tris_, v, n, a, e);
    uint64_t* nodes = nullptr;
    size_t nn = 0;
    float box[6];
    if (!cache.empty() &&
        tn_tree_cache_load(cache.c_str(), v.data(), (uint32_t)tris_.size(), &nodes, &nn, box) == TN_OK) {
        create(nodes, nn, box);
        tn_buffer_free(nodes);
        return;
    }
    uint32_t h = 0;
    check(tn_kdtree_build(v.data(), (uint32_t)tris_.size(), 0, &nodes, &nn, box, &h));
    if (!cache.empty()) tn_tree_cache_save(cache.c_str(), v.data(), (uint32_t)tris_.size(), nodes, nn, box);
    create(nodes, nn, box);
    tn_buffer_free(nodes);
}

KDTree::KDTree(Triangles tris, uint32_t scene_flags) : tris_(std::move(tris)) {
    create(nullptr, 0, nullptr, scene_flags);
}

void KDTree::create(const uint64_t* nodes, size_t nn, const float* box, uint32_t flags) {
    std::vector<float> v, n, a, e;
    flatten(tris_, v, n, a, e);
    if (nodes)
        check(tn_scene_create_with_tree(v.data(), n.data(), a.data(), e.data(), (uint32_t)tris_.size(),
                                        nodes, nn, box, &scene_));
    else
        check(tn_scene_create(v.data(), n.data(), a.data(), e.data(), (uint32_t)tris_.size(), flags, &scene_));
    check(tn_scene_get_info(scene_, &info_));
    bool reflects = false;
    for (const Triangle& t : tris_) reflects = reflects || t.reflectivity != 0;
    if (reflects) {
        std::vector<float> rc(4 * tris_.size()), rf(tris_.size());
        for (size_t i = 0; i < tris_.size(); ++i) {
            std::memcpy(&rc[4 * i], &tris_[i].reflective, 16);
            rf[i] = tris_[i].reflectivity;
        }
        check(tn_scene_set_reflection(scene_, rc.data(), rf.data()));
    }
}

KDTree::KDTree(KDTree&& o) noexcept : tris_(std::move(o.tris_)), scene_(o.scene_), info_(o.info_) {
    o.scene_ = nullptr;
}
KDTree& KDTree::operator=(KDTree&& o) noexcept {
    if (this != &o) {
        if (scene_) tn_scene_destroy(scene_);
        tris_ = std::move(o.tris_);
        scene_ = o.scene_;
        info_ = o.info_;
        o.scene_ = nullptr;
    }
    return *this;
}
KDTree::~KDTree() {
    if (scene_) tn_scene_destroy(scene_);
}

namespace {

// Root slab test of lib/intersection.h:105-128 on the direction with zero
// components replaced by 1e-5 (fix_direction, lib/kdtree.cpp:503-511).
bool enters_root_box(const Ray& ray, const float* box) {
    const float o[3] = {ray.o.x, ray.o.y, ray.o.z};
    float d[3] = {ray.d.x, ray.d.y, ray.d.z};
    float tmin = 0, tmax = 0;
    for (int ax = 0; ax < 3; ++ax) {
        if (d[ax] == 0) d[ax] = 0.00001f;
        const float inv = 1 / d[ax];
        const float t1 = (box[ax] - o[ax]) * inv, t2 = (box[3 + ax] - o[ax]) * inv;
        if (ax == 0) {
            tmin = std::fmin(t1, t2);
            tmax = std::fmax(t1, t2);
        } else {
            tmin = std::fmax(tmin, std::fmin(t1, t2));
            tmax = std::fmin(tmax, std::fmax(t1, t2));
        }
    }
    return !(tmax < tmin);
}

}  // namespace

OptionalId KDTreeIntersection::intersect(const Ray& ray, float& r, float& a, float& b) {
    uint32_t id;
    float rr, ss, tt;
    check(tn_intersect_host(tree_->handle(), &ray.o.x, &ray.d.x, 1, &id, &rr, &ss, &tt,
                            tree_->refined() ? TN_ISECT_REFINED : 0u));
    if (id == TN_INVALID_ID) {
        // the reference sets r to FLT_MAX once the ray enters the root box
        // (lib/kdtree.cpp:535) and leaves it there on a miss
        if (enters_root_box(ray, tree_->box())) r = rr;
        return OptionalId{};
    }
    r = rr;
    a = ss;
    b = tt;
    return OptionalId{id};
}

void KDTreeIntersection::intersect(const std::vector<Ray>& rays, std::vector<uint32_t>& ids,
                                   std::vector<float>& r, std::vector<float>& s,
                                   std::vector<float>& t) {
    const size_t n = rays.size();
    std::vector<float> o(3 * n), d(3 * n);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(&o[3 * i], &rays[i].o, 12);
        std::memcpy(&d[3 * i], &rays[i].d, 12);
    }
    ids.resize(n);
    r.resize(n);
    s.resize(n);
    t.resize(n);
    check(tn_intersect_host(tree_->handle(), o.data(), d.data(), (uint32_t)n, ids.data(), r.data(),
                            s.data(), t.data(), tree_->refined() ? TN_ISECT_REFINED : 0u));
}

std::vector<Color> trace(const std::vector<Ray>& rays, KDTreeIntersection& ti,
                         const std::vector<Light>& lights, int depth, const TracerConfig& conf) {
    const size_t n = rays.size();
    std::vector<Color> out(n);
    if (n == 0) return out;
    DeviceBuffer buf(n * (24 + 16));
    std::vector<float> od(6 * n);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(&od[3 * i], &rays[i].o, 12);
        std::memcpy(&od[3 * n + 3 * i], &rays[i].d, 12);
    }
    to_device(buf.p, od.data(), 24 * n);
    float* rgba = buf.as<float>() + 6 * n;
    if (hipMemset(rgba, 0, 16 * n) != hipSuccess) throw Error(TN_ERR_HIP, "hipMemset failed");
    tn_render_params p = params_of(conf, 1, 1);
    // The reference draws hemisphere samples from one stream per worker
    // thread that advances with every trace() call (lib/sampling.h:12); here
    // each call of a thread takes the next range of tn_trace's streams, so
    // repeated calls for the same ray give fresh samples.
    static thread_local uint32_t calls = 0;
    p.sample_offset = calls++;
    const auto ls = lights_of(lights);
    check(tn_trace(ti.tree().handle(), buf.as<float>(), buf.as<float>() + 3 * n, (uint32_t)n, depth,
                   ls.data(), (uint32_t)ls.size(), &p, rgba, nullptr, nullptr));
    if (hipDeviceSynchronize() != hipSuccess) throw Error(TN_ERR_HIP, "device synchronisation failed");
    to_host(out.data(), rgba, 16 * n);
    return out;
}

Color trace(const Ray& ray, KDTreeIntersection& ti, const std::vector<Light>& lights, int depth,
            const TracerConfig& conf) {
    return trace(std::vector<Ray>{ray}, ti, lights, depth, conf)[0];
}

RenderResult render(const KDTree& tree, const Camera& cam, const std::vector<Light>& lights,
                    const TracerConfig& conf, size_t height) {
    RenderResult res;
    const size_t np = conf.width * height;
    res.linear.resize(np);
    res.image.resize(np);
    tn_camera c{};
    std::memcpy(c.pos, &cam.position, 12);
    std::memcpy(c.trafo, cam.trafo, sizeof c.trafo);
    c.delta_x = cam.delta_x;
    c.delta_y = cam.delta_y;
    const tn_render_params p = params_of(conf, conf.width, height);
    const auto ls = lights_of(lights);
    check(tn_render(tree.handle(), &c, ls.data(), (uint32_t)ls.size(), &p,
                    reinterpret_cast<float*>(res.linear.data()),
                    reinterpret_cast<float*>(res.image.data()), &res.stats));
    return res;
}

std::string to_ppm(const std::vector<Color>& image, size_t width, size_t height) {
    long n = tn_write_ppm(reinterpret_cast<const float*>(image.data()), (uint32_t)width,
                          (uint32_t)height, nullptr, 0);
    check((int)(n < 0 ? n : 0));
    std::string s((size_t)n + 1, '\0');
    check((int)tn_write_ppm(reinterpret_cast<const float*>(image.data()), (uint32_t)width,
                            (uint32_t)height, &s[0], s.size()) < 0 ? -1 : 0);
    s.resize((size_t)n);
    return s;
}

void write_pfm(const std::string& path, const std::vector<Color>& image, size_t width, size_t height) {
    if (image.size() != width * height) throw Error(TN_ERR_INVALID, "image size does not match width x height");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error(TN_ERR_INVALID, "cannot write " + path);
    f << "PF\n" << width << " " << height << "\n-1.0\n";  // negative scale: little-endian
    std::vector<float> row(3 * width);
    for (size_t y = height; y-- > 0;) {
        for (size_t x = 0; x < width; ++x) {
            const Color& c = image[y * width + x];
            row[3 * x] = c.r;
            row[3 * x + 1] = c.g;
            row[3 * x + 2] = c.b;
        }
        f.write(reinterpret_cast<const char*>(row.data()), (std::streamsize)(row.size() * 4));
    }
    if (!f) throw Error(TN_ERR_INVALID, "cannot write " + path);
}

Scene load_tns(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(TN_ERR_INVALID, "cannot open scene " + path);
    char magic[4];
    uint32_t version = 0, ntris = 0, nlights = 0;
    f.read(magic, 4);
    f.read(reinterpret_cast<char*>(&version), 4);
    f.read(reinterpret_cast<char*>(&ntris), 4);
    f.read(reinterpret_cast<char*>(&nlights), 4);
    if (!f || std::memcmp(magic, "TNSC", 4) != 0 || (version != 1 && version != 2))
        throw Error(TN_ERR_INVALID, path + " is not a turner_amd scene file");
    std::vector<float> v(9 * (size_t)ntris), n(9 * (size_t)ntris), a(4 * (size_t)ntris),
        e(4 * (size_t)ntris), rc, rf;
    f.read(reinterpret_cast<char*>(v.data()), v.size() * 4);
    f.read(reinterpret_cast<char*>(n.data()), n.size() * 4);
    f.read(reinterpret_cast<char*>(a.data()), a.size() * 4);
    f.read(reinterpret_cast<char*>(e.data()), e.size() * 4);
    if (version == 2) {  // reflective colour (4N) and reflectivity (N)
        rc.resize(4 * (size_t)ntris);
        rf.resize(ntris);
        f.read(reinterpret_cast<char*>(rc.data()), rc.size() * 4);
        f.read(reinterpret_cast<char*>(rf.data()), rf.size() * 4);
    }
    Scene sc;
    float cam[15];
    f.read(reinterpret_cast<char*>(cam), sizeof cam);
    sc.camera.position = Vec3{cam[0], cam[1], cam[2]};
    std::memcpy(sc.camera.trafo, cam + 3, 36);
    sc.camera.delta_x = cam[12];
    sc.camera.delta_y = cam[13];
    sc.camera.aspect = cam[14];
    for (uint32_t i = 0; i < nlights; ++i) {
        float l[7];
        f.read(reinterpret_cast<char*>(l), sizeof l);
        sc.lights.push_back(Light{Vec3{l[0], l[1], l[2]}, Color{l[3], l[4], l[5], l[6]}});
    }
    if (!f) throw Error(TN_ERR_INVALID, path + " is truncated");
    sc.triangles.resize(ntris);
    for (size_t i = 0; i < ntris; ++i) {
        Triangle& t = sc.triangles[i];
        for (int k = 0; k < 3; ++k) {
            t.vertices[k] = Vec3{v[9 * i + 3 * k], v[9 * i + 3 * k + 1], v[9 * i + 3 * k + 2]};
            t.normals[k] = Vec3{n[9 * i + 3 * k], n[9 * i + 3 * k + 1], n[9 * i + 3 * k + 2]};
        }
        std::memcpy(&t.diffuse, &a[4 * i], 16);
        std::memcpy(&t.emissive, &e[4 * i], 16);
        if (!rf.empty()) {
            std::memcpy(&t.reflective, &rc[4 * i], 16);
            t.reflectivity = rf[i];
        }
    }
    return sc;
}

Guide::Guide(const KDTree& tree, uint64_t seed, float lr, uint32_t batch, uint32_t epochs, float mis_guided) {
    const float* box = tree.box();
    check(tn_guide_create(box, box + 3, seed, &g_));
    const int rc = tn_guide_set_options(g_, lr, batch, mis_guided, epochs);
    if (rc < 0) {
        const std::string msg = tn_last_error();
        tn_guide_destroy(g_);
        g_ = nullptr;
        throw Error(rc, "turner_amd: " + msg);
    }
}

Guide::~Guide() {
    if (g_) tn_guide_destroy(g_);
}

RenderResult render_guided(const KDTree& tree, const Camera& cam, const std::vector<Light>& lights,
                           const TracerConfig& conf, size_t height, Guide& guide, const GuidedSchedule& sch) {
    if (conf.num_monte_carlo_samples != 1)
        throw Error(TN_ERR_INVALID, "guided rendering trains on one path per sample (-m 1)");
    RenderResult res;
    const size_t np = conf.width * height;
    res.linear.resize(np);
    res.image.resize(np);
    tn_camera c{};
    std::memcpy(c.pos, &cam.position, 12);
    std::memcpy(c.trafo, cam.trafo, sizeof c.trafo);
    c.delta_x = cam.delta_x;
    c.delta_y = cam.delta_y;
    const auto ls = lights_of(lights);
    DeviceBuffer acc(np * 16 * 3);  // image sums, training scratch, tone-mapped
    float* img = acc.as<float>();
    float* scratch = img + np * 4;
    float* shown = img + np * 8;
    if (hipMemset(img, 0, np * 16 * 2) != hipSuccess) throw Error(TN_ERR_HIP, "hipMemset failed");
    tn_render_params p = params_of(conf, conf.width, height);
    uint32_t offset = 0;
    const uint32_t plain_flags = p.flags;
    auto run = [&](int spp, float* into, uint32_t train) {
        p.pixel_samples = spp;
        p.sample_offset = offset;
        p.flags = plain_flags | ((train && sch.defer_training) ? TN_RENDER_DEFER_TRAINING : 0u);
        tn_render_stats st{};
        check(tn_render_guided(tree.handle(), &c, ls.data(), (uint32_t)ls.size(), &p, guide.handle(), train,
                               into, &st, nullptr));
        offset += (uint32_t)spp;
        res.stats.rays += st.rays;
        res.stats.primary_rays += st.primary_rays;
        res.stats.shadow_rays += st.shadow_rays;
        res.stats.ms += st.ms;
    };
    for (int k = 0; k < sch.train_passes; ++k) run(sch.train_spp, scratch, sch.train_samples);
    check(tn_guide_flush(guide.handle(), nullptr));  // a deferred last set trains before the image passes
    int left = conf.num_pixel_samples;
    while (left > 0) {
        const int n = std::min(left, std::max(1, sch.pass_spp));
        run(n, img, 0);
        left -= n;
    }
    std::vector<float> tmp(np * 4);
    to_host(tmp.data(), img, np * 16);
    const float inv = 1.0f / (float)conf.num_pixel_samples;
    for (size_t k = 0; k < np; ++k)
        res.linear[k] = Color{tmp[4 * k] * inv, tmp[4 * k + 1] * inv, tmp[4 * k + 2] * inv, tmp[4 * k + 3] * inv};
    check(tn_tonemap(img, shown, (uint32_t)np, inv, conf.exposure, conf.inverse_gamma,
                     conf.gamma_correction_enabled ? 1 : 0, nullptr));
    to_host(res.image.data(), shown, np * 16);
    return res;
}

}  // namespace turner_amd
