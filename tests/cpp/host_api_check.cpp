// host_api_check.cpp -- calls the C++ host interface (turner_amd/host/turner.hpp)
// the way the reference's callers do, for tests/test_gpu_host_api.py to check
// against the oracle:
//
//   intersect: KDTreeIntersection::intersect(ray, r, a, b) one ray at a time
//              (lib/kdtree.h:291; r, a, b written on a hit only) and the
//              batched twin;
//   trace:     trace(ray, ti, lights, depth, conf) one ray at a time (the call
//              of main.cpp:212-214) and the batched twin;
//   pixels:    the pixel loop of main.cpp:193-216 (jitter xorshift64* seeded
//              42 per row, raster2cam of lib/types.h:119-123, one trace() call
//              per pixel sample, the mean per pixel).
//
// Files are raw little-endian arrays.  rays.bin: uint32 n, then n x 3 float
// origins, n x 3 float directions.
//
// usage:
//   host_api_check intersect <scene.tns> <rays.bin> <out.bin> <singles> [scene_flags]
//       out: n uint32 ids, n x 3 float (r, s, t) batched; then `singles` x
//       (uint32 id, float r, a, b) of the one-ray calls (r, a, b start as NaN);
//       scene_flags: TN_SCENE_* for KDTree(triangles, flags) (default: KDTree(triangles))
//   host_api_check trace <scene.tns> <rays.bin> <out.bin> <singles> <depth> <max_depth>
//                  <mc> <flags> <seed>
//       out: n x RGBA batched, then `singles` x RGBA of one-ray calls
//   host_api_check pixels <scene.tns> <out.bin> <width> <height> <spp> <max_depth> <mc> <flags>
//       out: width x height x RGBA means
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "turner.hpp"

using namespace turner_amd;

namespace {

std::vector<Ray> read_rays(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    uint32_t n = 0;
    f.read(reinterpret_cast<char*>(&n), 4);
    std::vector<float> od(6 * (size_t)n);
    f.read(reinterpret_cast<char*>(od.data()), (std::streamsize)(od.size() * 4));
    if (!f) throw Error(TN_ERR_INVALID, "cannot read " + path);
    std::vector<Ray> rays(n);
    for (uint32_t i = 0; i < n; ++i) {
        rays[i].o = Vec3{od[3 * i], od[3 * i + 1], od[3 * i + 2]};
        rays[i].d = Vec3{od[3 * (n + i)], od[3 * (n + i) + 1], od[3 * (n + i) + 2]};
    }
    return rays;
}

void write(std::ofstream& f, const void* p, size_t bytes) { f.write(static_cast<const char*>(p), (std::streamsize)bytes); }

// xorshift64*, float = low 24 bits * 2^-24 (lib/xorshift.h:36-62)
float jitter(uint64_t& s) {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return std::ldexp((float)((s * 2685821657736338717ull) & 0xFFFFFFull), -24);
}

int run(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: host_api_check intersect|trace|pixels ...\n");
        return 2;
    }
    const std::string mode = argv[1];
    Scene sc = load_tns(argv[2]);
    const uint32_t scene_flags =
        (mode == "intersect" && argc == 7) ? (uint32_t)std::strtoul(argv[6], nullptr, 10) : 0u;
    KDTree tree = scene_flags ? KDTree(sc.triangles, scene_flags) : KDTree(sc.triangles);
    KDTreeIntersection ti(tree);
    if (mode == "intersect" && (argc == 6 || argc == 7)) {
        const std::vector<Ray> rays = read_rays(argv[3]);
        const size_t singles = std::min<size_t>(std::strtoul(argv[5], nullptr, 10), rays.size());
        std::vector<uint32_t> ids;
        std::vector<float> r, s, t;
        ti.intersect(rays, ids, r, s, t);
        std::ofstream f(argv[4], std::ios::binary);
        write(f, ids.data(), 4 * ids.size());
        write(f, r.data(), 4 * r.size());
        write(f, s.data(), 4 * s.size());
        write(f, t.data(), 4 * t.size());
        for (size_t i = 0; i < singles; ++i) {
            float rr = NAN, a = NAN, b = NAN;
            const OptionalId id = ti.intersect(rays[i], rr, a, b);
            const uint32_t v = id ? id.value() : TN_INVALID_ID;
            write(f, &v, 4);
            write(f, &rr, 4);
            write(f, &a, 4);
            write(f, &b, 4);
        }
        return f ? 0 : 1;
    }
    if (mode == "trace" && argc == 11) {
        const std::vector<Ray> rays = read_rays(argv[3]);
        const size_t singles = std::min<size_t>(std::strtoul(argv[5], nullptr, 10), rays.size());
        TracerConfig conf;
        const int depth = std::atoi(argv[6]);
        conf.max_recursion_depth = std::atoi(argv[7]);
        conf.num_monte_carlo_samples = std::atoi(argv[8]);
        conf.flags = (uint32_t)std::strtoul(argv[9], nullptr, 10);
        conf.seed = std::strtoull(argv[10], nullptr, 10);
        const std::vector<Color> batch = trace(rays, ti, sc.lights, depth, conf);
        std::ofstream f(argv[4], std::ios::binary);
        write(f, batch.data(), 16 * batch.size());
        for (size_t i = 0; i < singles; ++i) {
            const Color c = trace(rays[i], ti, sc.lights, depth, conf);
            write(f, &c, 16);
        }
        return f ? 0 : 1;
    }
    if (mode == "pixels" && argc == 10) {
        const size_t w = std::strtoul(argv[4], nullptr, 10), h = std::strtoul(argv[5], nullptr, 10);
        const int spp = std::atoi(argv[6]);
        TracerConfig conf;
        conf.width = w;
        conf.max_recursion_depth = std::atoi(argv[7]);
        conf.num_monte_carlo_samples = std::atoi(argv[8]);
        conf.flags = (uint32_t)std::strtoul(argv[9], nullptr, 10);
        const Camera& cam = sc.camera;
        std::vector<Color> img(w * h);
        for (size_t y = 0; y < h; ++y) {
            uint64_t gen = 42;  // main.cpp:201
            for (size_t x = 0; x < w; ++x) {
                Color acc;
                for (int k = 0; k < spp; ++k) {
                    const float dx = jitter(gen), dy = jitter(gen);
                    const float px = (float)x + dx, py = (float)y + dy;
                    const float cx = -cam.delta_x * (1 - 2 * px / (float)w);
                    const float cy = cam.delta_y * (1 - 2 * py / (float)h);
                    const float cz = -1;
                    const float* m = cam.trafo;
                    Ray ray;
                    ray.o = cam.position;
                    ray.d = Vec3{m[0] * cx + m[1] * cy + m[2] * cz, m[3] * cx + m[4] * cy + m[5] * cz,
                                 m[6] * cx + m[7] * cy + m[8] * cz};
                    const Color c = trace(ray, ti, sc.lights, 0, conf);  // main.cpp:212-214
                    acc.r += c.r;
                    acc.g += c.g;
                    acc.b += c.b;
                    acc.a += c.a;
                }
                const float inv = 1.0f / (float)spp;
                img[y * w + x] = Color{acc.r * inv, acc.g * inv, acc.b * inv, acc.a * inv};
            }
        }
        std::ofstream f(argv[3], std::ios::binary);
        write(f, img.data(), 16 * img.size());
        return f ? 0 : 1;
    }
    std::fprintf(stderr, "bad arguments for mode %s\n", mode.c_str());
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
}
