// pathtracer_main.cpp -- the `pathtracer`, `raycaster` and `raytracer` front
// ends (main.cpp:88-244 with the USAGE of pathtracer.h:4-25, raycaster.h:4-22
// and raytracer.h:4-22) over the MI355X renderer; TN_CLI_RENDERER picks the
// trace() plugin at compile time, as the reference links one renderer per
// binary (CMakeLists.txt:62-67).  Scene files are the turner_amd .tns format
// (the reference's assimp .blend import is not available); the image goes to
// stdout as PPM P3, statistics to stderr in the reference's format
// (lib/output.h:101-112).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "turner.hpp"

#ifndef TN_CLI_RENDERER
#define TN_CLI_RENDERER TN_RENDERER_PATHTRACER
#endif

namespace {

constexpr uint32_t kRenderer = TN_CLI_RENDERER;

const char* kCommon =
    R"(
Options:
  -w --width=<px>                   Width of the image [default: 640].
  -a --aspect=<num>                 Aspect ratio of the image. If the model has
                                    specified the aspect ratio, it will be used
                                    [default: 1].
  --background=<3x float>           Background color [default: 0 0 0].
  -t --threads=<int>                Number of threads [default: 1].
  --inverse-gamma=<float>           Inverse of gamma for gamma correction
                                    [default: 0.454545].
  --no-gamma-correction             Disables gamma correction.
  --exposure=<float>                Exposure [default: 1].
  -v --verbose                      Verbose output.
)";

const char* kPathtracer =
    R"(
Pathtracer options:
  -d --max-depth=<int>              Maximum recursion depth for raytracing
                                    [default: 3].
  -p --pixel-samples=<int>          Number of samples per pixel [default: 1].
  -m --monte-carlo-samples=<int>    Monto Carlo samples per ray [default: 8].

MI355X options:
  --seed=<int>                      Seed of the per-path random streams [default: 1].
  --cosine                          Cosine-weighted hemisphere sampling.
  --area-lights                     Emissive triangles: next-event estimation.
  --ref-jitter                      Pixel jitter of the reference (xorshift64*, seed 42 per row).
  --exhaustive                      Reference traversal order without early exit.
  --tree-cache=<path>               kd-tree cache file (validated against the scene).
  --render-tree                     Rendering kd-tree (drops the reference builder's fully clipped
                                    triangles and empty-side quirks; same closest hits).
  --reference-tree                  The reference kd-tree as built, without the refined copy of
                                    it that renders by default (same hits, about 3x slower).
  --graph                           Replay the per-chunk launch sequence from a hipGraph.
  --guided                          Learned path guiding (needs -m 1): trains online, then renders.
  --train-passes=<int>              Guiding: training passes of 1 spp [default: 6].
  --pfm=<path>                      Also write the linear image (before exposure) as a float map.
  --print-config                    Print the parsed configuration and exit.
)";

const char* kRaycaster =
    R"(
Raycaster options:
  --max-visibility=<float>   Any object farther away is dark [default: 2.0].

MI355X options (the pixel jitter is the reference's: xorshift64*, seed 42 per row):
  -p --pixel-samples=<int>          Number of samples per pixel [default: 1].
  --exhaustive                      Reference traversal order without early exit.
  --tree-cache=<path>               kd-tree cache file (validated against the scene).
  --render-tree                     Rendering kd-tree (same closest hits).
  --reference-tree                  The reference kd-tree without its refined copy (same hits).
  --graph                           Replay the per-chunk launch sequence from a hipGraph.
  --pfm=<path>                      Also write the linear image (before exposure) as a float map.
  --print-config                    Print the parsed configuration and exit.
)";

const char* kRaytracer =
    R"(
Raytracer options:
  -d --max-depth=<int>      Maximum recursion depth for raytracing [default: 3].
  --shadow=<float>          Intensity of shadow [default: 0.5].

MI355X options (the pixel jitter is the reference's: xorshift64*, seed 42 per row):
  -p --pixel-samples=<int>          Number of samples per pixel [default: 1].
  --exhaustive                      Reference traversal order without early exit.
  --tree-cache=<path>               kd-tree cache file (validated against the scene).
  --render-tree                     Rendering kd-tree (same closest hits).
  --reference-tree                  The reference kd-tree without its refined copy (same hits).
  --graph                           Replay the per-chunk launch sequence from a hipGraph.
  --pfm=<path>                      Also write the linear image (before exposure) as a float map.
  --print-config                    Print the parsed configuration and exit.
)";

std::string usage() {
    const char* name = kRenderer == TN_RENDERER_RAYCASTER   ? "raycaster"
                       : kRenderer == TN_RENDERER_RAYTRACER ? "raytracer"
                                                            : "pathtracer";
    const char* own = kRenderer == TN_RENDERER_RAYCASTER   ? kRaycaster
                      : kRenderer == TN_RENDERER_RAYTRACER ? kRaytracer
                                                           : kPathtracer;
    return std::string("Usage: ") + name + " <filename> [options]\n" + kCommon + own;
}

struct Args {
    std::map<std::string, std::string> opt;
    std::vector<std::string> pos;
};

// short -> long names of options that take a value, and flags
const std::map<std::string, std::string> kShort = {
    {"-w", "--width"}, {"-a", "--aspect"}, {"-t", "--threads"}, {"-d", "--max-depth"},
    {"-p", "--pixel-samples"}, {"-m", "--monte-carlo-samples"}, {"-v", "--verbose"}};
const char* kFlags[] = {"--no-gamma-correction", "--verbose", "--cosine", "--area-lights",
                        "--ref-jitter", "--exhaustive", "--print-config", "--help",
                        "--render-tree", "--reference-tree", "--graph", "--guided"};

bool is_flag(const std::string& name) {
    for (const char* f : kFlags)
        if (name == f) return true;
    return false;
}

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        if (s.size() > 1 && s[0] == '-' && !(s.size() > 1 && (std::isdigit((unsigned char)s[1]) || s[1] == '.'))) {
            std::string name = s, value;
            const size_t eq = s.find('=');
            if (s.rfind("--", 0) == 0 && eq != std::string::npos) {
                name = s.substr(0, eq);
                value = s.substr(eq + 1);
            } else if (s.rfind("--", 0) != 0 && s.size() > 2) {  // -w320
                name = s.substr(0, 2);
                value = s.substr(2);
            }
            auto it = kShort.find(name);
            if (it != kShort.end()) name = it->second;
            if (is_flag(name)) {
                a.opt[name] = "1";
                continue;
            }
            if (value.empty()) {
                if (i + 1 >= argc) throw std::runtime_error("option " + name + " needs a value");
                value = argv[++i];
            }
            a.opt[name] = value;
        } else {
            a.pos.push_back(s);
        }
    }
    return a;
}

turner_amd::Color parse_color(const std::string& s) {  // config.h:45-60
    std::stringstream ss(s);
    std::vector<float> v;
    std::string item;
    while (std::getline(ss, item, ' '))
        if (!item.empty()) v.push_back(std::stof(item));
    if (v.size() == 1) return {v[0], v[0], v[0], 1};
    if (v.size() == 3) return {v[0], v[1], v[2], 1};
    throw std::runtime_error("--background needs one or three numbers");
}

std::string get(const Args& a, const std::string& k, const std::string& def) {
    auto it = a.opt.find(k);
    return it == a.opt.end() ? def : it->second;
}

}  // namespace

int main(int argc, char** argv) {
    using namespace turner_amd;
    TracerConfig conf;
    bool render_tree = false, reference_tree = false, guided = false;
    std::string pfm;
    int train_passes = 6;
    try {
        Args a = parse(argc, argv);
        if (a.opt.count("--help") || a.pos.size() != 1) {
            std::cout << usage();
            return a.opt.count("--help") ? 0 : 1;
        }
        conf.filename = a.pos[0];
        conf.width = std::stoul(get(a, "--width", "640"));
        conf.aspect = std::stof(get(a, "--aspect", "1"));
        conf.bg_color = parse_color(get(a, "--background", "0 0 0"));
        conf.num_threads = std::stoul(get(a, "--threads", "1"));
        conf.inverse_gamma = std::stof(get(a, "--inverse-gamma", "0.454545"));
        conf.gamma_correction_enabled = !a.opt.count("--no-gamma-correction");
        conf.exposure = std::stof(get(a, "--exposure", "1"));
        conf.verbose = a.opt.count("--verbose") > 0;
        conf.max_recursion_depth = std::stoi(get(a, "--max-depth", "3"));
        conf.num_pixel_samples = std::stoi(get(a, "--pixel-samples", "1"));
        conf.num_monte_carlo_samples = std::stoi(get(a, "--monte-carlo-samples", "8"));
        conf.renderer = kRenderer;
        conf.max_visibility = std::stof(get(a, "--max-visibility", "2.0"));
        conf.shadow_intensity = std::stof(get(a, "--shadow", "0.5"));
        if (kRenderer != TN_RENDERER_PATHTRACER) {
            conf.num_monte_carlo_samples = 1;
            conf.flags |= TN_RENDER_REF_JITTER;
        }
        conf.seed = std::stoull(get(a, "--seed", "1"));
        conf.tree_cache = get(a, "--tree-cache", "");
        if (a.opt.count("--cosine")) conf.flags |= TN_RENDER_COSINE;
        if (a.opt.count("--area-lights")) conf.flags |= TN_RENDER_AREA_LIGHTS;
        if (a.opt.count("--ref-jitter")) conf.flags |= TN_RENDER_REF_JITTER;
        if (a.opt.count("--exhaustive")) conf.flags |= TN_RENDER_EXHAUSTIVE;
        if (a.opt.count("--graph")) conf.flags |= TN_RENDER_GRAPH;
        render_tree = a.opt.count("--render-tree") > 0;
        reference_tree = a.opt.count("--reference-tree") > 0;
        pfm = get(a, "--pfm", "");
        guided = kRenderer == TN_RENDERER_PATHTRACER && a.opt.count("--guided") > 0;
        train_passes = std::stoi(get(a, "--train-passes", "6"));
        if (guided && (conf.num_monte_carlo_samples != 1 || train_passes < 0))
            throw std::runtime_error("--guided needs -m 1 and --train-passes >= 0");
        // config.h:29-33,119-126
        if (!(0 < conf.aspect && 1 <= conf.num_threads && 0 <= conf.exposure &&
              0 < conf.max_recursion_depth && 1 <= conf.num_pixel_samples &&
              0 <= conf.num_monte_carlo_samples && 0 <= conf.max_visibility &&
              0 <= conf.shadow_intensity && conf.shadow_intensity <= 1))
            throw std::runtime_error("invalid configuration");
        if (conf.verbose || a.opt.count("--print-config")) {
            std::cerr << "Filename: " << conf.filename << "\nAspect ratio: " << conf.aspect
                      << "\nImage width: " << conf.width << "\n\nCommon parameters:\n  Number of threads: "
                      << conf.num_threads << "\n  Inverse gamma: " << conf.inverse_gamma
                      << "\n  Exposure: " << conf.exposure << "\n  Gamma correction enabled: "
                      << conf.gamma_correction_enabled
                      << "\n\nTracer parameters (not all applicable):\n  Max recursion depth: "
                      << conf.max_recursion_depth << "\n  Max visibility: " << conf.max_visibility
                      << "\n  Shadow intensity: " << conf.shadow_intensity << "\n  Number of pixel samples: "
                      << conf.num_pixel_samples << "\n  Number of Monte-Carlo samples: "
                      << conf.num_monte_carlo_samples << std::endl;
            if (a.opt.count("--print-config")) return 0;
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n" << usage();
        return 1;
    }
    try {
        std::cerr << "Loading scene..." << std::endl;
        auto t0 = std::chrono::steady_clock::now();
        Scene sc = load_tns(conf.filename);
        // main.cpp:112-116: the scene's aspect wins when it has one
        float aspect = sc.camera.aspect > 0 ? sc.camera.aspect : conf.aspect;
        std::cerr << "Loading triangles and building kd-tree..." << std::endl;
        const size_t ntri = sc.triangles.size();
        // default: the reference tree with its refined copy (TN_SCENE_REFINED),
        // whose renders give the reference tree's hits bit for bit; a cached
        // tree is the reference tree as stored
        KDTree tree = render_tree      ? KDTree(std::move(sc.triangles), (uint32_t)TN_SCENE_EMPTY_LEAVES)
                      : !conf.tree_cache.empty() ? KDTree(std::move(sc.triangles), conf.tree_cache)
                      : reference_tree ? KDTree(std::move(sc.triangles))
                                       : KDTree(std::move(sc.triangles), (uint32_t)TN_SCENE_REFINED);
        auto t1 = std::chrono::steady_clock::now();
        std::cerr << "KDTree runtime: " << (size_t)tree.build_ms() << std::endl;
        const size_t height = (size_t)(conf.width / aspect);
        std::cerr << "Rendering ";
        RenderResult res;
        if (guided) {
            Guide g(tree, conf.seed);
            GuidedSchedule sched;
            sched.train_passes = train_passes;
            res = render_guided(tree, sc.camera, sc.lights, conf, height, g, sched);
        } else {
            res = render(tree, sc.camera, sc.lights, conf, height);
        }
        std::cerr << std::endl;
        const double load_s = std::chrono::duration<double>(t1 - t0).count();
        const double ms = res.stats.ms;
        // lib/output.h:101-112
        std::cerr << "Triangles      : " << ntri << "\nKd-Tree Height : " << tree.height()
                  << "\nRays           : " << res.stats.rays << "\nRays (primary) : "
                  << res.stats.primary_rays << "\nRays/sec       : "
                  << (ms > 0 ? (unsigned long long)(1000.0 * res.stats.rays / ms) : 0ull)
                  << "\nLoading time   : " << load_s << " sec\nRendering time : " << ms / 1000.0
                  << " sec" << std::endl;
        std::cout << to_ppm(res.image, conf.width, height) << std::endl;
        if (!pfm.empty()) write_pfm(pfm, res.linear, conf.width, height);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
