// sol_render: command-line renderer over the C++ API, following the flow of sol's driver
// (reference test/driver.cpp:196-245): build a scene, make the renderer, run a RenderJob, scale
// the accumulated image by 1 / spp and save it. The scene is a TOML scene file, loaded as sol's
// driver loads it (Scene::load with the aspect ratio width / height as default), or one of the
// procedural scenes; images are written as PFM.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>

#include "sol/algorithms/gpu_path_tracer.h"
#include "sol/image.h"
#include "sol/render_job.h"
#include "sol/scene_loader.h"
#include "sol/scenes.h"

namespace {

void usage() {
    std::fprintf(stderr,
                 "usage: sol_render [options]\n"
                 "  --scene FILE.toml|cornell|spheres  scene file or procedural scene (default cornell)\n"
                 "  --width W --height H        image size (default 256 x 256)\n"
                 "  --spp N                     samples per pixel (default 16)\n"
                 "  --spf N                     samples per frame of the render job (default spp)\n"
                 "  --max-path-len N            (default 64, as sol's PathTracer::Config)\n"
                 "  --renderer wavefront|megakernel (default wavefront)\n"
                 "  --device D                  HIP device (default 0)\n"
                 "  -o FILE.pfm                 output image\n");
}

bool parse_size(const char* s, size_t& out) {
    char* end = nullptr;
    unsigned long long v = std::strtoull(s, &end, 10);
    if (!end || *end != '\0' || v == 0) return false;
    out = static_cast<size_t>(v);
    return true;
}

} // namespace

int main(int argc, char** argv) {
    std::string scene_name = "cornell", renderer_name = "wavefront", out;
    size_t width = 256, height = 256, spp = 16, spf = 0, max_path_len = 64, device = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto value = [&](size_t& dst) {
            if (i + 1 >= argc || !parse_size(argv[i + 1], dst)) {
                std::fprintf(stderr, "invalid value for %s\n", a.c_str());
                std::exit(1);
            }
            ++i;
        };
        if (a == "--scene" && i + 1 < argc) scene_name = argv[++i];
        else if (a == "--renderer" && i + 1 < argc) renderer_name = argv[++i];
        else if (a == "-o" && i + 1 < argc) out = argv[++i];
        else if (a == "--width") value(width);
        else if (a == "--height") value(height);
        else if (a == "--spp") value(spp);
        else if (a == "--spf") value(spf);
        else if (a == "--max-path-len") value(max_path_len);
        else if (a == "--device") {
            if (i + 1 >= argc) { usage(); return 1; }
            device = static_cast<size_t>(std::atoi(argv[++i]));
        } else {
            usage();
            return a == "-h" || a == "--help" ? 0 : 1;
        }
    }
    if (spf == 0) spf = spp;
    using sol::GpuPathTracer;
    GpuPathTracer::Kind kind;
    if (renderer_name == "wavefront") kind = GpuPathTracer::Kind::Wavefront;
    else if (renderer_name == "megakernel") kind = GpuPathTracer::Kind::Megakernel;
    else { usage(); return 1; }

    try {
        const float aspect = static_cast<float>(width) / static_cast<float>(height);
        sol::bvh::BuildConfig build;
        auto t0 = std::chrono::steady_clock::now();
        const bool from_file = scene_name.size() > 5 && scene_name.ends_with(".toml");
        if (!from_file && scene_name != "cornell" && scene_name != "spheres") {
            usage();
            return 1;
        }
        sol::Scene scene = [&] {
            if (from_file) {
                sol::SceneDefaults defaults;
                defaults.aspect_ratio = aspect;
                std::ostringstream err;
                auto loaded = sol::load_scene(scene_name, defaults, build, &err);
                if (!loaded) throw std::runtime_error("cannot load scene: " + err.str());
                return std::move(*loaded);
            }
            if (scene_name == "cornell") return sol::scenes::cornell_box(aspect, build);
            sol_sphere_field_params p;
            sol_default_sphere_field_params(&p);
            p.aspect = aspect;
            return sol::scenes::sphere_field(p, build);
        }();
        GpuPathTracer::Config config;
        config.max_path_len = max_path_len;
        GpuPathTracer renderer(scene, config, kind, static_cast<int>(device));
        auto t1 = std::chrono::steady_clock::now();
        std::printf("Scene %s built and uploaded in %lld ms\n", scene_name.c_str(),
                    static_cast<long long>(std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()));

        sol::Image output(width, height, 3);
        sol::RenderJob job(renderer, output);
        job.sample_count = spp;
        job.samples_per_frame = spf;
        auto r0 = std::chrono::steady_clock::now();
        job.start();
        job.wait();
        auto r1 = std::chrono::steady_clock::now();
        std::printf("Rendering finished in %lld ms (%zu frames, %zu samples per pixel)\n",
                    static_cast<long long>(std::chrono::duration_cast<std::chrono::milliseconds>(r1 - r0).count()),
                    job.frames_rendered(), job.samples_rendered());
        if (!out.empty()) {
            output.scale(1.0f / static_cast<float>(spp));
            if (!output.save_pfm(out)) {
                std::fprintf(stderr, "cannot write %s\n", out.c_str());
                return 1;
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
