// sol_render: command-line renderer over the C++ API with sol's driver options and flow
// (reference test/driver.cpp:16-245): load a TOML scene (aspect ratio width / height by default),
// make the renderer, run a RenderJob, scale the accumulated image by 1 / spp and save it (EXR by
// default, PFM on request). The renderer is the GPU path tracer, wavefront by default. Options
// beyond sol's: the procedural scenes (--scene cornell|spheres), --renderer, --spf (samples per
// frame of the render job), --device.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <optional>
#include <sstream>
#include <string>

#include "sol/algorithms/gpu_path_tracer.h"
#include "sol/image.h"
#include "sol/render_job.h"
#include "sol/scene_loader.h"
#include "sol/scenes.h"

namespace {

struct Options {
    std::string scene = "";
    std::string out_file = "render.exr";
    sol::Image::Format out_format = sol::Image::Format::Auto;
    std::string renderer = "wavefront";
    size_t width = 1080, height = 720, spp = 16, spf = 0, device = 0;
    size_t max_path_len = 64, min_rr_path_len = 3;
    float min_survival_prob = 0.05f, max_survival_prob = 0.75f, ray_offset = 1e-5f;
};

void usage() {
    Options d;
    std::printf(
        "Usage: sol_render [options] scene.toml\n"
        "Available options:\n"
        "  -h         --help                      Shows this message\n"
        "  -o <file>  --output <file>             Output image file (default: '%s')\n"
        "  -f <fmt>   --format <fmt>              Output format: auto, exr, pfm, png, jpeg, tiff (default: auto)\n"
        "  -a <algo>  --algorithm <algo>          path_tracer (the GPU path tracer)\n"
        "  -spp <n>   --samples-per-pixel <n>     Samples per pixel (default: %zu; also --spp)\n"
        "  -w <n>     --width <n>                 Output width (default: %zu)\n"
        "             --height <n>                Output height (default: %zu)\n"
        "             --max-path-len <len>        Maximum path length (default: %zu)\n"
        "             --min-survival-prob <prob>  Minimum Russian roulette survival probability (default: %g)\n"
        "             --max-survival-prob <prob>  Maximum Russian roulette survival probability (default: %g)\n"
        "             --min-rr-path-len <len>     Path length from which Russian roulette runs (default: %zu)\n"
        "             --ray-offset <off>          Ray offset against self intersection (default: %g)\n"
        "             --scene <file|cornell|spheres>  Scene file or procedural scene\n"
        "             --renderer <wavefront|megakernel>  GPU renderer (default: wavefront)\n"
        "             --spf <n>                   Samples per frame of the render job (default: all)\n"
        "             --device <n>                HIP device (default: 0)\n",
        d.out_file.c_str(), d.spp, d.width, d.height, d.max_path_len, d.min_survival_prob, d.max_survival_prob,
        d.min_rr_path_len, d.ray_offset);
}

bool parse_size(const char* s, size_t& out) {
    char* end = nullptr;
    unsigned long long v = std::strtoull(s, &end, 10);
    if (!end || *end != '\0' || v == 0) return false;
    out = static_cast<size_t>(v);
    return true;
}

bool parse_float(const char* s, float& out) {
    char* end = nullptr;
    out = std::strtof(s, &end);
    return end && *end == '\0';
}

std::optional<Options> parse_options(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto arg = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "Missing argument for option '%s'\n", a.c_str());
                return nullptr;
            }
            return argv[++i];
        };
        auto size_opt = [&](size_t& dst) {
            const char* v = arg();
            if (!v || !parse_size(v, dst)) {
                if (v) std::fprintf(stderr, "Invalid value for option '%s'\n", a.c_str());
                return false;
            }
            return true;
        };
        auto float_opt = [&](float& dst) {
            const char* v = arg();
            if (!v || !parse_float(v, dst)) {
                if (v) std::fprintf(stderr, "Invalid value for option '%s'\n", a.c_str());
                return false;
            }
            return true;
        };
        bool ok = true;
        if (a == "-h" || a == "--help") {
            usage();
            return std::nullopt;
        } else if (a == "-o" || a == "--output") {
            const char* v = arg();
            ok = v != nullptr;
            if (v) o.out_file = v;
        } else if (a == "-f" || a == "--format") {
            const char* v = arg();
            if (!v) return std::nullopt;
            std::string f = v;
            if (f == "auto") o.out_format = sol::Image::Format::Auto;
            else if (f == "exr") o.out_format = sol::Image::Format::Exr;
            else if (f == "pfm") o.out_format = sol::Image::Format::Pfm;
            else if (f == "png") o.out_format = sol::Image::Format::Png;
            else if (f == "jpeg") o.out_format = sol::Image::Format::Jpeg;
            else if (f == "tiff") o.out_format = sol::Image::Format::Tiff;
            else {
                std::fprintf(stderr, "Unknown image format '%s'\n", v);
                return std::nullopt;
            }
        } else if (a == "-a" || a == "--algorithm") {
            const char* v = arg();
            if (!v) return std::nullopt;
            if (std::string(v) != "path_tracer") {
                std::fprintf(stderr, "Unknown rendering algorithm '%s'\n", v);
                return std::nullopt;
            }
        } else if (a == "-w" || a == "--width") {
            ok = size_opt(o.width);
        } else if (a == "--height") {
            ok = size_opt(o.height);
        } else if (a == "-spp" || a == "--samples-per-pixel" || a == "--spp") {
            ok = size_opt(o.spp);
        } else if (a == "--spf") {
            ok = size_opt(o.spf);
        } else if (a == "--max-path-len") {
            ok = size_opt(o.max_path_len);
        } else if (a == "--min-rr-path-len") {
            const char* v = arg();
            ok = v != nullptr;
            if (v) o.min_rr_path_len = std::strtoul(v, nullptr, 10);
        } else if (a == "--min-survival-prob") {
            ok = float_opt(o.min_survival_prob);
        } else if (a == "--max-survival-prob") {
            ok = float_opt(o.max_survival_prob);
        } else if (a == "--ray-offset") {
            ok = float_opt(o.ray_offset);
        } else if (a == "--scene") {
            const char* v = arg();
            ok = v != nullptr;
            if (v) o.scene = v;
        } else if (a == "--renderer") {
            const char* v = arg();
            ok = v != nullptr;
            if (v) o.renderer = v;
        } else if (a == "--device") {
            const char* v = arg();
            ok = v != nullptr;
            if (v) o.device = static_cast<size_t>(std::atoi(v));
        } else if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "Unknown option '%s'\n", a.c_str());
            return std::nullopt;
        } else if (o.scene.empty()) {
            o.scene = a;
        } else {
            std::fprintf(stderr, "Too many input files\n");
            return std::nullopt;
        }
        if (!ok) return std::nullopt;
    }
    if (o.scene.empty()) {
        std::fprintf(stderr, "Missing scene file\nType 'sol_render -h' to show usage\n");
        return std::nullopt;
    }
    if (o.renderer != "wavefront" && o.renderer != "megakernel") {
        std::fprintf(stderr, "Unknown renderer '%s'\n", o.renderer.c_str());
        return std::nullopt;
    }
    if (o.spf == 0) o.spf = o.spp;
    return o;
}

// save_image (driver.cpp:166-190): the format from the extension when automatic; a format that
// cannot be written falls back to the default (EXR).
bool save_image(const sol::Image& image, const Options& o) {
    auto format = o.out_format;
    if (format == sol::Image::Format::Auto) {
        const std::string& f = o.out_file;
        if (f.ends_with(".png")) format = sol::Image::Format::Png;
        if (f.ends_with(".jpg") || f.ends_with(".jpeg")) format = sol::Image::Format::Jpeg;
        if (f.ends_with(".tiff")) format = sol::Image::Format::Tiff;
        if (f.ends_with(".exr")) format = sol::Image::Format::Exr;
        if (f.ends_with(".pfm")) format = sol::Image::Format::Pfm;
    }
    if (!image.save(o.out_file, format)) {
        if (format != sol::Image::Format::Auto && image.save(o.out_file, sol::Image::Format::Auto)) {
            std::printf("Image could not be saved in the given format, so the default format was used instead\n");
            return true;
        }
        std::printf("Could not save image to '%s'\n", o.out_file.c_str());
        return false;
    }
    std::printf("Image was saved to '%s'\n", o.out_file.c_str());
    return true;
}

} // namespace

int main(int argc, char** argv) {
    auto options = parse_options(argc, argv);
    if (!options) return 1;
    const Options& o = *options;
    using sol::GpuPathTracer;
    const GpuPathTracer::Kind kind = o.renderer == "megakernel" ? GpuPathTracer::Kind::Megakernel
                                                                : GpuPathTracer::Kind::Wavefront;
    try {
        const float aspect = static_cast<float>(o.width) / static_cast<float>(o.height);
        sol::bvh::BuildConfig build;
        auto t0 = std::chrono::steady_clock::now();
        std::optional<sol::Scene> scene;
        if (o.scene == "cornell") {
            scene.emplace(sol::scenes::cornell_box(aspect, build));
        } else if (o.scene == "spheres") {
            sol_sphere_field_params p;
            sol_default_sphere_field_params(&p);
            p.aspect = aspect;
            scene.emplace(sol::scenes::sphere_field(p, build));
        } else {
            sol::SceneDefaults defaults;
            defaults.aspect_ratio = aspect;
            std::ostringstream err;
            auto loaded = sol::load_scene(o.scene, defaults, build, &err);
            if (!loaded) {
                std::fprintf(stderr, "%s\n", err.str().c_str());
                return 1;
            }
            scene.emplace(std::move(*loaded));
        }
        std::printf("Scene summary:\n    %zu BSDF(s)\n    %zu light(s)\n    %zu texture(s)\n    %zu image(s)\n",
                    scene->bsdfs.size(), scene->lights.size(), scene->textures.size(), scene->images.size());
        GpuPathTracer::Config config;
        config.max_path_len = o.max_path_len;
        config.min_rr_path_len = o.min_rr_path_len;
        config.min_survival_prob = o.min_survival_prob;
        config.max_survival_prob = o.max_survival_prob;
        config.ray_offset = o.ray_offset;
        GpuPathTracer renderer(*scene, config, kind, static_cast<int>(o.device));
        auto t1 = std::chrono::steady_clock::now();
        std::printf("Scene %s built and uploaded in %lld ms\n", o.scene.c_str(),
                    static_cast<long long>(std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()));

        sol::Image output(o.width, o.height, 3);
        sol::RenderJob job(renderer, output);
        job.sample_count = o.spp;
        job.samples_per_frame = o.spf;
        auto r0 = std::chrono::steady_clock::now();
        job.start();
        std::printf("Rendering started...\n");
        job.wait();
        auto r1 = std::chrono::steady_clock::now();
        std::printf("Rendering finished in %lld ms (%zu frames, %zu samples per pixel)\n",
                    static_cast<long long>(std::chrono::duration_cast<std::chrono::milliseconds>(r1 - r0).count()),
                    job.frames_rendered(), job.samples_rendered());
        if (!o.out_file.empty()) {
            output.scale(1.0f / static_cast<float>(job.sample_count));
            if (!save_image(output, o)) return 1;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
