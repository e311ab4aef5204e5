// GpuPathTracer: the sol::Renderer over libsolhip.so.
#include <stdexcept>
#include <string>
#include <vector>

#include "sol/algorithms/gpu_path_tracer.h"
#include "sol/image.h"

namespace sol {

namespace {
void check(int status, const char* what) {
    if (status != 0)
        throw std::runtime_error(std::string(what) + ": " + solhip_last_error());
}
} // namespace

GpuPathTracer::GpuPathTracer(const Scene& scene, const Config& config, Kind kind, int device)
    : Renderer("GpuPathTracer", scene), config_(config), kind_(kind), flat_(flatten_scene(scene)) {
    check(solhip_create(device, &ctx_), "solhip_create");
    solhip_scene_desc desc = flat_.gpu_desc();
    int st = solhip_scene_upload(ctx_, &desc, &gpu_scene_);
    if (st != 0) {
        solhip_destroy(ctx_);
        ctx_ = nullptr;
        check(st, "solhip_scene_upload");
    }
}

GpuPathTracer::~GpuPathTracer() {
    if (gpu_scene_) solhip_scene_destroy(gpu_scene_);
    if (ctx_) solhip_destroy(ctx_);
}

void GpuPathTracer::render(Image& image, size_t sample_index, size_t sample_count) const {
    if (image.channel_count() != 3) throw std::invalid_argument("GpuPathTracer renders 3-channel images");
    if (sample_count == 0) return;
    sol_pt_config cfg{};
    cfg.max_path_len = static_cast<uint32_t>(config_.max_path_len);
    cfg.min_rr_path_len = static_cast<uint32_t>(config_.min_rr_path_len);
    cfg.min_survival_prob = config_.min_survival_prob;
    cfg.max_survival_prob = config_.max_survival_prob;
    cfg.ray_offset = config_.ray_offset;
    size_t w = image.width(), h = image.height(), n = w * h;
    // The render adds each pixel's sum of samples into a zeroed planar image, which is then added
    // to the image's channels: image + (0 + s) is image + s bit for bit (a sum is never -0), and the
    // image is not copied in.
    std::vector<float> planar(3 * n);
    check(solhip_render(gpu_scene_, &cfg, static_cast<int>(kind_), static_cast<uint32_t>(w), static_cast<uint32_t>(h),
                        sample_index, static_cast<uint32_t>(sample_count), planar.data(), 0),
          "solhip_render");
    for (int c = 0; c < 3; ++c) {
        float* dst = image.channel(c).get();
        const float* src = planar.data() + c * n;
        for (size_t i = 0; i < n; ++i) dst[i] += src[i];
    }
}

} // namespace sol
