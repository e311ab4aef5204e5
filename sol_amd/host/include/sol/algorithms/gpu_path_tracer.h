// The GPU path tracer: a sol::Renderer whose render() runs the HIP kernels of libsolhip.so.
// It is the drop-in for sol::PathTracer (reference include/sol/algorithms/path_tracer.h:33-54):
// same Config, same accumulate-into-Image semantics, same per-(pixel, sample) random streams.
#ifndef SOL_ALGORITHMS_GPU_PATH_TRACER_H
#define SOL_ALGORITHMS_GPU_PATH_TRACER_H

#include <memory>

#include "sol/renderer.h"
#include "sol/scene.h"
#include "solhip.h"

namespace sol {

namespace detail {
struct PathTracerConfig {
    size_t max_path_len = 64;
    size_t min_rr_path_len = 3;
    float min_survival_prob = 0.05f;
    float max_survival_prob = 0.75f;
    float ray_offset = 1.e-5f;
};
} // namespace detail

class GpuPathTracer final : public Renderer {
public:
    using Config = detail::PathTracerConfig;
    enum class Kind { Megakernel = SOLHIP_RENDERER_MEGAKERNEL, Wavefront = SOLHIP_RENDERER_WAVEFRONT };

    /// Flattens and uploads the scene. Throws std::runtime_error if the GPU library fails.
    GpuPathTracer(const Scene& scene, const Config& config = {}, Kind kind = Kind::Wavefront, int device = 0);
    ~GpuPathTracer() override;

    void render(Image& image, size_t sample_index, size_t sample_count = 1) const override;

    solhip_ctx* context() const { return ctx_; }
    solhip_scene* gpu_scene() const { return gpu_scene_; }

private:
    Config config_;
    Kind kind_;
    FlatScene flat_;
    solhip_ctx* ctx_ = nullptr;
    solhip_scene* gpu_scene_ = nullptr;
};

} // namespace sol

#endif
