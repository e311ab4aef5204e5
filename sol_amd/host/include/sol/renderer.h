// Rendering-algorithm plugin API, as sol::Renderer (reference include/sol/renderer.h:21-64).
#ifndef SOL_RENDERER_H
#define SOL_RENDERER_H

#include <cstddef>
#include <string>
#include <string_view>

namespace sol {

struct Image;
struct Scene;

class Renderer {
public:
    Renderer(std::string_view name, const Scene& scene) : name_(name), scene_(scene) {}
    virtual ~Renderer() = default;

    const std::string& name() const { return name_; }

    /// Adds the samples [sample_index, sample_index + sample_count) of every pixel into the
    /// image. Deterministic in (pixel, absolute sample index), so any range can be retraced.
    virtual void render(Image& image, size_t sample_index, size_t sample_count = 1) const = 0;

protected:
    std::string name_;
    const Scene& scene_;
};

} // namespace sol

#endif
