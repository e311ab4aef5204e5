// Lights, as sol's (reference include/sol/lights.h:47-139, include/sol/shapes.h:49-60):
// a point light and an area light over a uniformly sampled triangle.
#ifndef SOL_LIGHTS_H
#define SOL_LIGHTS_H

#include "sol/math.h"
#include "sol/color.h"
#include "sol/textures.h"
#include "sol_types.h"

namespace sol {

class Light {
public:
    enum class Tag { PointLight, UniformTriangleLight };
    const Tag tag;
    /// Parameter group of the emission for gradient computations (-1: not differentiated).
    int32_t group = -1;

    explicit Light(Tag tag) : tag(tag) {}
    virtual ~Light() = default;
    virtual bool has_area() const = 0;
    virtual sol_light flatten() const = 0;
};

class PointLight final : public Light {
public:
    PointLight(const Vec3f& pos, const Color& intensity)
        : Light(Tag::PointLight), pos_(pos), intensity_(intensity) {}
    bool has_area() const override { return false; }
    sol_light flatten() const override;
private:
    Vec3f pos_;
    Color intensity_;
};

// A triangle with its unit normal and inverse area, sampled uniformly (shapes.h:49-60).
struct UniformTriangle {
    Triangle shape;
    Vec3f normal;
    float inv_area;
    explicit UniformTriangle(const Triangle& t)
        : shape(t), normal(t.normal()), inv_area(1.0f / t.area()) {}
};

class UniformTriangleLight final : public Light {
public:
    UniformTriangleLight(const UniformTriangle& shape, const ColorTexture& intensity)
        : Light(Tag::UniformTriangleLight), shape_(shape), intensity_(intensity) {}
    bool has_area() const override { return true; }
    sol_light flatten() const override;
private:
    UniformTriangle shape_;
    const ColorTexture& intensity_;
};

} // namespace sol

#endif
