// Textures, as sol's (reference include/sol/textures.h:56-144). Only constant textures are on
// the GPU path: every configuration uses constant material parameters. A colour texture used
// as a scalar evaluates to its luminance (textures.h:97-99).
#ifndef SOL_TEXTURES_H
#define SOL_TEXTURES_H

#include "sol/math.h"
#include "sol/color.h"

namespace sol {

class Texture {
public:
    enum class Tag { ConstantTexture, ConstantColorTexture };
    const Tag tag;
    explicit Texture(Tag tag) : tag(tag) {}
    virtual ~Texture() = default;
    virtual float sample(const Vec2f& uv) const = 0;
};

class ColorTexture : public Texture {
public:
    explicit ColorTexture(Tag tag) : Texture(tag) {}
    float sample(const Vec2f& uv) const override final { return sample_color(uv).luminance(); }
    virtual Color sample_color(const Vec2f& uv) const = 0;
};

class ConstantTexture final : public Texture {
public:
    explicit ConstantTexture(float constant) : Texture(Tag::ConstantTexture), constant_(constant) {}
    float sample(const Vec2f&) const override { return constant_; }
private:
    float constant_;
};

class ConstantColorTexture final : public ColorTexture {
public:
    explicit ConstantColorTexture(const Color& color)
        : ColorTexture(Tag::ConstantColorTexture), color_(color) {}
    Color sample_color(const Vec2f&) const override { return color_; }
private:
    Color color_;
};

} // namespace sol

#endif
