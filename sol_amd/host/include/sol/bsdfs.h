// BSDFs, as sol's class family (reference include/sol/bsdfs.h:28-182). On this side a BSDF is
// a description: the shading arithmetic runs on the GPU. flatten() writes the constant-texture
// parameters into the flat record the GPU library reads.
#ifndef SOL_BSDFS_H
#define SOL_BSDFS_H

#include "sol/textures.h"
#include "sol_types.h"

namespace sol {

class Bsdf {
public:
    enum class Tag { DiffuseBsdf, PhongBsdf, MirrorBsdf, GlassBsdf, InterpBsdf };
    // Classification used to skip next-event estimation on specular surfaces
    // (bsdfs.h:39-43, path_tracer.cpp:66 of the reference).
    enum class Type { Diffuse = 0, Glossy = 1, Specular = 2 };

    const Tag tag;
    const Type type;

    /// Parameter group of the material for gradient computations (-1: not differentiated).
    int32_t group = -1;

    Bsdf(Tag tag, Type type) : tag(tag), type(type) {}
    virtual ~Bsdf() = default;

    virtual sol_material flatten() const = 0;
};

class DiffuseBsdf final : public Bsdf {
public:
    explicit DiffuseBsdf(const ColorTexture& kd) : Bsdf(Tag::DiffuseBsdf, Type::Diffuse), kd_(kd) {}
    sol_material flatten() const override;
private:
    const ColorTexture& kd_;
};

class PhongBsdf final : public Bsdf {
public:
    PhongBsdf(const ColorTexture& ks, const Texture& ns) : Bsdf(Tag::PhongBsdf, Type::Glossy), ks_(ks), ns_(ns) {}
    sol_material flatten() const override;
private:
    const ColorTexture& ks_;
    const Texture& ns_;
};

class MirrorBsdf final : public Bsdf {
public:
    explicit MirrorBsdf(const ColorTexture& ks) : Bsdf(Tag::MirrorBsdf, Type::Specular), ks_(ks) {}
    sol_material flatten() const override;
private:
    const ColorTexture& ks_;
};

class GlassBsdf final : public Bsdf {
public:
    GlassBsdf(const ColorTexture& ks, const ColorTexture& kt, const Texture& eta)
        : Bsdf(Tag::GlassBsdf, Type::Specular), ks_(ks), kt_(kt), eta_(eta) {}
    sol_material flatten() const override;
private:
    const ColorTexture& ks_;
    const ColorTexture& kt_;
    const Texture& eta_;
};

/// Mix of two BSDFs with weight k for b (reference bsdfs.cpp:179-230). Its type is Diffuse if
/// either side is, else Glossy if either side is, else Specular. On the GPU path a must be a
/// DiffuseBsdf and b a PhongBsdf - the mix the OBJ loader builds for materials with both a
/// diffuse and a specular colour - and flatten() rejects other pairs.
class InterpBsdf final : public Bsdf {
public:
    InterpBsdf(const Bsdf* a, const Bsdf* b, const Texture& k)
        : Bsdf(Tag::InterpBsdf, infer_type(a->type, b->type)), a_(a), b_(b), k_(k) {}
    sol_material flatten() const override;
    const Bsdf* a() const { return a_; }
    const Bsdf* b() const { return b_; }
private:
    static Type infer_type(Type a, Type b) {
        if (a == Type::Diffuse || b == Type::Diffuse) return Type::Diffuse;
        if (a == Type::Glossy || b == Type::Glossy) return Type::Glossy;
        return Type::Specular;
    }
    const Bsdf* a_;
    const Bsdf* b_;
    const Texture& k_;
};

} // namespace sol

#endif
