// The intersection boundary, as sol::Geometry (reference include/sol/geometry.h:14-43): the
// surface record of a hit, and closest / any-hit queries that a scene's root answers. sol answers
// them on the CPU with its bvh v2 traverser; here TriangleMesh answers them on the GPU through
// solhip_intersect (one ray per call as sol's interface has it, or many per call through the
// batched overloads), once the mesh is attached to a GPU context.
#ifndef SOL_GEOMETRY_H
#define SOL_GEOMETRY_H

#include <optional>

#include "sol/math.h"

namespace sol {

class Bsdf;
class Light;

/// Surface information at a hit point (geometry.h:14-24): the side of the surface the ray came
/// from, the point, texture and surface coordinates, the geometric normal and the local frame of
/// the shading normal (third column), both turned towards the incoming ray.
struct SurfaceInfo {
    bool is_front_side = false;
    Vec3f point;
    Vec2f tex_coords;
    Vec2f surf_coords;
    Vec3f face_normal;
    Mat3x3f local;

    Vec3f normal() const { return local.col(2); }
};

/// Result of a closest-hit query (geometry.h:27-31).
struct Hit {
    SurfaceInfo surf_info;
    const Light* light = nullptr;   ///< emitter at the hit point, if any
    const Bsdf* bsdf = nullptr;     ///< BSDF at the hit point, if any
};

class Geometry {
public:
    virtual ~Geometry() = default;

    /// The closest hit along the ray within (tmin, tmax), or nothing; on a hit the ray's tmax
    /// becomes the hit distance (geometry.h:36-39).
    virtual std::optional<Hit> intersect_closest(Rayf& ray) const = 0;
    /// Whether anything lies on the ray within (tmin, tmax) (geometry.h:40-41).
    virtual bool intersect_any(const Rayf& ray) const = 0;
};

} // namespace sol

#endif
