// Pinhole camera, as sol::PerspectiveCamera (reference include/sol/cameras.h:39-56,
// src/cameras.cpp:5-25). The GPU generates the rays; this side computes the basis.
#ifndef SOL_CAMERAS_H
#define SOL_CAMERAS_H

#include "sol/math.h"
#include "sol_types.h"

namespace sol {

class Camera {
public:
    virtual ~Camera() = default;
    virtual sol_camera basis() const = 0;
    virtual sol_camera_params params() const = 0;
};

class PerspectiveCamera final : public Camera {
public:
    PerspectiveCamera(const Vec3f& eye, const Vec3f& dir, const Vec3f& up, float horz_fov, float aspect_ratio);
    sol_camera basis() const override;
    sol_camera_params params() const override { return params_; }
private:
    sol_camera_params params_;
};

// cameras.cpp:5-21: dir normalised; right = normalize(cross(dir, up)) * w;
// up' = cross(right_n, dir_n) * h; w = tan(fov * pi / 360), h = w / aspect.
sol_camera camera_basis(const sol_camera_params& p);

} // namespace sol

#endif
