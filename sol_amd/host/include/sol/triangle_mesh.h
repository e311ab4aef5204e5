// Triangle mesh with its BVH, as sol::TriangleMesh (reference include/sol/triangle_mesh.h:15-60,
// src/triangle_mesh.cpp:44-60): the constructor takes the same arrays and builds the BVH on
// the host. Flattening for the GPU happens in flatten_scene (scene.h).
#ifndef SOL_TRIANGLE_MESH_H
#define SOL_TRIANGLE_MESH_H

#include <cstddef>
#include <memory>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "sol/geometry.h"
#include "sol/math.h"
#include "sol/bvh.h"

namespace sol {

class Bsdf;
class Light;

class TriangleMesh final : public Geometry {
public:
    struct BuildTimes { double build_seconds = 0, optimize_seconds = 0; float sah_built = 0, sah_final = 0; };

    TriangleMesh(
        std::vector<size_t>&& indices,
        std::vector<Vec3f>&& vertices,
        std::vector<Vec3f>&& normals,
        std::vector<Vec2f>&& tex_coords,
        std::vector<const Bsdf*>&& bsdfs,
        std::unordered_map<size_t, const Light*>&& lights,
        const bvh::BuildConfig& config = {});

    size_t triangle_count() const { return indices_.size() / 3; }
    std::tuple<size_t, size_t, size_t> triangle_indices(size_t t) const {
        return {indices_[t * 3 + 0], indices_[t * 3 + 1], indices_[t * 3 + 2]};
    }

    const std::vector<size_t>& indices() const { return indices_; }
    const std::vector<Vec3f>& vertices() const { return vertices_; }
    const std::vector<Vec3f>& normals() const { return normals_; }
    const std::vector<Vec2f>& tex_coords() const { return tex_coords_; }
    const std::vector<const Bsdf*>& bsdfs() const { return bsdfs_; }
    const std::unordered_map<size_t, const Light*>& lights() const { return lights_; }
    const bvh::Bvh& bvh() const { return bvh_; }
    const BuildTimes& build_times() const { return times_; }

private:
    std::vector<size_t> indices_;
    std::vector<Vec3f> vertices_;
    std::vector<Vec3f> normals_;
    std::vector<Vec2f> tex_coords_;
    std::vector<const Bsdf*> bsdfs_;
    std::unordered_map<size_t, const Light*> lights_;
    bvh::Bvh bvh_;
    BuildTimes times_;
};

} // namespace sol

#endif
