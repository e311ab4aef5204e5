// Triangle mesh with its BVH, as sol::TriangleMesh (reference include/sol/triangle_mesh.h:15-60,
// src/triangle_mesh.cpp:44-60): the constructor takes the same arrays and builds the BVH on
// the host. Flattening for the GPU happens in flatten_scene (scene.h). The Geometry queries
// (src/triangle_mesh.cpp:64-117) run on the GPU: attach_gpu uploads the mesh's BVH and triangles
// to a context, and every query is a solhip_intersect launch whose result is decoded on the host
// as sol decodes its hits. A query on a mesh that is not attached throws std::logic_error: there
// is no CPU traversal behind this interface.
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
#include "solhip.h"

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
    TriangleMesh(const TriangleMesh&) = delete;
    TriangleMesh& operator=(const TriangleMesh&) = delete;
    ~TriangleMesh() override;

    /// Uploads the mesh's BVH and triangles to the GPU context for intersection queries (no
    /// materials or lights), replacing a previous attachment. The context must outlive the
    /// attachment; queries on one context are serialised by the caller. Throws
    /// std::runtime_error if the GPU library fails.
    void attach_gpu(solhip_ctx* ctx);
    void detach_gpu();
    solhip_scene* gpu_geometry() const { return gpu_; }

    /// Geometry queries (triangle_mesh.cpp:64-117), one ray per call.
    std::optional<Hit> intersect_closest(Rayf& ray) const override;
    bool intersect_any(const Rayf& ray) const override;
    /// The same queries for n rays in one GPU launch: hits[i] (and rays[i].tmax on a hit) for
    /// rays[i]; occluded[i] for rays[i].
    void intersect_closest(Rayf* rays, size_t n, std::optional<Hit>* hits) const;
    void intersect_any(const Rayf* rays, size_t n, bool* occluded) const;
    /// The hit record of triangle t (original index) at barycentrics (u, v) for a ray whose tmax
    /// is the hit distance, as triangle_mesh.cpp:78-104 builds it: shading normal and texture
    /// coordinates interpolated, both normals turned towards the ray, the frame of the shading
    /// normal, the triangle's emitter and BSDF.
    Hit decode_hit(const Rayf& ray, size_t t, float u, float v) const;

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
    solhip_scene* gpu_ = nullptr;   // the attached GPU copy, owned
};

} // namespace sol

#endif
