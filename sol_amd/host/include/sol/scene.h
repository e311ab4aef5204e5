// Scene container, as sol::Scene (reference include/sol/scene.h:34-57), and its flattening into
// the arrays the GPU library and the C ABI consume.
#ifndef SOL_SCENE_H
#define SOL_SCENE_H

#include <memory>
#include <vector>

#include "sol/geometry.h"
#include "sol/cameras.h"
#include "sol/bsdfs.h"
#include "sol/lights.h"
#include "sol/textures.h"
#include "sol/image.h"
#include "solhip.h"

namespace sol {

struct Scene {
    Scene() = default;
    Scene(Scene&&) = default;
    ~Scene() = default;

    std::unique_ptr<Geometry> root;
    std::unique_ptr<Camera> camera;

    template <typename T> using unique_vector = std::vector<std::unique_ptr<T>>;
    unique_vector<Bsdf> bsdfs;
    unique_vector<Light> lights;
    unique_vector<Texture> textures;
    unique_vector<Image> images;
};

// Everything a renderer reads, as flat arrays. Materials and lights are numbered in the order
// of Scene::bsdfs and Scene::lights.
struct FlatScene {
    // Raw mesh in original triangle order.
    std::vector<float> positions, normals;
    std::vector<uint32_t> indices;
    std::vector<int32_t> tri_material, tri_light;
    std::vector<sol_material> materials;
    std::vector<sol_light> lights;
    sol_camera_params camera_params{};
    sol_camera camera{};
    // BVH in bvh v2 layout.
    std::vector<sol_bvh_node> bvh_nodes;
    std::vector<uint32_t> prim_indices;
    // GPU layout.
    std::vector<sol_gpu_node> gpu_nodes;
    std::vector<sol_gpu_node4> gpu_nodes4;
    uint32_t bvh4_depth = 0, top_nodes4 = 0;
    std::vector<sol_gpu_tri> gpu_tris;
    std::vector<sol_gpu_shade> gpu_shade;
    uint32_t bvh_depth = 0, top_nodes = 0;
    double flatten_seconds = 0;

    solhip_scene_desc gpu_desc() const;
};

// Throws std::invalid_argument when the scene cannot be rendered (no TriangleMesh root, no
// light, a BSDF or light outside the scene's tables).
FlatScene flatten_scene(const Scene& scene, uint32_t top_levels = 6);

class TriangleMesh;
// The mesh part of flatten_scene: raw arrays, the bvh v2 layout and the GPU layouts (binary and
// 4-wide nodes, leaf-order triangles and shading records). f.tri_material and f.tri_light must
// already hold one entry per triangle (-1 for none), and f.materials the table they index.
void flatten_geometry(const TriangleMesh& mesh, FlatScene& f, uint32_t top_levels = 6);

} // namespace sol

#endif
