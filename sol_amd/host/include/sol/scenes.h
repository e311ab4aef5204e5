// Procedural scenes: the Cornell box, the sphere field and the triangle soup.
#ifndef SOL_SCENES_H
#define SOL_SCENES_H

#include "sol/scene.h"
#include "sol/bvh.h"
#include "sol_host.h"

namespace sol::scenes {

Scene cornell_box(float aspect, const bvh::BuildConfig& config);
Scene sphere_field(const sol_sphere_field_params& params, const bvh::BuildConfig& config);
Scene triangle_soup(uint32_t num_tris, uint64_t seed, const bvh::BuildConfig& config);

// Icosphere of the given subdivision level (20 * 4^level triangles, shared vertices,
// counter-clockwise seen from outside).
void icosphere(uint32_t level, std::vector<Vec3f>& vertices, std::vector<uint32_t>& indices);

} // namespace sol::scenes

#endif
