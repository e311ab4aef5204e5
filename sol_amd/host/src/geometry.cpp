// TriangleMesh's Geometry queries (reference src/triangle_mesh.cpp:64-117) over the GPU: the mesh's
// BVH and triangles are uploaded once (attach_gpu), each query batch is one solhip_intersect
// launch, and every closest hit is decoded on the host as sol decodes it.
#include <stdexcept>
#include <string>
#include <vector>

#include "sol/scene.h"
#include "sol/triangle_mesh.h"

namespace sol {

namespace {

void check(int status, const char* what) {
    if (status != 0) throw std::runtime_error(std::string(what) + ": " + solhip_last_error());
}

sol_ray to_sol_ray(const Rayf& r) {
    sol_ray o;
    o.org[0] = r.org.x; o.org[1] = r.org.y; o.org[2] = r.org.z;
    o.dir[0] = r.dir.x; o.dir[1] = r.dir.y; o.dir[2] = r.dir.z;
    o.tmin = r.tmin;
    o.tmax = r.tmax;
    return o;
}

} // namespace

TriangleMesh::~TriangleMesh() { detach_gpu(); }

void TriangleMesh::attach_gpu(solhip_ctx* ctx) {
    if (!ctx) throw std::invalid_argument("attach_gpu needs a context");
    FlatScene f;
    f.tri_material.assign(triangle_count(), -1);
    f.tri_light.assign(triangle_count(), -1);
    flatten_geometry(*this, f);
    solhip_scene_desc desc = f.gpu_desc();   // no materials, no lights: intersection only
    solhip_scene* s = nullptr;
    check(solhip_scene_upload(ctx, &desc, &s), "solhip_scene_upload");
    detach_gpu();
    gpu_ = s;
}

void TriangleMesh::detach_gpu() {
    if (gpu_) solhip_scene_destroy(gpu_);
    gpu_ = nullptr;
}

Hit TriangleMesh::decode_hit(const Rayf& ray, size_t t, float u, float v) const {
    auto [i0, i1, i2] = triangle_indices(t);
    const Vec3f& p0 = vertices_[i0];
    Vec3f face_normal = cross(vertices_[i1] - p0, vertices_[i2] - p0);
    Vec3f normal = lerp(normals_[i0], normals_[i1], normals_[i2], u, v);
    Vec2f tex = tex_coords_.size() == vertices_.size() ? lerp(tex_coords_[i0], tex_coords_[i1], tex_coords_[i2], u, v)
                                                       : Vec2f(0.0f, 0.0f);
    // both normals face the side the ray comes from (triangle_mesh.cpp:86-91)
    const bool front = dot(face_normal, ray.dir) < 0.0f;
    if (!front) {
        face_normal = -face_normal;
        normal = -normal;
    }
    Hit hit;
    hit.surf_info.is_front_side = front;
    hit.surf_info.point = ray.point_at(ray.tmax);
    hit.surf_info.tex_coords = tex;
    hit.surf_info.surf_coords = Vec2f(u, v);
    hit.surf_info.face_normal = face_normal;
    hit.surf_info.local = ortho_basis(normalize(normal));
    if (auto it = lights_.find(t); it != lights_.end()) hit.light = it->second;
    hit.bsdf = bsdfs_[t];
    return hit;
}

void TriangleMesh::intersect_closest(Rayf* rays, size_t n, std::optional<Hit>* hits) const {
    if (n == 0) return;
    if (!gpu_) throw std::logic_error("TriangleMesh::intersect_closest: the mesh is not attached to a GPU context");
    std::vector<sol_ray> in(n);
    for (size_t i = 0; i < n; ++i) in[i] = to_sol_ray(rays[i]);
    std::vector<sol_hit> out(n);
    check(solhip_intersect(gpu_, in.data(), n, 0, out.data(), 0), "solhip_intersect");
    for (size_t i = 0; i < n; ++i) {
        if (out[i].prim < 0) {
            hits[i] = std::nullopt;
            continue;
        }
        rays[i].tmax = out[i].t;
        hits[i] = decode_hit(rays[i], static_cast<size_t>(out[i].prim), out[i].u, out[i].v);
    }
}

void TriangleMesh::intersect_any(const Rayf* rays, size_t n, bool* occluded) const {
    if (n == 0) return;
    if (!gpu_) throw std::logic_error("TriangleMesh::intersect_any: the mesh is not attached to a GPU context");
    std::vector<sol_ray> in(n);
    for (size_t i = 0; i < n; ++i) in[i] = to_sol_ray(rays[i]);
    std::vector<int32_t> out(n);
    check(solhip_intersect(gpu_, in.data(), n, 1, out.data(), 0), "solhip_intersect");
    for (size_t i = 0; i < n; ++i) occluded[i] = out[i] != 0;
}

std::optional<Hit> TriangleMesh::intersect_closest(Rayf& ray) const {
    std::optional<Hit> hit;
    intersect_closest(&ray, 1, &hit);
    return hit;
}

bool TriangleMesh::intersect_any(const Rayf& ray) const {
    bool occluded = false;
    intersect_any(&ray, 1, &occluded);
    return occluded;
}

} // namespace sol
