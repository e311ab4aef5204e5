// Scene model: BSDF/light/camera descriptions, the triangle mesh with its BVH, images, and the
// flattening of a Scene into the arrays of the C ABI.
#include <chrono>
#include <cstdio>
#include <numbers>
#include <stdexcept>
#include <unordered_map>

#include "sol/scene.h"
#include "sol/triangle_mesh.h"

namespace sol {

namespace {
void copy3(float* dst, const Vec3f& v) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; }
void copy3(float* dst, const Color& c) { dst[0] = c.r; dst[1] = c.g; dst[2] = c.b; }
sol_material blank_material(int32_t type) {
    sol_material m{};
    m.type = type;
    m.group = -1;
    return m;
}
double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
} // namespace

// -- BSDFs: constant textures are evaluated once (every texture on the GPU path is constant) --

sol_material DiffuseBsdf::flatten() const {
    sol_material m = blank_material(SOL_BSDF_DIFFUSE);
    copy3(m.kd, kd_.sample_color(Vec2f(0, 0)));
    return m;
}

sol_material PhongBsdf::flatten() const {
    sol_material m = blank_material(SOL_BSDF_PHONG);
    copy3(m.ks, ks_.sample_color(Vec2f(0, 0)));
    m.ns = ns_.sample(Vec2f(0, 0));
    m.ns_sample = m.ns;
    return m;
}

sol_material MirrorBsdf::flatten() const {
    sol_material m = blank_material(SOL_BSDF_MIRROR);
    copy3(m.ks, ks_.sample_color(Vec2f(0, 0)));
    return m;
}

sol_material GlassBsdf::flatten() const {
    sol_material m = blank_material(SOL_BSDF_GLASS);
    copy3(m.ks, ks_.sample_color(Vec2f(0, 0)));
    copy3(m.kt, kt_.sample_color(Vec2f(0, 0)));
    m.eta = eta_.sample(Vec2f(0, 0));
    return m;
}

sol_material InterpBsdf::flatten() const {
    if (!a_ || !b_ || a_->tag != Tag::DiffuseBsdf || b_->tag != Tag::PhongBsdf)
        throw std::invalid_argument("InterpBsdf: the GPU path mixes a DiffuseBsdf (a) and a PhongBsdf (b) only");
    sol_material d = a_->flatten(), p = b_->flatten();
    sol_material m = blank_material(SOL_BSDF_INTERP);
    for (int c = 0; c < 3; ++c) {
        m.kd[c] = d.kd[c];
        m.ks[c] = p.ks[c];
    }
    m.ns = p.ns;
    m.ns_sample = p.ns_sample;
    m.mix = k_.sample(Vec2f(0, 0));
    return m;
}

// -- Lights -------------------------------------------------------------------------------------

sol_light PointLight::flatten() const {
    sol_light l{};
    l.type = SOL_LIGHT_POINT;
    copy3(l.v0, pos_);
    copy3(l.intensity, intensity_);
    l.inv_area = 1.0f;
    l.group = -1;
    return l;
}

sol_light UniformTriangleLight::flatten() const {
    sol_light l{};
    l.type = SOL_LIGHT_TRIANGLE;
    copy3(l.v0, shape_.shape.v0);
    copy3(l.v1, shape_.shape.v1);
    copy3(l.v2, shape_.shape.v2);
    copy3(l.normal, shape_.normal);
    l.inv_area = shape_.inv_area;
    copy3(l.intensity, intensity_.sample_color(Vec2f(0, 0)));
    l.group = -1;
    return l;
}

// -- Camera ---------------------------------------------------------------------------------------

sol_camera camera_basis(const sol_camera_params& p) {
    Vec3f eye(p.eye[0], p.eye[1], p.eye[2]);
    Vec3f dir(p.dir[0], p.dir[1], p.dir[2]);
    Vec3f up(p.up[0], p.up[1], p.up[2]);
    Vec3f dir_n = normalize(dir);
    Vec3f right = normalize(cross(dir, up));
    Vec3f up_n = cross(right, dir_n);
    float w = std::tan(p.fov * std::numbers::pi_v<float> / 360.0f);
    float h = w / p.aspect;
    right *= w;
    up_n *= h;
    sol_camera c{};
    copy3(c.eye, eye);
    copy3(c.dir, dir_n);
    copy3(c.right, right);
    copy3(c.up, up_n);
    return c;
}

PerspectiveCamera::PerspectiveCamera(const Vec3f& eye, const Vec3f& dir, const Vec3f& up, float horz_fov, float aspect_ratio) {
    copy3(params_.eye, eye);
    copy3(params_.dir, dir);
    copy3(params_.up, up);
    params_.fov = horz_fov;
    params_.aspect = aspect_ratio;
}

sol_camera PerspectiveCamera::basis() const { return camera_basis(params_); }

// -- Triangle mesh --------------------------------------------------------------------------------

TriangleMesh::TriangleMesh(
    std::vector<size_t>&& indices,
    std::vector<Vec3f>&& vertices,
    std::vector<Vec3f>&& normals,
    std::vector<Vec2f>&& tex_coords,
    std::vector<const Bsdf*>&& bsdfs,
    std::unordered_map<size_t, const Light*>&& lights,
    const bvh::BuildConfig& config)
    : indices_(std::move(indices))
    , vertices_(std::move(vertices))
    , normals_(std::move(normals))
    , tex_coords_(std::move(tex_coords))
    , bsdfs_(std::move(bsdfs))
    , lights_(std::move(lights))
{
    size_t n = triangle_count();
    if (n == 0 || indices_.size() % 3 != 0) throw std::invalid_argument("mesh needs at least one triangle");
    if (normals_.size() != vertices_.size()) throw std::invalid_argument("mesh needs one normal per vertex");
    if (bsdfs_.size() != n) throw std::invalid_argument("mesh needs one BSDF pointer per triangle");
    for (size_t i : indices_)
        if (i >= vertices_.size()) throw std::invalid_argument("vertex index out of range");

    auto t0 = std::chrono::steady_clock::now();
    std::vector<BBox> bboxes(n);
    std::vector<Vec3f> centers(n);
    #pragma omp parallel for schedule(static)
    for (long long i = 0; i < static_cast<long long>(n); ++i) {
        Triangle t{vertices_[indices_[i * 3 + 0]], vertices_[indices_[i * 3 + 1]], vertices_[indices_[i * 3 + 2]]};
        bboxes[i] = t.bbox();
        centers[i] = t.center();
    }
    bvh_ = bvh::build_sweep_sah(bboxes.data(), centers.data(), n, config);
    times_.build_seconds = seconds_since(t0);
    times_.sah_built = bvh_.sah_cost(config.traversal_cost, config.intersection_cost);
    auto t1 = std::chrono::steady_clock::now();
    bvh::optimize_reinsertion(bvh_, config);
    times_.optimize_seconds = seconds_since(t1);
    times_.sah_final = bvh_.sah_cost(config.traversal_cost, config.intersection_cost);
}

// -- Flattening -------------------------------------------------------------------------------------

solhip_scene_desc FlatScene::gpu_desc() const {
    solhip_scene_desc d{};
    d.num_nodes = static_cast<uint32_t>(gpu_nodes.size());
    d.nodes = gpu_nodes.data();
    d.num_tris = static_cast<uint32_t>(gpu_tris.size());
    d.tris = gpu_tris.data();
    d.shade = gpu_shade.data();
    d.num_materials = static_cast<uint32_t>(materials.size());
    d.materials = materials.data();
    d.num_lights = static_cast<uint32_t>(lights.size());
    d.lights = lights.data();
    d.camera = camera;
    d.bvh_depth = bvh_depth;
    d.top_nodes = top_nodes;
    d.num_nodes4 = static_cast<uint32_t>(gpu_nodes4.size());
    d.nodes4 = gpu_nodes4.data();
    d.bvh4_depth = bvh4_depth;
    d.top_nodes4 = top_nodes4;
    return d;
}

FlatScene flatten_scene(const Scene& scene, uint32_t top_levels) {
    auto t0 = std::chrono::steady_clock::now();
    const auto* mesh = dynamic_cast<const TriangleMesh*>(scene.root.get());
    if (!mesh) throw std::invalid_argument("scene root must be a TriangleMesh");
    if (!scene.camera) throw std::invalid_argument("scene has no camera");
    if (scene.lights.empty()) throw std::invalid_argument("scene has no light");

    FlatScene f;
    std::unordered_map<const Bsdf*, int32_t> bsdf_index;
    for (size_t i = 0; i < scene.bsdfs.size(); ++i) {
        bsdf_index[scene.bsdfs[i].get()] = static_cast<int32_t>(i);
        f.materials.push_back(scene.bsdfs[i]->flatten());
        f.materials.back().group = scene.bsdfs[i]->group;
    }
    std::unordered_map<const Light*, int32_t> light_index;
    for (size_t i = 0; i < scene.lights.size(); ++i) {
        light_index[scene.lights[i].get()] = static_cast<int32_t>(i);
        f.lights.push_back(scene.lights[i]->flatten());
        f.lights.back().group = scene.lights[i]->group;
    }
    f.camera_params = scene.camera->params();
    f.camera = scene.camera->basis();

    size_t nt = mesh->triangle_count();
    f.tri_material.assign(nt, -1);
    f.tri_light.assign(nt, -1);
    for (size_t t = 0; t < nt; ++t) {
        const Bsdf* b = mesh->bsdfs()[t];
        if (b) {
            auto it = bsdf_index.find(b);
            if (it == bsdf_index.end()) throw std::invalid_argument("mesh BSDF is not owned by the scene");
            f.tri_material[t] = it->second;
        }
    }
    for (auto& [t, l] : mesh->lights()) {
        auto it = light_index.find(l);
        if (t >= nt || it == light_index.end()) throw std::invalid_argument("mesh light is not owned by the scene");
        f.tri_light[t] = it->second;
    }
    flatten_geometry(*mesh, f, top_levels);
    f.flatten_seconds = seconds_since(t0);
    return f;
}

void flatten_geometry(const TriangleMesh& mesh_ref, FlatScene& f, uint32_t top_levels) {
    const TriangleMesh* mesh = &mesh_ref;
    size_t nt = mesh->triangle_count();
    size_t nv = mesh->vertices().size();
    if (f.tri_material.size() != nt || f.tri_light.size() != nt)
        throw std::invalid_argument("per-triangle material and light tables must match the mesh");
    f.positions.resize(nv * 3);
    f.normals.resize(nv * 3);
    for (size_t i = 0; i < nv; ++i) {
        copy3(&f.positions[i * 3], mesh->vertices()[i]);
        copy3(&f.normals[i * 3], mesh->normals()[i]);
    }
    f.indices.resize(nt * 3);
    for (size_t i = 0; i < nt * 3; ++i) f.indices[i] = static_cast<uint32_t>(mesh->indices()[i]);

    const bvh::Bvh& b = mesh->bvh();
    f.bvh_nodes.resize(b.nodes.size());
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        const auto& nd = b.nodes[i];
        sol_bvh_node& o = f.bvh_nodes[i];
        o.bounds[0] = nd.box.lo.x; o.bounds[1] = nd.box.hi.x;
        o.bounds[2] = nd.box.lo.y; o.bounds[3] = nd.box.hi.y;
        o.bounds[4] = nd.box.lo.z; o.bounds[5] = nd.box.hi.z;
        o.index = nd.index;
        o.prim_count = nd.prim_count;
    }
    f.prim_indices = b.prim_indices;

    bvh::GpuLayout layout = bvh::to_gpu_layout(b, top_levels);
    f.gpu_nodes = std::move(layout.nodes);
    f.top_nodes = layout.top_nodes;
    f.bvh_depth = layout.depth;
    // 4-wide: four levels breadth-first (at most 85 nodes; the kernels stage up to 80 in LDS)
    bvh::GpuLayout4 layout4 = bvh::to_gpu_layout4(b, 4);
    f.gpu_nodes4 = std::move(layout4.nodes);
    f.top_nodes4 = layout4.top_nodes;
    f.bvh4_depth = layout4.depth;
    f.gpu_tris.resize(nt);
    f.gpu_shade.resize(nt);
    #pragma omp parallel for schedule(static)
    for (long long slot = 0; slot < static_cast<long long>(nt); ++slot) {
        uint32_t t = b.prim_indices[slot];
        sol_gpu_tri& g = f.gpu_tris[slot];
        sol_gpu_shade& s = f.gpu_shade[slot];
        g = sol_gpu_tri{};
        s = sol_gpu_shade{};
        uint32_t i0 = f.indices[t * 3 + 0], i1 = f.indices[t * 3 + 1], i2 = f.indices[t * 3 + 2];
        for (int k = 0; k < 3; ++k) {
            g.v0[k] = f.positions[i0 * 3 + k];
            g.v1[k] = f.positions[i1 * 3 + k];
            g.v2[k] = f.positions[i2 * 3 + k];
            s.n0[k] = f.normals[i0 * 3 + k];
            s.n1[k] = f.normals[i1 * 3 + k];
            s.n2[k] = f.normals[i2 * 3 + k];
        }
        g.prim = static_cast<int32_t>(t);
        Vec3f v0(g.v0[0], g.v0[1], g.v0[2]), v1(g.v1[0], g.v1[1], g.v1[2]), v2(g.v2[0], g.v2[1], g.v2[2]);
        copy3(s.face_n, cross(v1 - v0, v2 - v0));
        s.material = f.tri_material[t];
        s.light = f.tri_light[t];
        s.bsdf = s.material >= 0 ? f.materials[s.material].type : SOL_BSDF_NONE;
    }
}

} // namespace sol
