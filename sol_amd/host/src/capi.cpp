// C ABI of the host library (include/sol_host.h).
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <optional>
#include <string>
#include <vector>

#include "sol_host.h"
#include "sol/image.h"
#include "sol/scene_loader.h"
#include "sol/scenes.h"
#include "sol/samplers.h"
#include "sol/triangle_mesh.h"

struct sol_scene {
    sol::Scene scene;
    sol::FlatScene flat;
    sol_build_stats stats{};
};

namespace {

thread_local std::string g_error;

// A file that cannot be opened or parsed (reported as SOL_E_LOAD).
struct LoadFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

sol::bvh::BuildConfig to_build_config(const sol_bvh_config* c, uint32_t* top_levels) {
    sol_bvh_config d;
    sol_default_bvh_config(&d);
    if (!c) c = &d;
    sol::bvh::BuildConfig b;
    b.min_leaf_size = c->min_leaf_size;
    b.max_leaf_size = c->max_leaf_size;
    b.traversal_cost = c->traversal_cost;
    b.intersection_cost = c->intersection_cost;
    b.reinsertion_iters = c->reinsertion_iters;
    b.reinsertion_batch = c->reinsertion_batch;
    *top_levels = c->top_levels;
    return b;
}

template <typename F>
int make_scene(const sol_bvh_config* cfg, sol_scene** out, F&& make) {
    if (!out) return fail(SOL_E_INVALID, "null output pointer");
    *out = nullptr;
    try {
        uint32_t top_levels = 0;
        sol::bvh::BuildConfig b = to_build_config(cfg, &top_levels);
        auto* s = new sol_scene{make(b), {}, {}};
        try {
            s->flat = sol::flatten_scene(s->scene, top_levels);
        } catch (...) {
            delete s;
            throw;
        }
        const auto* mesh = static_cast<const sol::TriangleMesh*>(s->scene.root.get());
        s->stats.build_seconds = mesh->build_times().build_seconds;
        s->stats.optimize_seconds = mesh->build_times().optimize_seconds;
        s->stats.flatten_seconds = s->flat.flatten_seconds;
        s->stats.sah_cost_built = mesh->build_times().sah_built;
        s->stats.sah_cost_final = mesh->build_times().sah_final;
        s->stats.depth = mesh->bvh().depth();
        s->stats.num_nodes = static_cast<uint32_t>(mesh->bvh().nodes.size());
        s->stats.num_leaves = mesh->bvh().leaf_count();
        *out = s;
        return SOL_OK;
    } catch (const std::bad_alloc&) {
        return fail(SOL_E_ALLOC, "out of host memory");
    } catch (const LoadFailure& e) {
        return fail(SOL_E_LOAD, e.what());
    } catch (const std::invalid_argument& e) {
        std::string m = e.what();
        return fail(m.find("no light") != std::string::npos ? SOL_E_NO_LIGHTS : SOL_E_INVALID, m);
    } catch (const std::exception& e) {
        return fail(SOL_E_INTERNAL, e.what());
    }
}

} // namespace

extern "C" {

const char* sol_host_last_error(void) { return g_error.c_str(); }

void sol_default_bvh_config(sol_bvh_config* out) {
    if (!out) return;
    out->min_leaf_size = 1;
    out->max_leaf_size = 8;
    out->traversal_cost = 1.0f;
    out->intersection_cost = 1.0f;
    out->reinsertion_iters = 3;
    out->reinsertion_batch = 0.05f;
    out->top_levels = 6;
}

void sol_default_pt_config(sol_pt_config* out) {
    if (!out) return;
    out->max_path_len = 64;
    out->min_rr_path_len = 3;
    out->min_survival_prob = 0.05f;
    out->max_survival_prob = 0.75f;
    out->ray_offset = 1.e-5f;
    out->flags = 0;
}

void sol_default_sphere_field_params(sol_sphere_field_params* out) {
    if (!out) return;
    out->num_spheres = 400;
    out->subdivision = 4;
    out->seed = 2;
    out->num_groups = 0;
    out->aspect = 16.0f / 9.0f;
    out->extent = 10.0f;
    out->radius_min = 0.05f;
    out->radius_max = 0.5f;
}

int sol_scene_cornell_box(float aspect, const sol_bvh_config* cfg, sol_scene** out) {
    if (!(aspect > 0)) return fail(SOL_E_INVALID, "aspect must be positive");
    return make_scene(cfg, out, [&](const sol::bvh::BuildConfig& b) { return sol::scenes::cornell_box(aspect, b); });
}

int sol_scene_sphere_field(const sol_sphere_field_params* p, const sol_bvh_config* cfg, sol_scene** out) {
    if (!p) return fail(SOL_E_INVALID, "null parameters");
    return make_scene(cfg, out, [&](const sol::bvh::BuildConfig& b) { return sol::scenes::sphere_field(*p, b); });
}

int sol_scene_triangle_soup(uint32_t num_tris, uint64_t seed, const sol_bvh_config* cfg, sol_scene** out) {
    return make_scene(cfg, out, [&](const sol::bvh::BuildConfig& b) { return sol::scenes::triangle_soup(num_tris, seed, b); });
}

int sol_image_save(const char* path, int format, const float* pixels, uint32_t width, uint32_t height,
                   uint32_t channels) {
    if (!path || !pixels || !width || !height || !channels) return fail(SOL_E_INVALID, "null or empty image");
    if (format < SOL_IMAGE_AUTO || format > SOL_IMAGE_PFM) return fail(SOL_E_INVALID, "unknown image format");
    if (format == SOL_IMAGE_PFM && channels != 3) return fail(SOL_E_INVALID, "PFM images have three channels");
    try {
        sol::Image img(width, height, channels);
        const size_t plane = size_t(width) * height;
        for (uint32_t c = 0; c < channels; ++c) std::memcpy(img.channel(c).get(), pixels + c * plane, plane * sizeof(float));
        if (!img.save(path, static_cast<sol::Image::Format>(format)))
            return fail(SOL_E_LOAD, std::string("cannot save image '") + path + "' in this format");
        return SOL_OK;
    } catch (const std::bad_alloc&) {
        return fail(SOL_E_ALLOC, "out of host memory");
    }
}

void sol_default_scene_defaults(sol_scene_defaults* out) {
    if (!out) return;
    sol::SceneDefaults d;
    out->fov = d.fov;
    out->aspect_ratio = d.aspect_ratio;
    for (int i = 0; i < 3; ++i) {
        out->eye[i] = d.eye_pos[i];
        out->dir[i] = d.dir_vector[i];
        out->up[i] = d.up_vector[i];
    }
}

int sol_scene_load(const char* file_name, const sol_scene_defaults* defaults, const sol_bvh_config* cfg,
                   sol_scene** out) {
    if (!file_name) return fail(SOL_E_INVALID, "null file name");
    sol_scene_defaults dd;
    sol_default_scene_defaults(&dd);
    if (!defaults) defaults = &dd;
    return make_scene(cfg, out, [&](const sol::bvh::BuildConfig& b) {
        sol::SceneDefaults d;
        d.fov = defaults->fov;
        d.aspect_ratio = defaults->aspect_ratio;
        d.eye_pos = sol::Vec3f(defaults->eye[0], defaults->eye[1], defaults->eye[2]);
        d.dir_vector = sol::Vec3f(defaults->dir[0], defaults->dir[1], defaults->dir[2]);
        d.up_vector = sol::Vec3f(defaults->up[0], defaults->up[1], defaults->up[2]);
        sol::Scene s;
        try {
            sol::SceneLoader loader(s, d, b);
            loader.load(file_name);
        } catch (const std::runtime_error& e) {
            throw LoadFailure(e.what());
        }
        return s;
    });
}

int sol_scene_load_obj(const char* file_name, const sol_camera_params* camera, const sol_bvh_config* cfg,
                       sol_scene** out) {
    if (!file_name || !camera) return fail(SOL_E_INVALID, "null argument");
    return make_scene(cfg, out, [&](const sol::bvh::BuildConfig& b) {
        sol::Scene s;
        try {
            sol::SceneLoader loader(s, sol::SceneDefaults{}, b);
            s.root = sol::obj::load(loader, file_name);
        } catch (const std::runtime_error& e) {
            throw LoadFailure(e.what());
        }
        const sol_camera_params& c = *camera;
        s.camera = std::make_unique<sol::PerspectiveCamera>(sol::Vec3f(c.eye[0], c.eye[1], c.eye[2]),
                                                            sol::Vec3f(c.dir[0], c.dir[1], c.dir[2]),
                                                            sol::Vec3f(c.up[0], c.up[1], c.up[2]), c.fov, c.aspect);
        return s;
    });
}

int sol_scene_from_arrays(const sol_mesh_view* mv, const sol_bvh_config* cfg, sol_scene** out) {
    if (!mv || !mv->positions || !mv->normals || !mv->indices || !mv->tri_material || !mv->tri_light ||
        (mv->num_materials && !mv->materials) || (mv->num_lights && !mv->lights))
        return fail(SOL_E_INVALID, "null array in mesh view");
    if (mv->num_tris == 0 || mv->num_vertices == 0) return fail(SOL_E_INVALID, "empty mesh");
    return make_scene(cfg, out, [&](const sol::bvh::BuildConfig& b) {
        sol::Scene s;
        // Materials and lights become constant-texture BSDFs and lights of the scene.
        auto color = [&](const float* c) {
            s.textures.push_back(std::make_unique<sol::ConstantColorTexture>(sol::Color(c[0], c[1], c[2])));
            return static_cast<const sol::ColorTexture*>(s.textures.back().get());
        };
        auto scalar = [&](float v) {
            s.textures.push_back(std::make_unique<sol::ConstantTexture>(v));
            return s.textures.back().get();
        };
        std::vector<std::unique_ptr<sol::Bsdf>> parts;
        for (uint32_t i = 0; i < mv->num_materials; ++i) {
            const sol_material& m = mv->materials[i];
            std::unique_ptr<sol::Bsdf> bsdf;
            switch (m.type) {
                case SOL_BSDF_DIFFUSE: bsdf = std::make_unique<sol::DiffuseBsdf>(*color(m.kd)); break;
                case SOL_BSDF_PHONG: bsdf = std::make_unique<sol::PhongBsdf>(*color(m.ks), *scalar(m.ns)); break;
                case SOL_BSDF_GLASS: bsdf = std::make_unique<sol::GlassBsdf>(*color(m.ks), *color(m.kt), *scalar(m.eta)); break;
                case SOL_BSDF_MIRROR: bsdf = std::make_unique<sol::MirrorBsdf>(*color(m.ks)); break;
                case SOL_BSDF_INTERP: {
                    // the two mixed BSDFs join the table after the caller's materials
                    parts.push_back(std::make_unique<sol::DiffuseBsdf>(*color(m.kd)));
                    parts.push_back(std::make_unique<sol::PhongBsdf>(*color(m.ks), *scalar(m.ns)));
                    bsdf = std::make_unique<sol::InterpBsdf>(parts[parts.size() - 2].get(), parts.back().get(),
                                                             *scalar(m.mix));
                    break;
                }
                default: throw std::invalid_argument("unknown material type");
            }
            bsdf->group = m.group;
            s.bsdfs.push_back(std::move(bsdf));
        }
        for (auto& p : parts) s.bsdfs.push_back(std::move(p));
        for (uint32_t i = 0; i < mv->num_lights; ++i) {
            const sol_light& l = mv->lights[i];
            std::unique_ptr<sol::Light> light;
            if (l.type == SOL_LIGHT_POINT) {
                light = std::make_unique<sol::PointLight>(sol::Vec3f(l.v0[0], l.v0[1], l.v0[2]),
                                                          sol::Color(l.intensity[0], l.intensity[1], l.intensity[2]));
            } else if (l.type == SOL_LIGHT_TRIANGLE) {
                sol::Triangle t{sol::Vec3f(l.v0[0], l.v0[1], l.v0[2]), sol::Vec3f(l.v1[0], l.v1[1], l.v1[2]),
                                sol::Vec3f(l.v2[0], l.v2[1], l.v2[2])};
                light = std::make_unique<sol::UniformTriangleLight>(sol::UniformTriangle(t), *color(l.intensity));
            } else {
                throw std::invalid_argument("unknown light type");
            }
            light->group = l.group;
            s.lights.push_back(std::move(light));
        }
        std::vector<size_t> indices(mv->indices, mv->indices + size_t(mv->num_tris) * 3);
        std::vector<sol::Vec3f> verts(mv->num_vertices), norms(mv->num_vertices);
        for (uint32_t i = 0; i < mv->num_vertices; ++i) {
            verts[i] = sol::Vec3f(mv->positions[i * 3], mv->positions[i * 3 + 1], mv->positions[i * 3 + 2]);
            norms[i] = sol::Vec3f(mv->normals[i * 3], mv->normals[i * 3 + 1], mv->normals[i * 3 + 2]);
        }
        std::vector<const sol::Bsdf*> bsdfs(mv->num_tris, nullptr);
        std::unordered_map<size_t, const sol::Light*> lights;
        for (uint32_t t = 0; t < mv->num_tris; ++t) {
            int32_t m = mv->tri_material[t], l = mv->tri_light[t];
            if (m >= static_cast<int32_t>(mv->num_materials) || l >= static_cast<int32_t>(mv->num_lights))
                throw std::invalid_argument("material or light index out of range");
            if (m >= 0) bsdfs[t] = s.bsdfs[m].get();
            if (l >= 0) lights[t] = s.lights[l].get();
        }
        std::vector<sol::Vec2f> uvs(mv->num_vertices, sol::Vec2f(0, 0));
        s.root = std::make_unique<sol::TriangleMesh>(std::move(indices), std::move(verts), std::move(norms),
                                                     std::move(uvs), std::move(bsdfs), std::move(lights), b);
        const sol_camera_params& c = mv->camera;
        s.camera = std::make_unique<sol::PerspectiveCamera>(sol::Vec3f(c.eye[0], c.eye[1], c.eye[2]),
                                                            sol::Vec3f(c.dir[0], c.dir[1], c.dir[2]),
                                                            sol::Vec3f(c.up[0], c.up[1], c.up[2]), c.fov, c.aspect);
        return s;
    });
}

void sol_scene_destroy(sol_scene* scene) { delete scene; }

int sol_scene_attach_gpu(sol_scene* s, solhip_ctx* ctx) {
    if (!s || !ctx) return fail(SOL_E_INVALID, "null argument");
    auto* mesh = dynamic_cast<sol::TriangleMesh*>(s->scene.root.get());
    if (!mesh) return fail(SOL_E_INVALID, "scene root is not a TriangleMesh");
    try {
        mesh->attach_gpu(ctx);
    } catch (const std::exception& e) {
        return fail(SOL_E_INTERNAL, e.what());
    }
    return SOL_OK;
}

namespace {

sol::Rayf to_rayf(const sol_ray& r) {
    return sol::Rayf(sol::Vec3f(r.org[0], r.org[1], r.org[2]), sol::Vec3f(r.dir[0], r.dir[1], r.dir[2]), r.tmin, r.tmax);
}

void put3(float* d, const sol::Vec3f& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

// The hit record with the BSDF and light as indices of the scene's tables (Scene::bsdfs and
// Scene::lights order, the order of the flat material and light tables).
sol_surface_hit to_surface_hit(const sol::Scene& scene, const sol::Rayf& ray, const std::optional<sol::Hit>& hit) {
    sol_surface_hit o{};
    o.material = -1;
    o.light = -1;
    o.t = ray.tmax;
    if (!hit) return o;
    const sol::SurfaceInfo& si = hit->surf_info;
    o.hit = 1;
    o.front = si.is_front_side ? 1 : 0;
    for (size_t i = 0; i < scene.bsdfs.size(); ++i)
        if (scene.bsdfs[i].get() == hit->bsdf) o.material = static_cast<int32_t>(i);
    for (size_t i = 0; i < scene.lights.size(); ++i)
        if (scene.lights[i].get() == hit->light) o.light = static_cast<int32_t>(i);
    put3(o.point, si.point);
    put3(o.face_normal, si.face_normal);
    put3(o.tangent, si.local.col(0));
    put3(o.bitangent, si.local.col(1));
    put3(o.normal, si.local.col(2));
    o.surf_coords[0] = si.surf_coords.x;
    o.surf_coords[1] = si.surf_coords.y;
    o.tex_coords[0] = si.tex_coords.x;
    o.tex_coords[1] = si.tex_coords.y;
    return o;
}

} // namespace

int sol_scene_intersect_closest(const sol_scene* s, const sol_ray* rays, uint64_t n, int batched, sol_surface_hit* out) {
    if (!s || (n && (!rays || !out))) return fail(SOL_E_INVALID, "null argument");
    const sol::Geometry* geom = s->scene.root.get();
    if (!geom) return fail(SOL_E_INVALID, "scene without root");
    try {
        std::vector<sol::Rayf> rs(n);
        for (uint64_t i = 0; i < n; ++i) rs[i] = to_rayf(rays[i]);
        std::vector<std::optional<sol::Hit>> hits(n);
        if (batched) {
            const auto* mesh = dynamic_cast<const sol::TriangleMesh*>(geom);
            if (!mesh) return fail(SOL_E_INVALID, "scene root is not a TriangleMesh");
            mesh->intersect_closest(rs.data(), n, hits.data());
        } else {
            for (uint64_t i = 0; i < n; ++i) hits[i] = geom->intersect_closest(rs[i]);   // virtual, one ray
        }
        for (uint64_t i = 0; i < n; ++i) out[i] = to_surface_hit(s->scene, rs[i], hits[i]);
    } catch (const std::exception& e) {
        return fail(SOL_E_INTERNAL, e.what());
    }
    return SOL_OK;
}

int sol_scene_intersect_any(const sol_scene* s, const sol_ray* rays, uint64_t n, int batched, int32_t* out) {
    if (!s || (n && (!rays || !out))) return fail(SOL_E_INVALID, "null argument");
    const sol::Geometry* geom = s->scene.root.get();
    if (!geom) return fail(SOL_E_INVALID, "scene without root");
    try {
        std::vector<sol::Rayf> rs(n);
        for (uint64_t i = 0; i < n; ++i) rs[i] = to_rayf(rays[i]);
        if (batched) {
            const auto* mesh = dynamic_cast<const sol::TriangleMesh*>(geom);
            if (!mesh) return fail(SOL_E_INVALID, "scene root is not a TriangleMesh");
            std::unique_ptr<bool[]> occ(new bool[n ? n : 1]);
            mesh->intersect_any(rs.data(), n, occ.get());
            for (uint64_t i = 0; i < n; ++i) out[i] = occ[i] ? 1 : 0;
        } else {
            for (uint64_t i = 0; i < n; ++i) out[i] = geom->intersect_any(rs[i]) ? 1 : 0;   // virtual, one ray
        }
    } catch (const std::exception& e) {
        return fail(SOL_E_INTERNAL, e.what());
    }
    return SOL_OK;
}

int sol_scene_mesh(const sol_scene* s, sol_mesh_view* out) {
    if (!s || !out) return fail(SOL_E_INVALID, "null argument");
    const sol::FlatScene& f = s->flat;
    out->num_vertices = static_cast<uint32_t>(f.positions.size() / 3);
    out->positions = f.positions.data();
    out->normals = f.normals.data();
    out->num_tris = static_cast<uint32_t>(f.indices.size() / 3);
    out->indices = f.indices.data();
    out->tri_material = f.tri_material.data();
    out->tri_light = f.tri_light.data();
    out->num_materials = static_cast<uint32_t>(f.materials.size());
    out->materials = f.materials.data();
    out->num_lights = static_cast<uint32_t>(f.lights.size());
    out->lights = f.lights.data();
    out->camera = f.camera_params;
    return SOL_OK;
}

int sol_scene_bvh(const sol_scene* s, sol_bvh_view* out) {
    if (!s || !out) return fail(SOL_E_INVALID, "null argument");
    out->num_nodes = static_cast<uint32_t>(s->flat.bvh_nodes.size());
    out->nodes = s->flat.bvh_nodes.data();
    out->num_prims = static_cast<uint32_t>(s->flat.prim_indices.size());
    out->prim_indices = s->flat.prim_indices.data();
    return SOL_OK;
}

int sol_scene_gpu_desc(const sol_scene* s, solhip_scene_desc* out) {
    if (!s || !out) return fail(SOL_E_INVALID, "null argument");
    *out = s->flat.gpu_desc();
    return SOL_OK;
}

int sol_scene_build_stats(const sol_scene* s, sol_build_stats* out) {
    if (!s || !out) return fail(SOL_E_INVALID, "null argument");
    *out = s->stats;
    return SOL_OK;
}

int sol_scene_set_materials(sol_scene* s, const sol_material* m, uint32_t n) {
    if (!s || !m || n != s->flat.materials.size()) return fail(SOL_E_INVALID, "material count mismatch");
    std::memcpy(s->flat.materials.data(), m, sizeof(sol_material) * n);
    for (sol_gpu_shade& r : s->flat.gpu_shade) r.bsdf = r.material >= 0 ? m[r.material].type : SOL_BSDF_NONE;
    return SOL_OK;
}

int sol_scene_set_lights(sol_scene* s, const sol_light* l, uint32_t n) {
    if (!s || !l || n != s->flat.lights.size()) return fail(SOL_E_INVALID, "light count mismatch");
    std::memcpy(s->flat.lights.data(), l, sizeof(sol_light) * n);
    return SOL_OK;
}

void sol_camera_from_params(const sol_camera_params* p, sol_camera* out) {
    if (p && out) *out = sol::camera_basis(*p);
}

int sol_gen_rays(uint32_t kind, uint64_t num_rays, uint64_t seed, float offset, sol_ray* out) {
    if (!out || kind > 1) return fail(SOL_E_INVALID, "invalid ray batch request");
    sol::PcgSampler org_rng(seed), dir_rng(seed + 1), tgt_rng(seed + 2);
    const float two_pi = 6.283185307179586f;
    for (uint64_t i = 0; i < num_rays; ++i) {
        sol_ray& r = out[i];
        float ox = org_rng.uniform(-0.5f, 1.5f), oy = org_rng.uniform(-0.5f, 1.5f), oz = org_rng.uniform(-0.5f, 1.5f);
        r.org[0] = ox; r.org[1] = oy; r.org[2] = oz;
        if (kind == 0) {
            float u1 = dir_rng(), u2 = dir_rng();
            float z = 1.0f - 2.0f * u1;
            float rr = std::sqrt(std::max(0.0f, 1.0f - z * z));
            float phi = two_pi * u2;
            r.dir[0] = rr * std::cos(phi);
            r.dir[1] = rr * std::sin(phi);
            r.dir[2] = z;
            r.tmin = 0.0f;
            r.tmax = INFINITY;
        } else {
            float tx = tgt_rng(), ty = tgt_rng(), tz = tgt_rng();
            r.dir[0] = tx - ox; r.dir[1] = ty - oy; r.dir[2] = tz - oz;
            r.tmin = offset;
            r.tmax = 1.0f - offset;
        }
    }
    return SOL_OK;
}

} // extern "C"
