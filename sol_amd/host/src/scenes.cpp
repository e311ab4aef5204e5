// Procedural scenes. All random choices come from the project's PCG streams with fixed seeds,
// so a scene is a pure function of its parameters.
#include <cmath>
#include <map>
#include <numbers>
#include <stdexcept>

#include "sol/scenes.h"
#include "sol/samplers.h"
#include "sol/triangle_mesh.h"

namespace sol::scenes {

namespace {

struct MeshBuilder {
    std::vector<size_t> indices;
    std::vector<Vec3f> vertices, normals;
    std::vector<const Bsdf*> bsdfs;
    std::unordered_map<size_t, const Light*> lights;

    size_t tri_count() const { return indices.size() / 3; }

    // Quad p0 p1 p2 p3, counter-clockwise seen from its front; two triangles (p0,p1,p2),
    // (p0,p2,p3) with the flat normal on every vertex.
    void quad(const Vec3f& p0, const Vec3f& p1, const Vec3f& p2, const Vec3f& p3, const Bsdf* bsdf) {
        Vec3f n = normalize(cross(p1 - p0, p2 - p0));
        size_t base = vertices.size();
        for (const Vec3f& p : {p0, p1, p2, p3}) { vertices.push_back(p); normals.push_back(n); }
        for (size_t i : {0, 1, 2, 0, 2, 3}) indices.push_back(base + i);
        bsdfs.push_back(bsdf);
        bsdfs.push_back(bsdf);
    }
};

const ConstantColorTexture* color_tex(Scene& s, const Color& c) {
    auto t = std::make_unique<ConstantColorTexture>(c);
    auto* p = t.get();
    s.textures.push_back(std::move(t));
    return p;
}

const ConstantTexture* scalar_tex(Scene& s, float v) {
    auto t = std::make_unique<ConstantTexture>(v);
    auto* p = t.get();
    s.textures.push_back(std::move(t));
    return p;
}

template <typename B, typename... Args>
B* add_bsdf(Scene& s, Args&&... args) {
    auto b = std::make_unique<B>(std::forward<Args>(args)...);
    auto* p = b.get();
    s.bsdfs.push_back(std::move(b));
    return p;
}

// Adds the two triangles of the last quad as triangle lights of the given emission.
void quad_lights(Scene& s, MeshBuilder& m, const ColorTexture* emission) {
    size_t t1 = m.tri_count() - 1, t0 = t1 - 1;
    for (size_t t : {t0, t1}) {
        Triangle tri{m.vertices[m.indices[t * 3 + 0]], m.vertices[m.indices[t * 3 + 1]], m.vertices[m.indices[t * 3 + 2]]};
        auto l = std::make_unique<UniformTriangleLight>(UniformTriangle(tri), *emission);
        m.lights[t] = l.get();
        s.lights.push_back(std::move(l));
    }
}

// Axis-aligned box of half extents (hx, hz) and height h on the floor, rotated by `angle`
// (radians) about the vertical axis through (cx, cz); six outward-facing quads.
void rotated_box(MeshBuilder& m, float cx, float cz, float hx, float hz, float h, float angle, const Bsdf* bsdf) {
    float c = std::cos(angle), s = std::sin(angle);
    auto P = [&](float sx, float y, float sz) {
        float x = sx * hx, z = sz * hz;
        return Vec3f(cx + x * c - z * s, y, cz + x * s + z * c);
    };
    m.quad(P(-1, h, -1), P(-1, h, 1), P(1, h, 1), P(1, h, -1), bsdf);   // top    (+y)
    m.quad(P(-1, 0, -1), P(1, 0, -1), P(1, 0, 1), P(-1, 0, 1), bsdf);   // bottom (-y)
    m.quad(P(-1, 0, 1), P(1, 0, 1), P(1, h, 1), P(-1, h, 1), bsdf);     // +z
    m.quad(P(-1, 0, -1), P(-1, h, -1), P(1, h, -1), P(1, 0, -1), bsdf); // -z
    m.quad(P(1, 0, -1), P(1, h, -1), P(1, h, 1), P(1, 0, 1), bsdf);     // +x
    m.quad(P(-1, 0, -1), P(-1, 0, 1), P(-1, h, 1), P(-1, h, -1), bsdf); // -x
}

Scene finish(Scene&& s, MeshBuilder&& m, const bvh::BuildConfig& config) {
    std::vector<Vec2f> uvs(m.vertices.size(), Vec2f(0, 0));
    s.root = std::make_unique<TriangleMesh>(std::move(m.indices), std::move(m.vertices), std::move(m.normals),
                                            std::move(uvs), std::move(m.bsdfs), std::move(m.lights), config);
    return std::move(s);
}

} // namespace

Scene cornell_box(float aspect, const bvh::BuildConfig& config) {
    Scene s;
    MeshBuilder m;
    auto* white = add_bsdf<DiffuseBsdf>(s, *color_tex(s, Color(0.73f, 0.73f, 0.73f)));
    auto* red = add_bsdf<DiffuseBsdf>(s, *color_tex(s, Color(0.65f, 0.05f, 0.05f)));
    auto* green = add_bsdf<DiffuseBsdf>(s, *color_tex(s, Color(0.12f, 0.45f, 0.15f)));
    auto* light_bsdf = add_bsdf<DiffuseBsdf>(s, *color_tex(s, Color(0.78f, 0.78f, 0.78f)));
    auto* emission = color_tex(s, Color(17.0f, 12.0f, 4.0f));

    // Room [-1,1] x [0,2] x [-1,1], open towards +z, walls facing inwards.
    m.quad(Vec3f(-1, 0, -1), Vec3f(-1, 0, 1), Vec3f(1, 0, 1), Vec3f(1, 0, -1), white);  // floor
    m.quad(Vec3f(-1, 2, -1), Vec3f(1, 2, -1), Vec3f(1, 2, 1), Vec3f(-1, 2, 1), white);  // ceiling
    m.quad(Vec3f(-1, 0, -1), Vec3f(1, 0, -1), Vec3f(1, 2, -1), Vec3f(-1, 2, -1), white); // back
    m.quad(Vec3f(-1, 0, -1), Vec3f(-1, 2, -1), Vec3f(-1, 2, 1), Vec3f(-1, 0, 1), red);   // left
    m.quad(Vec3f(1, 0, -1), Vec3f(1, 0, 1), Vec3f(1, 2, 1), Vec3f(1, 2, -1), green);     // right
    rotated_box(m, -0.33f, -0.30f, 0.30f, 0.30f, 1.20f, 0.2618f, white);  // tall box, 15 degrees
    rotated_box(m, 0.35f, 0.35f, 0.30f, 0.30f, 0.60f, -0.3142f, white);   // short box, -18 degrees
    // Ceiling light, facing down.
    m.quad(Vec3f(-0.25f, 1.98f, -0.2f), Vec3f(0.25f, 1.98f, -0.2f), Vec3f(0.25f, 1.98f, 0.2f),
           Vec3f(-0.25f, 1.98f, 0.2f), light_bsdf);
    quad_lights(s, m, emission);

    s.camera = std::make_unique<PerspectiveCamera>(Vec3f(0, 0.9f, 2.5f), Vec3f(0, 0, -1), Vec3f(0, 1, 0), 60.0f, aspect);
    return finish(std::move(s), std::move(m), config);
}

void icosphere(uint32_t level, std::vector<Vec3f>& vertices, std::vector<uint32_t>& indices) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    vertices = {
        {-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
        {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
        {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& v : vertices) v = normalize(v);
    indices = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
               1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
               3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
               4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (uint32_t l = 0; l < level; ++l) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            uint32_t id = static_cast<uint32_t>(vertices.size());
            vertices.push_back(normalize((vertices[a] + vertices[b]) * 0.5f));
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> next;
        next.reserve(indices.size() * 4);
        for (size_t f = 0; f < indices.size(); f += 3) {
            uint32_t a = indices[f], b = indices[f + 1], c = indices[f + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            for (uint32_t v : {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca}) next.push_back(v);
        }
        indices.swap(next);
    }
}

Scene sphere_field(const sol_sphere_field_params& p, const bvh::BuildConfig& config) {
    if (p.num_spheres == 0 || p.radius_min <= 0 || p.radius_max < p.radius_min || p.aspect <= 0)
        throw std::invalid_argument("invalid sphere field parameters");
    Scene s;
    MeshBuilder m;
    PcgSampler rng(p.seed);

    auto* ground = add_bsdf<DiffuseBsdf>(s, *color_tex(s, Color(0.5f, 0.5f, 0.5f)));
    float g = p.extent + 10.0f;
    m.quad(Vec3f(-g, 0, -g), Vec3f(-g, 0, g), Vec3f(g, 0, g), Vec3f(g, 0, -g), ground);

    // One material per group (per sphere when num_groups == 0): diffuse, Phong or glass,
    // each with probability 1/3.
    uint32_t groups = p.num_groups ? p.num_groups : p.num_spheres;
    std::vector<const Bsdf*> group_bsdf(groups);
    const auto* one = color_tex(s, Color(1.0f, 1.0f, 1.0f));
    const auto* glass_eta = scalar_tex(s, 1.0f / 1.5f);
    for (uint32_t gi = 0; gi < groups; ++gi) {
        float pick = rng();
        Bsdf* b;
        if (pick < 1.0f / 3.0f) {
            float r = rng.uniform(0.1f, 0.9f), gg = rng.uniform(0.1f, 0.9f), bb = rng.uniform(0.1f, 0.9f);
            b = add_bsdf<DiffuseBsdf>(s, *color_tex(s, Color(r, gg, bb)));
        } else if (pick < 2.0f / 3.0f) {
            float ns = rng.log_uniform(1.0f, 1000.0f);
            float ks = rng.uniform(0.04f, 0.5f);
            b = add_bsdf<PhongBsdf>(s, *color_tex(s, Color(ks, ks, ks)), *scalar_tex(s, ns));
        } else {
            b = add_bsdf<GlassBsdf>(s, *one, *one, *glass_eta);
        }
        if (p.num_groups) b->group = static_cast<int32_t>(gi);
        group_bsdf[gi] = b;
    }

    std::vector<Vec3f> unit_v;
    std::vector<uint32_t> unit_i;
    icosphere(p.subdivision, unit_v, unit_i);
    for (uint32_t i = 0; i < p.num_spheres; ++i) {
        float radius = rng.log_uniform(p.radius_min, p.radius_max);
        float cx = rng.uniform(-p.extent, p.extent);
        float cz = rng.uniform(-p.extent, p.extent);
        Vec3f c(cx, radius, cz);
        size_t base = m.vertices.size();
        for (const Vec3f& v : unit_v) {
            m.vertices.push_back(c + v * radius);
            m.normals.push_back(v);
        }
        for (uint32_t idx : unit_i) m.indices.push_back(base + idx);
        const Bsdf* b = group_bsdf[i % groups];
        for (size_t f = 0; f < unit_i.size() / 3; ++f) m.bsdfs.push_back(b);
    }

    // Three quad lights facing down (no BSDF) and a point light.
    const float lh = 8.0f, half = 1.5f;
    const Vec3f centres[3] = {Vec3f(-6, lh, -4), Vec3f(0, lh, 3), Vec3f(6, lh, -4)};
    const Color emissions[3] = {Color(8.0f, 7.0f, 6.0f), Color(6.0f, 7.0f, 8.0f), Color(7.0f, 8.0f, 6.0f)};
    for (int l = 0; l < 3; ++l) {
        const Vec3f& q = centres[l];
        m.quad(Vec3f(q.x - half, lh, q.z - half), Vec3f(q.x + half, lh, q.z - half),
               Vec3f(q.x + half, lh, q.z + half), Vec3f(q.x - half, lh, q.z + half), nullptr);
        auto* e = color_tex(s, emissions[l]);
        quad_lights(s, m, e);
        if (p.num_groups) {
            s.lights[s.lights.size() - 2]->group = l;
            s.lights[s.lights.size() - 1]->group = l;
        }
    }
    auto pl = std::make_unique<PointLight>(Vec3f(0, 6, 8), Color(60.0f, 60.0f, 60.0f));
    if (p.num_groups) pl->group = 3;
    s.lights.push_back(std::move(pl));

    s.camera = std::make_unique<PerspectiveCamera>(Vec3f(0, 5, 18), Vec3f(0, -4.5f, -18), Vec3f(0, 1, 0), 60.0f, p.aspect);
    return finish(std::move(s), std::move(m), config);
}

Scene triangle_soup(uint32_t num_tris, uint64_t seed, const bvh::BuildConfig& config) {
    if (num_tris == 0) throw std::invalid_argument("soup needs triangles");
    Scene s;
    MeshBuilder m;
    auto* white = add_bsdf<DiffuseBsdf>(s, *color_tex(s, Color(0.5f, 0.5f, 0.5f)));
    PcgSampler rng(seed);
    m.vertices.reserve(size_t(num_tris) * 3);
    m.normals.reserve(size_t(num_tris) * 3);
    m.indices.reserve(size_t(num_tris) * 3);
    m.bsdfs.assign(num_tris, white);
    for (uint32_t t = 0; t < num_tris; ++t) {
        float cx = rng(), cy = rng(), cz = rng();
        Vec3f c(cx, cy, cz);
        Vec3f v[3];
        for (auto& vi : v) {
            float ox = rng.uniform(-0.01f, 0.01f), oy = rng.uniform(-0.01f, 0.01f), oz = rng.uniform(-0.01f, 0.01f);
            vi = c + Vec3f(ox, oy, oz);
        }
        Vec3f n = cross(v[1] - v[0], v[2] - v[0]);
        float len = length(n);
        n = len > 0 ? n * (1.0f / len) : Vec3f(0, 0, 1);
        for (auto& vi : v) {
            m.indices.push_back(m.vertices.size());
            m.vertices.push_back(vi);
            m.normals.push_back(n);
        }
    }
    s.lights.push_back(std::make_unique<PointLight>(Vec3f(0.5f, 2.0f, 0.5f), Color(4.0f, 4.0f, 4.0f)));
    s.camera = std::make_unique<PerspectiveCamera>(Vec3f(0.5f, 0.5f, 3.0f), Vec3f(0, 0, -1), Vec3f(0, 1, 0), 60.0f, 1.0f);
    return finish(std::move(s), std::move(m), config);
}

} // namespace sol::scenes
