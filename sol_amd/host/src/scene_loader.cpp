// TOML scene files (reference src/scene_loader.cpp:30-108) and the shared-object tables of the
// loader (src/scene_loader.h:57-105).
#include "sol/scene_loader.h"

#include <bit>
#include <filesystem>
#include <fstream>
#include <sstream>

#include "sol/toml.h"

namespace sol {

namespace {

uint32_t bits(float f) { return std::bit_cast<uint32_t>(f); }

// parse_vec3 (scene_loader.cpp:15-23): an array's first three elements, each defaulting
// separately when missing or not a number; anything else gives the default.
Vec3f vec3_or(const toml::Value* v, const Vec3f& d) {
    if (!v || !v->is_array()) return d;
    Vec3f r = d;
    for (int i = 0; i < 3; ++i) {
        const toml::Value* e = v->at(static_cast<size_t>(i));
        if (auto x = e ? e->number() : std::nullopt) r[i] = static_cast<float>(*x);
    }
    return r;
}

float float_or(const toml::Value* v, float d) {
    if (auto x = v ? v->number() : std::nullopt) return static_cast<float>(*x);
    return d;
}

std::string string_or(const toml::Value* v, const std::string& d) {
    if (auto s = v ? v->string() : std::nullopt) return *s;
    return d;
}

} // namespace

// Texture maps: the GPU path has constant textures only. A map whose image exists is refused; one
// that cannot be opened falls back to the constant, as sol's get_texture does when Image::load
// fails (obj.cpp:357-372; images resolve against the working directory, scene_loader.cpp:66-78).
void SceneLoader::check_map(const std::string& map) const {
    if (map.empty()) return;
    std::error_code ec;
    std::ifstream probe(std::filesystem::absolute(map, ec));
    if (probe) throw std::invalid_argument("image texture '" + map + "': the GPU path supports constant textures only");
}

const ColorTexture* SceneLoader::color_texture(const std::string& map, const Color& c) {
    check_map(map);
    auto key = std::make_tuple(bits(c.r), bits(c.g), bits(c.b));
    auto it = color_textures_.find(key);
    if (it != color_textures_.end()) return it->second;
    scene_.textures.push_back(std::make_unique<ConstantColorTexture>(c));
    auto* t = static_cast<const ColorTexture*>(scene_.textures.back().get());
    color_textures_.emplace(key, t);
    return t;
}

const Texture* SceneLoader::scalar_texture(const std::string& map, float v) {
    check_map(map);
    auto it = scalar_textures_.find(bits(v));
    if (it != scalar_textures_.end()) return it->second;
    scene_.textures.push_back(std::make_unique<ConstantTexture>(v));
    const Texture* t = scene_.textures.back().get();
    scalar_textures_.emplace(bits(v), t);
    return t;
}

template <typename T, typename... Args>
static const Bsdf* shared_bsdf(Scene& scene, std::map<std::tuple<int, const void*, const void*, const void*>, const Bsdf*>& table,
                               Bsdf::Tag tag, const void* a, const void* b, const void* c, Args&&... args) {
    auto key = std::make_tuple(static_cast<int>(tag), a, b, c);
    auto it = table.find(key);
    if (it != table.end()) return it->second;
    scene.bsdfs.push_back(std::make_unique<T>(std::forward<Args>(args)...));
    const Bsdf* p = scene.bsdfs.back().get();
    table.emplace(key, p);
    return p;
}

const Bsdf* SceneLoader::diffuse_bsdf(const ColorTexture* kd) {
    return shared_bsdf<DiffuseBsdf>(scene_, bsdfs_, Bsdf::Tag::DiffuseBsdf, kd, nullptr, nullptr, *kd);
}

const Bsdf* SceneLoader::phong_bsdf(const ColorTexture* ks, const Texture* ns) {
    return shared_bsdf<PhongBsdf>(scene_, bsdfs_, Bsdf::Tag::PhongBsdf, ks, ns, nullptr, *ks, *ns);
}

const Bsdf* SceneLoader::mirror_bsdf(const ColorTexture* ks) {
    return shared_bsdf<MirrorBsdf>(scene_, bsdfs_, Bsdf::Tag::MirrorBsdf, ks, nullptr, nullptr, *ks);
}

const Bsdf* SceneLoader::glass_bsdf(const ColorTexture* ks, const ColorTexture* kt, const Texture* eta) {
    return shared_bsdf<GlassBsdf>(scene_, bsdfs_, Bsdf::Tag::GlassBsdf, ks, kt, eta, *ks, *kt, *eta);
}

const Bsdf* SceneLoader::interp_bsdf(const Bsdf* a, const Bsdf* b, const Texture* k) {
    return shared_bsdf<InterpBsdf>(scene_, bsdfs_, Bsdf::Tag::InterpBsdf, a, b, k, a, b, *k);
}

const Light* SceneLoader::triangle_light(const Triangle& t, const ColorTexture* intensity) {
    auto key = std::make_tuple(bits(t.v0.x), bits(t.v0.y), bits(t.v0.z), bits(t.v1.x), bits(t.v1.y), bits(t.v1.z),
                               bits(t.v2.x), bits(t.v2.y), bits(t.v2.z), static_cast<const void*>(intensity));
    auto it = lights_.find(key);
    if (it != lights_.end()) return it->second;
    scene_.lights.push_back(std::make_unique<UniformTriangleLight>(UniformTriangle(t), *intensity));
    const Light* l = scene_.lights.back().get();
    lights_.emplace(key, l);
    return l;
}

void SceneLoader::insert_geom(const std::string& name, std::unique_ptr<Geometry>&& geom) {
    if (!geoms_.emplace(name, std::move(geom)).second)
        throw std::runtime_error("Duplicate geometry found with name '" + name + "'");
}

void SceneLoader::load(const std::string& file_name) {
    std::ifstream is(file_name, std::ios::binary);
    if (!is) throw std::runtime_error("Cannot open scene file '" + file_name + "'");
    std::stringstream ss;
    ss << is.rdbuf();
    const std::string text = ss.str();
    std::error_code ec;
    const std::string base_dir = std::filesystem::absolute(file_name, ec).parent_path().string();

    auto doc = toml::parse(text, file_name);
    if (const toml::Value* cam = doc->get("camera"); cam && cam->is_table()) {
        // create_camera (scene_loader.cpp:80-91)
        std::string type = string_or(cam->get("type"), "");
        if (type != "perspective")
            throw SourceError(file_name, cam->line, cam->col, "Unknown camera type '" + type + "'");
        scene_.camera = std::make_unique<PerspectiveCamera>(
            vec3_or(cam->get("eye"), defaults_.eye_pos), vec3_or(cam->get("dir"), defaults_.dir_vector),
            vec3_or(cam->get("up"), defaults_.up_vector), float_or(cam->get("fov"), defaults_.fov),
            float_or(cam->get("aspect"), defaults_.aspect_ratio));
    }
    if (const toml::Value* objects = doc->get("objects"); objects && objects->is_array()) {
        for (const auto& obj : objects->array) {
            if (!obj->is_table()) continue;
            // create_geom (scene_loader.cpp:93-102)
            std::string name = string_or(obj->get("name"), "");
            std::string type = string_or(obj->get("type"), "");
            if (type != "import") throw SourceError(file_name, obj->line, obj->col, "Unknown node type '" + type + "'");
            std::string file = string_or(obj->get("file"), "");
            if (!file.ends_with(".obj"))
                throw SourceError(file_name, obj->line, obj->col, "Unknown file format for '" + file + "'");
            insert_geom(name, obj::load(*this, base_dir + "/" + file));
        }
    }
    std::string root_name = string_or(doc->get("root"), "");
    auto it = geoms_.find(root_name);
    if (it == geoms_.end()) throw std::runtime_error("Root geometry named '" + root_name + "' cannot be found");
    scene_.root = std::move(it->second);
}

std::optional<Scene> load_scene(const std::string& file_name, const SceneDefaults& defaults,
                                const bvh::BuildConfig& bvh_config, std::ostream* err_out) {
    Scene scene;
    try {
        SceneLoader loader(scene, defaults, bvh_config);
        loader.load(file_name);
    } catch (const std::exception& e) {
        if (err_out) *err_out << e.what();
        return std::nullopt;
    }
    return std::optional<Scene>(std::move(scene));
}

} // namespace sol
