// Scene files, as sol's SceneLoader (reference src/scene_loader.h, src/scene_loader.cpp:30-115)
// and OBJ import (src/formats/obj.cpp). A scene file is TOML: a camera table, a list of objects
// (OBJ imports) and the name of the root object. Objects, textures, BSDFs and lights that compare
// equal are created once and shared, as sol's get_or_insert does.
//
// Only constant textures reach the GPU path: a texture map (map_Kd, map_Ks, map_Ns, map_Ke) whose
// image file exists is rejected with an error; a map whose image cannot be opened falls back to the
// constant colour, as sol does when Image::load fails.
#ifndef SOL_SCENE_LOADER_H
#define SOL_SCENE_LOADER_H

#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <ostream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <tuple>
#include <unordered_map>

#include "sol/bvh.h"
#include "sol/scene.h"
#include "sol/triangle_mesh.h"

namespace sol {

/// Values that a scene file may leave out (reference include/sol/scene.h:23-29).
struct SceneDefaults {
    float fov = 60.0f;
    float aspect_ratio = 1.0f;
    Vec3f eye_pos = Vec3f(0, 0, 0);
    Vec3f dir_vector = Vec3f(0, 0, 1);
    Vec3f up_vector = Vec3f(0, 1, 0);
};

/// An error in an input file, with its position: "file:line:column: message".
class SourceError : public std::runtime_error {
public:
    SourceError(const std::string& file, size_t line, size_t col, const std::string& msg)
        : std::runtime_error(file + ":" + std::to_string(line) + ":" + std::to_string(col) + ": " + msg) {}
};

class SceneLoader {
public:
    SceneLoader(Scene& scene, const SceneDefaults& defaults, const bvh::BuildConfig& bvh_config)
        : scene_(scene), defaults_(defaults), bvh_config_(bvh_config) {}

    /// Loads a TOML scene file into the scene; throws on any error.
    void load(const std::string& file_name);

    const SceneDefaults& defaults() const { return defaults_; }
    const bvh::BuildConfig& bvh_config() const { return bvh_config_; }

    // Shared scene objects: equal arguments give the same object.
    const ColorTexture* color_texture(const std::string& map, const Color& constant);
    const Texture* scalar_texture(const std::string& map, float constant);
    const Bsdf* diffuse_bsdf(const ColorTexture* kd);
    const Bsdf* phong_bsdf(const ColorTexture* ks, const Texture* ns);
    const Bsdf* mirror_bsdf(const ColorTexture* ks);
    const Bsdf* glass_bsdf(const ColorTexture* ks, const ColorTexture* kt, const Texture* eta);
    const Bsdf* interp_bsdf(const Bsdf* a, const Bsdf* b, const Texture* k);
    const Light* triangle_light(const Triangle& triangle, const ColorTexture* intensity);

    void insert_geom(const std::string& name, std::unique_ptr<Geometry>&& geom);

private:
    void check_map(const std::string& map) const;

    Scene& scene_;
    SceneDefaults defaults_;
    bvh::BuildConfig bvh_config_;
    std::map<std::tuple<float, float, float>, const ColorTexture*> color_textures_;
    std::map<float, const Texture*> scalar_textures_;
    std::map<std::tuple<int, const void*, const void*, const void*>, const Bsdf*> bsdfs_;
    std::map<std::tuple<float, float, float, float, float, float, float, float, float, const void*>, const Light*> lights_;
    std::unordered_map<std::string, std::unique_ptr<Geometry>> geoms_;
};

/// Scene::load (reference src/scene_loader.cpp:104-108): the scene, or nothing with the error
/// message written to err_out.
std::optional<Scene> load_scene(const std::string& file_name, const SceneDefaults& defaults,
                                const bvh::BuildConfig& bvh_config, std::ostream* err_out = nullptr);

namespace obj {

/// obj::load (reference src/formats/obj.cpp:512-531): the triangle mesh of an OBJ file and its
/// material libraries, in sol's non-strict mode (unknown commands, invalid faces and missing
/// material files are skipped). Throws if the file cannot be opened.
std::unique_ptr<TriangleMesh> load(SceneLoader& loader, std::string_view file_name);

} // namespace obj

} // namespace sol

#endif
