// Wavefront OBJ / MTL import (reference src/formats/obj.cpp), non-strict mode.
//
// Semantics kept from sol:
//   - lines are read into a 1024-byte buffer (obj.cpp:26); a longer line ends the file there, as
//     std::istream::getline failing does (obj.cpp:173);
//   - indices are 1-based with slot 0 a zero dummy, negative indices count back from the elements
//     declared so far, faces with fewer than three indices or a zero vertex index are skipped
//     (obj.cpp:192-215); faces are fanned into triangles (v0, vi, vi+1) (obj.cpp:465-481);
//   - each distinct (vertex, normal, texture) index triple becomes one mesh vertex, numbered in
//     order of first use (obj.cpp:445-455); vertices without a normal get the normalised sum of the
//     unit normals of the triangles that use them (obj.cpp:486-499);
//   - usemtl names unknown to the material libraries fall back to the all-zero "#dummy" material
//     (obj.cpp:414-422); MTL commands before the first newmtl also go to "#dummy", and a repeated
//     newmtl edits the existing entry (obj.cpp:290-299);
//   - materials map to BSDFs by illum: 5 -> MirrorBsdf(Ks), 7 -> GlassBsdf(Ks, Tf, 1 / Ni), else a
//     PhongBsdf(Ks, Ns) if Ks is non-zero, a DiffuseBsdf(Kd) if Kd is non-zero, and their mix with
//     weight lum(Ks) / (lum(Kd) + lum(Ks)) if both (obj.cpp:374-411); a texture map counts as
//     weight 1. A face with a non-zero Ke gets one triangle light per triangle (obj.cpp:469-474).
// Differences, on inputs where sol's behaviour is undefined: an index beyond the elements declared
// so far makes the face invalid (sol reads out of bounds); a lone '-' where an index is expected
// ends the face (sol loops forever); several mtllib files are read in order of first mention (sol
// iterates an unordered_set).
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "sol/scene_loader.h"

namespace sol::obj {

namespace {

constexpr size_t kLineBuffer = 1024;   // bytes, terminating NUL included

bool is_space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }
bool is_digit(char c) { return std::isdigit(static_cast<unsigned char>(c)) != 0; }
const char* skip_spaces(const char* p) { while (is_space(*p)) ++p; return p; }
const char* skip_token(const char* p) { while (*p && !is_space(*p)) ++p; return p; }

// The lines of a file as getline into a fixed buffer delivers them: stops at the first line that
// does not fit. Trailing white space is removed (obj.cpp:93-101).
template <typename F>
void for_each_line(std::istream& is, F&& f) {
    std::stringstream ss;
    ss << is.rdbuf();
    const std::string text = ss.str();
    size_t pos = 0, number = 0;
    std::string line;
    while (pos < text.size()) {
        size_t end = text.find('\n', pos);
        size_t len = (end == std::string::npos ? text.size() : end) - pos;
        if (len > kLineBuffer - 1) return;
        line.assign(text, pos, len);
        ++number;
        pos = end == std::string::npos ? text.size() : end + 1;
        const char* p = skip_spaces(line.c_str());
        if (*p == '\0' || *p == '#') continue;
        size_t n = std::strlen(p);
        while (n > 1 && is_space(p[n - 1])) --n;
        std::string trimmed(p, n);
        f(trimmed.c_str(), number);
    }
}

float read_float(const char*& p) {
    char* end;
    float v = std::strtof(p, &end);
    p = end;
    return v;
}

Vec3f read_vec3(const char* p) {
    float x = read_float(p), y = read_float(p), z = read_float(p);
    return Vec3f(x, y, z);
}

Vec2f read_vec2(const char* p) {
    float x = read_float(p), y = read_float(p);
    return Vec2f(x, y);
}

Color read_color(const char* p) {
    float r = read_float(p), g = read_float(p), b = read_float(p);
    return Color(std::max(r, 0.0f), std::max(g, 0.0f), std::max(b, 0.0f));
}

struct Index {
    int v = 0, n = 0, t = 0;
    bool operator==(const Index& o) const { return v == o.v && n == o.n && t == o.t; }
};

struct IndexHash {
    size_t operator()(const Index& i) const {
        uint64_t h = static_cast<uint32_t>(i.v);
        h = h * 0x9E3779B97F4A7C15ull ^ static_cast<uint32_t>(i.n);
        h = h * 0x9E3779B97F4A7C15ull ^ static_cast<uint32_t>(i.t);
        return static_cast<size_t>(h ^ (h >> 29));
    }
};

struct Face {
    std::vector<Index> indices;
    size_t material;
};

struct File {
    std::vector<Vec3f> vertices{Vec3f(0, 0, 0)};
    std::vector<Vec3f> normals{Vec3f(0, 0, 0)};
    std::vector<Vec2f> tex_coords{Vec2f(0, 0)};
    std::vector<Face> faces;                           // in file order
    std::vector<std::string> materials{"#dummy"};      // usemtl names, in order of first use
    std::vector<std::string> mtl_files;                // mtllib names, in order of first mention
};

// One "v[/[t][/n]]" index group (obj.cpp:111-138). Returns false at the end of the face.
bool read_index(const char*& p, Index& idx) {
    const char* q = skip_spaces(p);
    if (!is_digit(*q) && *q != '-') return false;
    char* end;
    long v = std::strtol(q, &end, 10);
    if (end == q) return false;
    idx = Index{};
    idx.v = static_cast<int>(v);
    q = skip_spaces(end);
    if (*q == '/') {
        ++q;
        if (*q != '/') {
            idx.t = static_cast<int>(std::strtol(q, &end, 10));
            q = end;
        }
        q = skip_spaces(q);
        if (*q == '/') {
            ++q;
            idx.n = static_cast<int>(std::strtol(q, &end, 10));
            q = end;
        }
    }
    p = q;
    return true;
}

// Relative (negative) index -> absolute; out of range -> -1 (invalid).
int resolve(int i, size_t count, bool allow_zero) {
    long long a = i < 0 ? static_cast<long long>(count) + i : i;
    if (a < (allow_zero ? 0 : 1) || a >= static_cast<long long>(count)) return -1;
    return static_cast<int>(a);
}

File parse_obj(const std::string& file_name) {
    std::ifstream is(file_name, std::ios::binary);
    if (!is) throw std::runtime_error("Cannot open OBJ file '" + file_name + "'");
    File file;
    size_t material = 0;
    for_each_line(is, [&](const char* p, size_t) {
        if (p[0] == 'v') {
            if (p[1] == ' ' || p[1] == '\t') file.vertices.push_back(read_vec3(p + 1));
            else if (p[1] == 'n') file.normals.push_back(read_vec3(p + 2));
            else if (p[1] == 't') file.tex_coords.push_back(read_vec2(p + 2));
        } else if (p[0] == 'f' && is_space(p[1])) {
            Face face;
            face.material = material;
            const char* q = p + 2;
            Index idx;
            while (read_index(q, idx)) face.indices.push_back(idx);
            bool valid = face.indices.size() >= 3;
            for (Index& i : face.indices) {
                i.v = resolve(i.v, file.vertices.size(), false);
                i.t = resolve(i.t, file.tex_coords.size(), true);
                i.n = resolve(i.n, file.normals.size(), true);
                valid = valid && i.v > 0 && i.t >= 0 && i.n >= 0;
            }
            if (valid) file.faces.push_back(std::move(face));
        } else if (!std::strncmp(p, "usemtl", 6) && is_space(p[6])) {
            const char* b = skip_spaces(p + 6);
            std::string name(b, skip_token(b));
            size_t i = 0;
            while (i < file.materials.size() && file.materials[i] != name) ++i;
            if (i == file.materials.size()) file.materials.push_back(name);
            material = i;
        } else if (!std::strncmp(p, "mtllib", 6) && is_space(p[6])) {
            const char* b = skip_spaces(p + 6);
            std::string name(b, skip_token(b));
            bool seen = false;
            for (auto& m : file.mtl_files) seen = seen || m == name;
            if (!seen) file.mtl_files.push_back(name);
        }
        // g, o, s and unknown commands: nothing (faces keep file order either way)
    });
    return file;
}

struct Material {
    Color ka, kd, ks, ke, tf;
    float ns = 0.0f, ni = 0.0f, tr = 0.0f, d = 0.0f;
    int illum = 0;
    std::string map_ka, map_kd, map_ks, map_ns, map_ke, map_d, bump;
};

using MaterialLib = std::unordered_map<std::string, Material>;

// parse_mtl (obj.cpp:281-355). A missing file is skipped.
void parse_mtl(const std::string& file_name, MaterialLib& lib) {
    std::ifstream is(file_name, std::ios::binary);
    if (!is) return;
    Material* m = &lib["#dummy"];
    for_each_line(is, [&](const char* p, size_t) {
        const size_t len = std::strlen(p);
        auto rest = [](const char* s) { return std::string(skip_spaces(s)); };
        if (!std::strncmp(p, "newmtl", 6) && is_space(p[6])) {
            const char* b = skip_spaces(p + 7);
            m = &lib[std::string(b, skip_token(b))];
        } else if (len >= 3 && p[0] == 'K' && is_space(p[2])) {
            if (p[1] == 'a') m->ka = read_color(p + 3);
            else if (p[1] == 'd') m->kd = read_color(p + 3);
            else if (p[1] == 's') m->ks = read_color(p + 3);
            else if (p[1] == 'e') m->ke = read_color(p + 3);
        } else if (len >= 3 && p[0] == 'N' && is_space(p[2])) {
            const char* q = p + 3;
            if (p[1] == 's') m->ns = read_float(q);
            else if (p[1] == 'i') m->ni = read_float(q);
        } else if (len >= 3 && p[0] == 'T' && is_space(p[2])) {
            const char* q = p + 3;
            if (p[1] == 'f') m->tf = read_color(p + 3);
            else if (p[1] == 'r') m->tr = read_float(q);
        } else if (p[0] == 'd' && is_space(p[1])) {
            const char* q = p + 2;
            m->d = read_float(q);
        } else if (!std::strncmp(p, "illum", 5) && is_space(p[5])) {
            const char* q = p + 6;
            m->illum = static_cast<int>(read_float(q));
        } else if (!std::strncmp(p, "map_Ka", 6) && is_space(p[6])) {
            m->map_ka = rest(p + 7);
        } else if (!std::strncmp(p, "map_Kd", 6) && is_space(p[6])) {
            m->map_kd = rest(p + 7);
        } else if (!std::strncmp(p, "map_Ks", 6) && is_space(p[6])) {
            m->map_ks = rest(p + 7);
        } else if (!std::strncmp(p, "map_Ke", 6) && is_space(p[6])) {
            m->map_ke = rest(p + 7);
        } else if (!std::strncmp(p, "map_Ns", 6) && is_space(p[6])) {
            m->map_ns = rest(p + 7);
        } else if (!std::strncmp(p, "map_d", 5) && is_space(p[5])) {
            m->map_d = rest(p + 6);
        } else if (!std::strncmp(p, "bump", 4) && is_space(p[4])) {
            m->bump = rest(p + 4);
        }
    });
}

// convert_material (obj.cpp:374-411).
const Bsdf* convert_material(SceneLoader& loader, const Material& m) {
    if (m.illum == 5) return loader.mirror_bsdf(loader.color_texture(m.map_ks, m.ks));
    if (m.illum == 7) {
        return loader.glass_bsdf(loader.color_texture(m.map_ks, m.ks), loader.color_texture("", m.tf),
                                 loader.scalar_texture("", 1.0f / m.ni));
    }
    const Bsdf* spec = nullptr;
    const Bsdf* diff = nullptr;
    float spec_k = 0.0f, diff_k = 0.0f;
    if (!(m.ks == Color::black()) || !m.map_ks.empty()) {
        spec = loader.phong_bsdf(loader.color_texture(m.map_ks, m.ks), loader.scalar_texture(m.map_ns, m.ns));
        spec_k = m.map_ks.empty() ? m.ks.luminance() : 1.0f;
    }
    if (!(m.kd == Color::black()) || !m.map_kd.empty()) {
        diff = loader.diffuse_bsdf(loader.color_texture(m.map_kd, m.kd));
        diff_k = m.map_kd.empty() ? m.kd.luminance() : 1.0f;
    }
    if (spec && diff) return loader.interp_bsdf(diff, spec, loader.scalar_texture("", spec_k / (diff_k + spec_k)));
    return diff ? diff : spec;
}

// build_mesh (obj.cpp:424-510).
std::unique_ptr<TriangleMesh> build_mesh(SceneLoader& loader, const File& file,
                                         const std::vector<const Material*>& face_materials) {
    std::unordered_map<Index, size_t, IndexHash> index_map;
    std::vector<size_t> indices;
    std::vector<Vec3f> vertices, normals;
    std::vector<Vec2f> tex_coords;
    std::vector<const Bsdf*> bsdfs;
    std::vector<size_t> normals_to_fix;
    std::unordered_map<size_t, const Light*> lights;
    std::vector<size_t> ids;
    for (size_t f = 0; f < file.faces.size(); ++f) {
        const Face& face = file.faces[f];
        ids.clear();
        for (const Index& i : face.indices) {
            auto [it, inserted] = index_map.emplace(i, vertices.size());
            if (inserted) {
                if (i.n == 0) normals_to_fix.push_back(normals.size());
                vertices.push_back(file.vertices[i.v]);
                normals.push_back(file.normals[i.n]);
                tex_coords.push_back(file.tex_coords[i.t]);
            }
            ids.push_back(it->second);
        }
        const Material& m = *face_materials[f];
        const Bsdf* bsdf = convert_material(loader, m);
        const bool emissive = !(m.ke == Color::black()) || !m.map_ke.empty();
        for (size_t k = 1; k + 1 < ids.size(); ++k) {
            if (emissive) {
                Triangle tri{vertices[ids[0]], vertices[ids[k]], vertices[ids[k + 1]]};
                lights.emplace(indices.size() / 3, loader.triangle_light(tri, loader.color_texture(m.map_ke, m.ke)));
            }
            indices.push_back(ids[0]);
            indices.push_back(ids[k]);
            indices.push_back(ids[k + 1]);
            bsdfs.push_back(bsdf);
        }
    }
    if (!normals_to_fix.empty()) {
        std::vector<Vec3f> smooth(normals.size(), Vec3f(0, 0, 0));
        for (size_t i = 0; i < indices.size(); i += 3) {
            Vec3f n = Triangle{vertices[indices[i]], vertices[indices[i + 1]], vertices[indices[i + 2]]}.normal();
            smooth[indices[i]] += n;
            smooth[indices[i + 1]] += n;
            smooth[indices[i + 2]] += n;
        }
        for (size_t v : normals_to_fix) normals[v] = normalize(smooth[v]);
    }
    return std::make_unique<TriangleMesh>(std::move(indices), std::move(vertices), std::move(normals),
                                          std::move(tex_coords), std::move(bsdfs), std::move(lights),
                                          loader.bvh_config());
}

} // namespace

std::unique_ptr<TriangleMesh> load(SceneLoader& loader, std::string_view file_name) {
    const std::string name(file_name);
    File file = parse_obj(name);
    MaterialLib lib;
    lib.emplace("#dummy", Material{});
    std::error_code ec;
    const std::string dir = std::filesystem::absolute(name, ec).parent_path().string();
    for (const std::string& mtl : file.mtl_files) parse_mtl(dir + "/" + mtl, lib);
    // check_materials (obj.cpp:414-422): unknown names use the dummy material
    std::vector<const Material*> by_index;
    for (const std::string& m : file.materials) {
        auto it = lib.find(m);
        by_index.push_back(it != lib.end() ? &it->second : &lib.at("#dummy"));
    }
    std::vector<const Material*> face_materials;
    face_materials.reserve(file.faces.size());
    for (const Face& f : file.faces) face_materials.push_back(by_index[f.material]);
    return build_mesh(loader, file, face_materials);
}

} // namespace sol::obj
