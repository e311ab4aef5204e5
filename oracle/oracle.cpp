// CPU oracle of the path-tracing hot path. TEST INFRASTRUCTURE ONLY: the checker that tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline use. The product never links it.
//
// A plain C++ restatement of sol, written from its sources (paths relative to the sol tree)
// and, where sol calls into its un-vendored dependencies, from the mathematics:
//
//   function                          follows                                   pinned by (tests/)
//   PcgGenerator / uniform float       include/sol/samplers.h:20-65, libstdc++   published pcg32 demo vector;
//                                      generate_canonical<float,24>              KATs; float edge case
//   pixel_seed (FNV-1a 32, 8 LE bytes) include/sol/renderer.h:44-46 (proto::fnv   FNV-1a published test
//                                      not vendored: FNV-1a assumed)             vectors
//   sample_pixel / camera ray          renderer.h:50-54, cameras.cpp:5-25        float64 closed forms: centre,
//                                                                               corners, fov and aspect angles,
//                                                                               footprint, jitter draw order,
//                                                                               an emitter's projection
//   watertight ray/triangle            JCGT 2013 (Woop et al.); replaces          hand-computed cases,
//                                      proto::PrecomputedTriangle::intersect     watertightness, brute force
//   BVH traversal (closest / any)      triangle_mesh.cpp:64-117                  brute force over all triangles
//   Diffuse / Phong / Glass / Mirror   bsdfs.cpp:11-165, bsdfs.h:89-94           closed forms: pdf integrals,
//                                                                               Fresnel at normal incidence,
//                                                                               TIR, mirror reflection law
//                                                                               and rejection below the face
//   Interp (diffuse + Phong mix)       bsdfs.cpp:179-230, color.h:66-71           mix 0 / 1 = the single lobe,
//                                      (proto::lerp not vendored:                 pdf integral, lobe-pick
//                                      a (1 - t) + b t assumed)                   frequency
//   triangle / point lights            lights.cpp:11-111, shapes.cpp:6-29        uniform-density, pdf = 1/A
//   trace_path / render                path_tracer.cpp:9-120                     furnace identity, MIS
//                                                                               consistency (NEE-only =
//                                                                               BSDF-only = MIS), closed-form
//                                                                               direct lighting
//
// Floating point: fp32 with -ffp-contract=off, correctly rounded / and sqrt; cos, sin and pow
// are evaluated in fp64 and rounded to fp32 once. Operation order is written out as sol's
// expressions read (left to right).
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include <omp.h>

namespace {

constexpr float kInvPi = 0.318309886183790671538f;
constexpr double kTwoPi = 6.283185307179586476925;
constexpr float kFarScale = 1.0f + 3.0f * 0x1p-23f;
// Boxes are culled against [0, tmax * (1 + 8 ulps)]: the triangle test's t has its own rounding
// (large near tmin, for rays leaving a surface), and a box must not be skipped because its slab
// distance rounded the other way. Triangles are still accepted only for tmin < t < tmax.
constexpr float kHighScale = 1.0f + 0x1p-20f;

struct Vec3 {
    float x, y, z;
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(Vec3 a, Vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator-(Vec3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Vec3 normalize(Vec3 a) { return a * (1.0f / length(a)); }
inline float get(Vec3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline Vec3 vec(const float* p) { return {p[0], p[1], p[2]}; }
inline Vec3 reflect(Vec3 d, Vec3 n) { return d - n * (2.0f * dot(d, n)); }
inline Vec3 lerp(Vec3 a, Vec3 b, Vec3 c, float u, float v) { return a * ((1.0f - u) - v) + b * u + c * v; }
inline float positive_dot(Vec3 a, Vec3 b) { return std::max(dot(a, b), 0.0f); }
inline float lerp(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline Vec3 lerp(Vec3 a, Vec3 b, float t) { return a * (1.0f - t) + b * t; }
inline float negative_dot(Vec3 a, Vec3 b) { return std::min(dot(a, b), 0.0f); }

using Color = Vec3;
inline float luminance(Color c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

// ---- PCG + float (samplers.h:20-65) ------------------------------------------------------------

struct Pcg {
    uint64_t state = 0;
    explicit Pcg(uint64_t seed) {
        state = 0;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * UINT64_C(6364136223846793005) + 1u;
        uint32_t xorshifted = static_cast<uint32_t>(((old >> 18) ^ old) >> 27);
        uint32_t rot = static_cast<uint32_t>(old >> 59);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
    }
    // generate_canonical<float, 24> over one 32-bit draw, clamped below 1.
    float operator()() {
        float u = static_cast<float>(next()) * 0x1p-32f;
        if (u >= 1.0f) u = std::nextafter(1.0f, 0.0f);
        return u;
    }
};

uint32_t fnv_bytes(uint32_t h, uint64_t v) {
    for (int i = 0; i < 8; ++i) {
        h ^= static_cast<uint32_t>(v >> (8 * i)) & 0xffu;
        h *= 16777619u;
    }
    return h;
}

uint32_t pixel_seed(uint64_t frame, uint64_t x, uint64_t y) {
    return fnv_bytes(fnv_bytes(fnv_bytes(2166136261u, x), y), frame);
}

// ---- sampling (proto::sample_* restated) --------------------------------------------------------

float cos2pi(float u) { return static_cast<float>(std::cos(kTwoPi * static_cast<double>(u))); }
float sin2pi(float u) { return static_cast<float>(std::sin(kTwoPi * static_cast<double>(u))); }
float fpow(float x, float y) { return static_cast<float>(std::pow(static_cast<double>(x), static_cast<double>(y))); }

Vec3 sample_cosine_hemisphere(float u, float v, float& pdf) {
    float cos_theta = std::sqrt(1.0f - v);
    float sin_theta = std::sqrt(v);
    pdf = cos_theta * kInvPi;
    return {cos2pi(u) * sin_theta, sin2pi(u) * sin_theta, cos_theta};
}
float cosine_hemisphere_pdf(float c) { return c * kInvPi; }
float cosine_power_hemisphere_pdf(float n, float c) { return (n + 1.0f) * (0.5f * kInvPi) * fpow(c, n); }
Vec3 sample_cosine_power_hemisphere(float n, float u, float v, float& pdf) {
    float cos_theta = fpow(v, 1.0f / (n + 1.0f));
    float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
    pdf = cosine_power_hemisphere_pdf(n, cos_theta);
    return {cos2pi(u) * sin_theta, sin2pi(u) * sin_theta, cos_theta};
}

// Orthonormal basis, Duff et al. 2017 (columns t, b, n).
struct Mat3 {
    Vec3 c0, c1, c2;
    Vec3 operator*(Vec3 v) const { return c0 * v.x + c1 * v.y + c2 * v.z; }
};
Mat3 ortho_basis(Vec3 n) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    return Mat3{{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x}, {b, sign + n.y * n.y * a, -n.y}, n};
}

// ---- geometry -------------------------------------------------------------------------------------

struct Ray {
    Vec3 org, dir;
    float tmin, tmax;
};

struct RayData {
    Vec3 org, dir, inv;
    int kx, ky, kz;
    float sx, sy, sz;
};

float nonzero(float d) { return std::fabs(d) < 1e-20f ? std::copysign(1e-20f, d) : d; }

RayData prepare(const Ray& r) {
    RayData d;
    d.org = r.org;
    d.dir = r.dir;
    d.inv = {1.0f / nonzero(r.dir.x), 1.0f / nonzero(r.dir.y), 1.0f / nonzero(r.dir.z)};
    float a[3] = {std::fabs(r.dir.x), std::fabs(r.dir.y), std::fabs(r.dir.z)};
    int kz = 0;
    if (a[1] > a[kz]) kz = 1;
    if (a[2] > a[kz]) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (get(r.dir, kz) < 0.0f) std::swap(kx, ky);
    d.kx = kx;
    d.ky = ky;
    d.kz = kz;
    d.sx = get(r.dir, kx) / get(r.dir, kz);
    d.sy = get(r.dir, ky) / get(r.dir, kz);
    d.sz = 1.0f / get(r.dir, kz);
    return d;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), with the fp64 fallback on
// edges. u weighs v1 and v weighs v2.
bool ray_triangle(const RayData& r, Vec3 p0, Vec3 p1, Vec3 p2, float& t, float& u, float& v) {
    Vec3 A = p0 - r.org, B = p1 - r.org, C = p2 - r.org;
    float Ax = get(A, r.kx) - r.sx * get(A, r.kz);
    float Ay = get(A, r.ky) - r.sy * get(A, r.kz);
    float Bx = get(B, r.kx) - r.sx * get(B, r.kz);
    float By = get(B, r.ky) - r.sy * get(B, r.kz);
    float Cx = get(C, r.kx) - r.sx * get(C, r.kz);
    float Cy = get(C, r.ky) - r.sy * get(C, r.kz);
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = static_cast<float>(static_cast<double>(Cx) * static_cast<double>(By) - static_cast<double>(Cy) * static_cast<double>(Bx));
        V = static_cast<float>(static_cast<double>(Ax) * static_cast<double>(Cy) - static_cast<double>(Ay) * static_cast<double>(Cx));
        W = static_cast<float>(static_cast<double>(Bx) * static_cast<double>(Ay) - static_cast<double>(By) * static_cast<double>(Ax));
    }
    bool any_neg = U < 0.0f || V < 0.0f || W < 0.0f;
    bool any_pos = U > 0.0f || V > 0.0f || W > 0.0f;
    if (any_neg && any_pos) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.sz * get(A, r.kz), Bz = r.sz * get(B, r.kz), Cz = r.sz * get(C, r.kz);
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    t = T * rcp;
    u = V * rcp;
    v = W * rcp;
    return true;
}

// Conservative slab test (Ize 2013 far-distance scaling).
bool ray_box(const RayData& r, const float* b, float /*tmin*/, float tmax, float& entry) {
    float tx0 = (b[0] - r.org.x) * r.inv.x, tx1 = (b[1] - r.org.x) * r.inv.x;
    float ty0 = (b[2] - r.org.y) * r.inv.y, ty1 = (b[3] - r.org.y) * r.inv.y;
    float tz0 = (b[4] - r.org.z) * r.inv.z, tz1 = (b[5] - r.org.z) * r.inv.z;
    float near = std::max(std::max(std::min(tx0, tx1), std::min(ty0, ty1)), std::max(std::min(tz0, tz1), 0.0f));
    float far = std::min(std::min(std::max(tx0, tx1), std::max(ty0, ty1)), std::max(tz0, tz1)) * kFarScale;
    far = std::min(far, tmax * kHighScale);
    entry = near;
    return near <= far;
}

struct Node {
    float b[6];
    uint32_t index, count;
};

struct ClosestHit {
    int32_t tri = -1;
    float t = 0, u = 0, v = 0;
};

struct Counters {
    uint64_t inner = 0, tris = 0;
};

} // namespace

struct oracle_scene {
    std::vector<Vec3> pos, nrm;
    std::vector<uint32_t> idx;
    std::vector<int32_t> tri_mat, tri_light;
    std::vector<sol_material> materials;
    std::vector<sol_light> lights;
    sol_camera_params cam_params{};
    Vec3 eye{}, cam_dir{}, cam_right{}, cam_up{};
    std::vector<Node> nodes;
    std::vector<uint32_t> prims;

    Vec3 vert(uint32_t t, int k) const { return pos[idx[t * 3 + k]]; }

    bool test_tri(const RayData& r, uint32_t t, float tmin, float& tmax, ClosestHit& best, bool any) const {
        float tt, u, v;
        if (!ray_triangle(r, vert(t, 0), vert(t, 1), vert(t, 2), tt, u, v)) return false;
        if (any) return tt > tmin && tt < tmax;
        bool closer = tt < tmax || (tt == tmax && best.tri >= 0 && static_cast<int32_t>(t) < best.tri);
        if (tt > tmin && closer) {
            tmax = tt;
            best.tri = static_cast<int32_t>(t);
            best.t = tt;
            best.u = u;
            best.v = v;
        }
        return false;
    }

    // bvh::SingleRayTraverser restated: from the root, test both children, visit the nearer
    // (the first on a tie), push the other.
    bool traverse(const Ray& ray, bool any, ClosestHit& best, Counters& c) const {
        RayData r = prepare(ray);
        float tmax = ray.tmax;
        best = ClosestHit{};
        if (nodes[0].count) {
            for (uint32_t i = 0; i < nodes[0].count; ++i) {
                ++c.tris;
                if (test_tri(r, prims[nodes[0].index + i], ray.tmin, tmax, best, any)) return true;
            }
            return !any && best.tri >= 0;
        }
        uint32_t stack[128];
        int sp = 0;
        uint32_t cur = 0;
        while (true) {
            const Node& n = nodes[cur];
            if (n.count == 0) {
                ++c.inner;
                uint32_t l = n.index, rr = n.index + 1;
                float tl, tr;
                bool hl = ray_box(r, nodes[l].b, ray.tmin, tmax, tl);
                bool hr = ray_box(r, nodes[rr].b, ray.tmin, tmax, tr);
                if (hl && hr) {
                    uint32_t near = l, far = rr;
                    if (tr < tl) std::swap(near, far);
                    stack[sp++] = far;
                    cur = near;
                    continue;
                }
                if (hl) { cur = l; continue; }
                if (hr) { cur = rr; continue; }
            } else {
                for (uint32_t i = 0; i < n.count; ++i) {
                    ++c.tris;
                    if (test_tri(r, prims[n.index + i], ray.tmin, tmax, best, any)) return true;
                }
            }
            if (sp == 0) break;
            cur = stack[--sp];
        }
        return !any && best.tri >= 0;
    }

    bool brute(const Ray& ray, bool any, ClosestHit& best) const {
        RayData r = prepare(ray);
        float tmax = ray.tmax;
        best = ClosestHit{};
        for (uint32_t t = 0; t < idx.size() / 3; ++t)
            if (test_tri(r, t, ray.tmin, tmax, best, any)) return true;
        return !any && best.tri >= 0;
    }

    // Object-median BVH, leaves of at most 4 triangles: deliberately unlike the product's
    // builder; closest-hit results do not depend on the tree.
    void build_own() {
        uint32_t n = static_cast<uint32_t>(idx.size() / 3);
        prims.resize(n);
        std::iota(prims.begin(), prims.end(), 0u);
        std::vector<Vec3> centre(n);
        for (uint32_t t = 0; t < n; ++t) centre[t] = (vert(t, 0) + vert(t, 1) + vert(t, 2)) * (1.0f / 3.0f);
        nodes.clear();
        nodes.push_back({});
        struct Item { uint32_t node, begin, end; };
        std::vector<Item> work{{0, 0, n}};
        while (!work.empty()) {
            Item it = work.back();
            work.pop_back();
            float b[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
            float cb[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
            for (uint32_t i = it.begin; i < it.end; ++i) {
                for (int k = 0; k < 3; ++k) {
                    Vec3 p = vert(prims[i], k);
                    for (int a = 0; a < 3; ++a) {
                        b[2 * a] = std::min(b[2 * a], get(p, a));
                        b[2 * a + 1] = std::max(b[2 * a + 1], get(p, a));
                    }
                }
                for (int a = 0; a < 3; ++a) {
                    cb[2 * a] = std::min(cb[2 * a], get(centre[prims[i]], a));
                    cb[2 * a + 1] = std::max(cb[2 * a + 1], get(centre[prims[i]], a));
                }
            }
            std::memcpy(nodes[it.node].b, b, sizeof b);
            if (it.end - it.begin <= 4) {
                nodes[it.node].index = it.begin;
                nodes[it.node].count = it.end - it.begin;
                continue;
            }
            int axis = 0;
            for (int a = 1; a < 3; ++a)
                if (cb[2 * a + 1] - cb[2 * a] > cb[2 * axis + 1] - cb[2 * axis]) axis = a;
            uint32_t mid = (it.begin + it.end) / 2;
            std::nth_element(prims.begin() + it.begin, prims.begin() + mid, prims.begin() + it.end,
                             [&](uint32_t x, uint32_t y) {
                                 float cx = get(centre[x], axis), cy = get(centre[y], axis);
                                 return cx < cy || (cx == cy && x < y);
                             });
            uint32_t first = static_cast<uint32_t>(nodes.size());
            nodes[it.node].index = first;
            nodes[it.node].count = 0;
            nodes.push_back({});
            nodes.push_back({});
            work.push_back({first, it.begin, mid});
            work.push_back({first + 1, mid, it.end});
        }
    }
};


namespace {

// ---- hit decoding (triangle_mesh.cpp:78-104) ----------------------------------------------------

struct SurfaceInfo {
    bool is_front_side;
    Vec3 point;
    float surf_u, surf_v;
    Vec3 face_normal;
    Mat3 local;
    Vec3 normal() const { return local.c2; }
};

struct Hit {
    SurfaceInfo surf_info;
    int light;   // -1: none
    int bsdf;    // -1: none
};

bool intersect_closest(const oracle_scene& s, Ray& ray, Hit& hit, Counters& c) {
    ClosestHit best;
    if (!s.traverse(ray, false, best, c)) return false;
    uint32_t t = static_cast<uint32_t>(best.tri);
    ray.tmax = best.t;
    Vec3 face_normal = cross(s.vert(t, 1) - s.vert(t, 0), s.vert(t, 2) - s.vert(t, 0));
    Vec3 normal = lerp(s.nrm[s.idx[t * 3]], s.nrm[s.idx[t * 3 + 1]], s.nrm[s.idx[t * 3 + 2]], best.u, best.v);
    bool front = dot(face_normal, ray.dir) < 0.0f;
    if (!front) {
        face_normal = -face_normal;
        normal = -normal;
    }
    SurfaceInfo& si = hit.surf_info;
    si.is_front_side = front;
    si.point = ray.org + ray.dir * ray.tmax;
    si.surf_u = best.u;
    si.surf_v = best.v;
    si.face_normal = face_normal;
    si.local = ortho_basis(normalize(normal));
    hit.light = s.tri_light[t];
    hit.bsdf = s.tri_mat[t];
    return true;
}

bool intersect_any(const oracle_scene& s, const Ray& ray, Counters& c) {
    ClosestHit unused;
    return s.traverse(ray, true, unused, c);
}

// ---- BSDFs (bsdfs.cpp) ----------------------------------------------------------------------------

enum class BsdfType { Diffuse, Glossy, Specular };

// InterpBsdf::infer_type (bsdfs.cpp:226-230): the mix of a diffuse and a Phong lobe is Diffuse.
BsdfType bsdf_type(const sol_material& m) {
    if (m.type == SOL_BSDF_DIFFUSE || m.type == SOL_BSDF_INTERP) return BsdfType::Diffuse;
    if (m.type == SOL_BSDF_PHONG) return BsdfType::Glossy;
    return BsdfType::Specular;
}

// The two BSDFs an InterpBsdf mixes (a = diffuse, b = Phong, as the OBJ loader builds it).
sol_material diffuse_part(const sol_material& m) {
    sol_material d = m;
    d.type = SOL_BSDF_DIFFUSE;
    return d;
}
sol_material phong_part(const sol_material& m) {
    sol_material p = m;
    p.type = SOL_BSDF_PHONG;
    return p;
}

struct BsdfSample {
    Vec3 in_dir;
    float pdf;
    float cos;
    Color color;
};

float reflect_cosine(Vec3 in_dir, Vec3 normal, Vec3 out_dir) { return -negative_dot(in_dir, reflect(out_dir, normal)); }

float fresnel_factor(float eta, float cos_i, float cos_t) {
    float rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return (rs * rs + rp * rp) * 0.5f;
}

Color phong_value(Color ks, float ns, float rc) { return ks * fpow(rc, ns) * (ns + 2.0f) * (0.5f * kInvPi); }

Color bsdf_eval(const sol_material& m, Vec3 in_dir, const SurfaceInfo& si, Vec3 out_dir) {
    switch (m.type) {
        case SOL_BSDF_DIFFUSE: return vec(m.kd) * kInvPi;
        case SOL_BSDF_PHONG: return phong_value(vec(m.ks), m.ns, reflect_cosine(in_dir, si.normal(), out_dir));
        case SOL_BSDF_INTERP:
            return lerp(bsdf_eval(diffuse_part(m), in_dir, si, out_dir), bsdf_eval(phong_part(m), in_dir, si, out_dir), m.mix);
        default: return {0, 0, 0};
    }
}

float bsdf_pdf(const sol_material& m, Vec3 in_dir, const SurfaceInfo& si, Vec3 out_dir) {
    switch (m.type) {
        case SOL_BSDF_DIFFUSE: return cosine_hemisphere_pdf(positive_dot(in_dir, si.normal()));
        case SOL_BSDF_PHONG: return cosine_power_hemisphere_pdf(m.ns_sample, reflect_cosine(in_dir, si.normal(), out_dir));
        case SOL_BSDF_INTERP:
            return lerp(bsdf_pdf(diffuse_part(m), in_dir, si, out_dir), bsdf_pdf(phong_part(m), in_dir, si, out_dir), m.mix);
        default: return 0.0f;
    }
}

bool validate(const SurfaceInfo& si, const BsdfSample& s, bool expect_below) {
    bool below = dot(s.in_dir, si.face_normal) < 0;
    return s.pdf > 0 && below == expect_below;
}

template <typename Rng>
bool bsdf_sample(const sol_material& m, Rng& rng, const SurfaceInfo& si, Vec3 out_dir, BsdfSample& s) {
    if (m.type == SOL_BSDF_DIFFUSE) {
        float u = rng();
        float v = rng();
        float pdf;
        Vec3 d = sample_cosine_hemisphere(u, v, pdf);
        s.in_dir = si.local * d;
        s.pdf = pdf;
        s.cos = d.z;
        s.color = bsdf_eval(m, s.in_dir, si, out_dir);
        return validate(si, s, false);
    }
    if (m.type == SOL_BSDF_PHONG) {
        Mat3 basis = ortho_basis(reflect(-out_dir, si.normal()));
        float u = rng();
        float v = rng();
        float pdf;
        Vec3 d = sample_cosine_power_hemisphere(m.ns_sample, u, v, pdf);
        s.in_dir = basis * d;
        s.pdf = pdf;
        s.cos = positive_dot(s.in_dir, si.normal());
        s.color = phong_value(vec(m.ks), m.ns, reflect_cosine(s.in_dir, si.normal(), out_dir));
        return validate(si, s, false);
    }
    if (m.type == SOL_BSDF_INTERP) {
        float k = m.mix;
        sol_material target = phong_part(m), other = diffuse_part(m);
        if (rng() > k) {
            std::swap(target, other);
            k = 1.0f - k;
        }
        if (!bsdf_sample(target, rng, si, out_dir, s)) return false;
        s.pdf = lerp(bsdf_pdf(other, s.in_dir, si, out_dir), s.pdf, k);
        s.color = lerp(bsdf_eval(other, s.in_dir, si, out_dir), s.color, k);
        return true;
    }
    if (m.type == SOL_BSDF_GLASS) {
        float eta = si.is_front_side ? m.eta : 1.0f / m.eta;
        float cos_i = dot(out_dir, si.normal());
        float cos2_t = 1.0f - eta * eta * (1.0f - cos_i * cos_i);
        if (cos2_t > 0.0f) {
            float cos_t = std::sqrt(cos2_t);
            if (rng() > fresnel_factor(eta, cos_i, cos_t)) {
                s.in_dir = si.normal() * (eta * cos_i - cos_t) - out_dir * eta;
                s.pdf = 1.0f;
                s.cos = 1.0f;
                s.color = vec(m.kt);
                return validate(si, s, true);
            }
        }
    }
    if (m.type == SOL_BSDF_GLASS || m.type == SOL_BSDF_MIRROR) {
        s.in_dir = reflect(-out_dir, si.normal());
        s.pdf = 1.0f;
        s.cos = 1.0f;
        s.color = vec(m.ks);
        return validate(si, s, false);
    }
    return false;
}

// ---- lights (lights.cpp, shapes.cpp) ---------------------------------------------------------------

struct LightAreaSample {
    Vec3 pos;
    Color intensity;
    float pdf_from;
    float cos;
};

bool sample_area(const sol_light& l, Pcg& rng, Vec3 from, LightAreaSample& out) {
    if (l.type == SOL_LIGHT_POINT) {
        out = {vec(l.v0), vec(l.intensity), 1.0f, 1.0f};
        return true;
    }
    float u = rng();
    float v = rng();
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    Vec3 pos = lerp(vec(l.v0), vec(l.v1), vec(l.v2), u, v);
    Vec3 dir = normalize(from - pos);
    float cos = positive_dot(dir, vec(l.normal));
    if (!(cos > 0)) return false;
    out = {pos, vec(l.intensity), l.inv_area, cos};
    return true;
}

struct Emission {
    Color intensity;
    float pdf_from;
};

Emission emission(const sol_light& l, Vec3 dir) {
    if (l.type == SOL_LIGHT_POINT) return {{0, 0, 0}, 1.0f};
    float cos = dot(dir, vec(l.normal));
    if (cos <= 0) return {{0, 0, 0}, 1.0f};
    return {vec(l.intensity), l.inv_area};
}

float balance_heuristic(float x, float y) { return 1.0f / (1.0f + y / x); }

// ---- trace_path (path_tracer.cpp:31-120) -----------------------------------------------------------

struct Tally {
    Counters ext, shadow;
    uint64_t ext_rays = 0, shadow_rays = 0, bsdf_hits = 0;
};

// What one surface vertex of a path contributes, recorded for the adjoint (oracle_backward).
// Radiance of the path: C = sum_k T_k * (emission_k + nee_k), T_0 = 1, T_{k+1} = T_k * f_k * a_k.
struct Vertex {
    int em_light = -1;       // light hit here whose emission applies, -1 = none
    float em_w = 0.0f;       // its MIS weight: emission_k = Le * em_w
    Color em_le{0, 0, 0};
    int mat = -1;            // material of the vertex (-1: none, path ends)
    int nee_light = -1;      // light of an unoccluded NEE sample, -1 = none
    Color nee_le{0, 0, 0};
    Color nee_f{0, 0, 0};    // BSDF value towards the light
    float nee_rc = 0.0f;     // Phong reflect cosine towards the light
    float nee_s = 0.0f;      // nee_k = Le * f * nee_s
    bool cont = false;       // the path continues from here
    Color f{0, 0, 0};        // BSDF value of the sampled direction
    float a = 0.0f;          // cos / (pdf * survival probability)
    float rc = 0.0f;         // Phong reflect cosine of the sampled direction
};

Color trace_path(const oracle_scene& s, const sol_pt_config& cfg, Pcg& sampler, Ray ray, Tally& tally,
                 std::vector<Vertex>* rec = nullptr) {
    const bool disable_mis = cfg.flags & SOL_PT_DISABLE_MIS;
    const bool disable_nee = cfg.flags & SOL_PT_DISABLE_NEE;
    const bool disable_rr = cfg.flags & SOL_PT_DISABLE_RR;
    const uint32_t num_lights = static_cast<uint32_t>(s.lights.size());
    float light_pick_prob = 1.0f / static_cast<float>(num_lights);
    float pdf_prev_bounce = 0.0f;
    Color throughput{1.0f, 1.0f, 1.0f};
    Color color{0.0f, 0.0f, 0.0f};

    for (uint32_t path_len = 0; path_len < cfg.max_path_len; path_len++) {
        Hit hit;
        ++tally.ext_rays;
        if (!intersect_closest(s, ray, hit, tally.ext)) break;
        Vec3 out_dir = -ray.dir;
        const SurfaceInfo& si = hit.surf_info;
        Vertex* vx = nullptr;
        if (rec) {
            rec->emplace_back();
            vx = &rec->back();
        }

        if (hit.light >= 0 && si.is_front_side) {
            float pdf_prev_bounce_area = pdf_prev_bounce * dot(out_dir, si.normal()) / (ray.tmax * ray.tmax);
            Emission e = emission(s.lights[hit.light], out_dir);
            float mis_weight = pdf_prev_bounce != 0.0f
                                   ? balance_heuristic(pdf_prev_bounce_area, e.pdf_from * light_pick_prob)
                                   : 1.0f;
            if (disable_mis || disable_nee) mis_weight = pdf_prev_bounce != 0 ? 0.0f : 1.0f;
            color = color + throughput * e.intensity * mis_weight;
            const sol_light& hl = s.lights[hit.light];
            if (vx && hl.type == SOL_LIGHT_TRIANGLE && dot(out_dir, vec(hl.normal)) > 0.0f) {
                vx->em_light = hit.light;
                vx->em_w = mis_weight;
                vx->em_le = e.intensity;
            }
        }

        if (hit.bsdf < 0) break;
        const sol_material& bsdf = s.materials[hit.bsdf];
        ++tally.bsdf_hits;
        if (vx) vx->mat = hit.bsdf;

        bool skip_nee = disable_nee || bsdf_type(bsdf) == BsdfType::Specular;
        if (!skip_nee) {
            float xi = sampler();
            uint32_t light_index = std::min(static_cast<uint32_t>(xi * static_cast<float>(num_lights)), num_lights - 1);
            const sol_light& light = s.lights[light_index];
            LightAreaSample ls;
            if (sample_area(light, sampler, si.point, ls)) {
                Vec3 in_dir = ls.pos - si.point;
                float cos_surf = dot(in_dir, si.normal());
                Ray shadow_ray{si.point, in_dir, cfg.ray_offset, 1.0f - cfg.ray_offset};
                ++tally.shadow_rays;
                if (!intersect_any(s, shadow_ray, tally.shadow)) {
                    float inv_light_dist = 1.0f / length(in_dir);
                    cos_surf *= inv_light_dist;
                    in_dir = in_dir * inv_light_dist;
                    bool has_area = light.type == SOL_LIGHT_TRIANGLE;
                    float pdf_bounce = has_area ? bsdf_pdf(bsdf, in_dir, si, out_dir) : 0.0f;
                    float pdf_light = ls.pdf_from * light_pick_prob;
                    float geom_term = ls.cos * inv_light_dist * inv_light_dist;
                    float mis_weight = has_area ? balance_heuristic(pdf_light, pdf_bounce * geom_term) : 1.0f;
                    if (disable_mis) mis_weight = 1.0f;
                    Color f = bsdf_eval(bsdf, in_dir, si, out_dir);
                    float scale = geom_term * cos_surf * mis_weight / pdf_light;
                    color = color + ls.intensity * throughput * f * scale;
                    if (vx) {
                        vx->nee_light = static_cast<int>(light_index);
                        vx->nee_le = ls.intensity;
                        vx->nee_f = f;
                        vx->nee_rc = reflect_cosine(in_dir, si.normal(), out_dir);
                        vx->nee_s = scale;
                    }
                }
            }
        }

        float survival_prob = 1.0f;
        if (!disable_rr && path_len >= cfg.min_rr_path_len) {
            survival_prob = std::min(std::max(luminance(throughput), cfg.min_survival_prob), cfg.max_survival_prob);
            if (sampler() >= survival_prob) break;
        }

        BsdfSample bs;
        if (!bsdf_sample(bsdf, sampler, si, out_dir, bs)) break;
        float a = bs.cos / (bs.pdf * survival_prob);
        throughput = throughput * (bs.color * a);
        if (vx) {
            vx->cont = true;
            vx->f = bs.color;
            vx->a = a;
            vx->rc = reflect_cosine(bs.in_dir, si.normal(), out_dir);
        }
        ray = Ray{si.point, bs.in_dir, cfg.ray_offset, FLT_MAX};
        pdf_prev_bounce = skip_nee ? 0.0f : bs.pdf;
    }
    return color;
}

void accumulate_stats(oracle_stats* out, const Tally& t, uint64_t samples) {
    if (!out) return;
    out->extension_rays += t.ext_rays;
    out->shadow_rays += t.shadow_rays;
    out->ext_inner_visits += t.ext.inner;
    out->ext_tri_tests += t.ext.tris;
    out->shadow_inner_visits += t.shadow.inner;
    out->shadow_tri_tests += t.shadow.tris;
    out->bsdf_hits += t.bsdf_hits;
    out->samples += samples;
}

// Renderer::sample_pixel (renderer.h:50-54): the pixel's area in camera space [-1, 1]^2, y down
// (pixel (0, 0) is the top-left corner, v = +1); the x jitter is drawn first.
struct Uv {
    float u, v;
};
template <typename Rng>
Uv sample_pixel(Rng& sampler, uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    float jx = sampler();
    float jy = sampler();
    return {(static_cast<float>(x) + jx) * (2.0f / static_cast<float>(w)) - 1.0f,
            1.0f - (static_cast<float>(y) + jy) * (2.0f / static_cast<float>(h))};
}

// PerspectiveCamera::generate_ray (cameras.cpp:23-25) over the vectors its constructor scaled
// (cameras.cpp:5-21, in oracle_scene_create).
Ray camera_ray(const oracle_scene& s, Uv uv) {
    return Ray{s.eye, normalize(s.cam_dir + s.cam_right * uv.u + s.cam_up * uv.v), 0.0f, FLT_MAX};
}

// Scripted random sequence for the BSDF closed-form checks.
struct Script {
    const float* values;
    int i = 0;
    float operator()() { return values[i++]; }
};

} // namespace

extern "C" {

int oracle_scene_create(const sol_mesh_view* m, const sol_bvh_view* bvh, oracle_scene** out) {
    if (!m || !out || !m->num_tris || !m->num_lights) return -1;
    auto* s = new oracle_scene();
    s->pos.resize(m->num_vertices);
    s->nrm.resize(m->num_vertices);
    for (uint32_t i = 0; i < m->num_vertices; ++i) {
        s->pos[i] = vec(m->positions + 3 * i);
        s->nrm[i] = vec(m->normals + 3 * i);
    }
    s->idx.assign(m->indices, m->indices + 3 * size_t(m->num_tris));
    s->tri_mat.assign(m->tri_material, m->tri_material + m->num_tris);
    s->tri_light.assign(m->tri_light, m->tri_light + m->num_tris);
    s->materials.assign(m->materials, m->materials + m->num_materials);
    s->lights.assign(m->lights, m->lights + m->num_lights);
    s->cam_params = m->camera;
    // PerspectiveCamera constructor (cameras.cpp:5-21).
    const sol_camera_params& c = m->camera;
    Vec3 dir = vec(c.dir), up = vec(c.up);
    s->eye = vec(c.eye);
    s->cam_dir = normalize(dir);
    Vec3 right = normalize(cross(dir, up));
    Vec3 cam_up = cross(right, s->cam_dir);
    float w = std::tan(c.fov * static_cast<float>(M_PI) / 360.0f);
    float h = w / c.aspect;
    s->cam_right = right * w;
    s->cam_up = cam_up * h;
    if (bvh) {
        s->nodes.resize(bvh->num_nodes);
        for (uint32_t i = 0; i < bvh->num_nodes; ++i) {
            std::memcpy(s->nodes[i].b, bvh->nodes[i].bounds, sizeof(float) * 6);
            s->nodes[i].index = bvh->nodes[i].index;
            s->nodes[i].count = bvh->nodes[i].prim_count;
        }
        s->prims.assign(bvh->prim_indices, bvh->prim_indices + bvh->num_prims);
    } else {
        s->build_own();
    }
    *out = s;
    return 0;
}

void oracle_scene_destroy(oracle_scene* s) { delete s; }

int oracle_scene_set_materials(oracle_scene* s, const sol_material* m, uint32_t n) {
    if (!s || n != s->materials.size()) return -1;
    s->materials.assign(m, m + n);
    return 0;
}

int oracle_scene_set_lights(oracle_scene* s, const sol_light* l, uint32_t n) {
    if (!s || n != s->lights.size()) return -1;
    s->lights.assign(l, l + n);
    return 0;
}

int oracle_intersect(const oracle_scene* s, const sol_ray* rays, uint64_t n, int any_hit, int brute_force, void* out,
                     oracle_stats* stats) {
    if (!s || (n && (!rays || !out))) return -1;
    uint64_t inner = 0, tris = 0;
    #pragma omp parallel for schedule(dynamic, 1024) reduction(+ : inner, tris)
    for (long long i = 0; i < static_cast<long long>(n); ++i) {
        const sol_ray& r = rays[i];
        Ray ray{vec(r.org), vec(r.dir), r.tmin, r.tmax};
        ClosestHit best;
        Counters c;
        bool found = brute_force ? s->brute(ray, any_hit != 0, best) : s->traverse(ray, any_hit != 0, best, c);
        inner += c.inner;
        tris += c.tris;
        if (any_hit) {
            static_cast<int32_t*>(out)[i] = found ? 1 : 0;
        } else {
            sol_hit& h = static_cast<sol_hit*>(out)[i];
            h.prim = found ? best.tri : -1;
            h.t = found ? best.t : r.tmax;
            h.u = found ? best.u : 0.0f;
            h.v = found ? best.v : 0.0f;
        }
    }
    if (stats) {
        if (any_hit) {
            stats->shadow_rays += n;
            stats->shadow_inner_visits += inner;
            stats->shadow_tri_tests += tris;
        } else {
            stats->extension_rays += n;
            stats->ext_inner_visits += inner;
            stats->ext_tri_tests += tris;
        }
    }
    return 0;
}

// PathTracer::render (path_tracer.cpp:9-23), par::omp::DynamicExecutor over the 2D pixel range.
int oracle_render(const oracle_scene* s, const sol_pt_config* cfg, uint32_t w, uint32_t h, uint64_t sample_index,
                  uint32_t sample_count, float* accum, int threads, uint32_t row_begin, uint32_t row_end,
                  oracle_stats* stats) {
    if (!s || !cfg || !accum || !w || !h) return -1;
    if (row_end == 0 || row_end > h) row_end = h;
    if (row_begin > row_end) return -1;
    if (threads <= 0) threads = omp_get_max_threads();
    const size_t plane = size_t(w) * h;
    const long long rows = row_end - row_begin;
    std::vector<Tally> tallies(threads);
    #pragma omp parallel for schedule(dynamic) collapse(2) num_threads(threads)
    for (long long yy = 0; yy < rows; ++yy) {
        for (long long x = 0; x < static_cast<long long>(w); ++x) {
            uint32_t y = row_begin + static_cast<uint32_t>(yy);
            Tally& tally = tallies[omp_get_thread_num()];
            Color color{0.0f, 0.0f, 0.0f};
            for (uint32_t i = 0; i < sample_count; ++i) {
                Pcg sampler(pixel_seed(sample_index + i, x, y));
                Ray ray = camera_ray(*s, sample_pixel(sampler, static_cast<uint32_t>(x), y, w, h));
                color = color + trace_path(*s, *cfg, sampler, ray, tally);
            }
            size_t p = size_t(y) * w + x;
            accum[p] += color.x;
            accum[plane + p] += color.y;
            accum[2 * plane + p] += color.z;
        }
    }
    for (const Tally& t : tallies) accumulate_stats(stats, t, 0);
    if (stats) stats->samples += uint64_t(rows) * w * sample_count;
    return 0;
}

// Gradients of sum(d_image * image) with respect to the diffuse albedo and Phong log-exponent of
// every material and the emission of every light, with sampling decisions, MIS weights and Russian
// roulette held constant. Reverse mode over each recorded path: the suffix sums
//   S_k = emission_k + nee_k + f_k * a_k * S_{k+1}
// give what vertex k's sampled BSDF value multiplies, and the forward throughput T_k what the
// parameters of vertex k multiply:
//   d/dtheta (material of k) += T_k * (df_nee/dtheta * Le * nee_s + df_k/dtheta * a_k * S_{k+1})
//   d/dLe    (emission at k) += T_k * em_w;   d/dLe (NEE light of k) += T_k * f_nee * nee_s
// with df/dkd_c = 1/pi (diffuse) and df/dlog ns = ks rc^ns ns ((ns + 2) ln rc + 1) / (2 pi) (Phong).
// Accumulated in double per thread, then added into the float outputs.
int oracle_backward(const oracle_scene* s, const sol_pt_config* cfg, uint32_t w, uint32_t h, uint64_t sample_index,
                    uint32_t sample_count, const float* d_image, float* d_materials, float* d_lights, int threads) {
    if (!s || !cfg || !d_image || !d_lights || !w || !h) return -1;
    const size_t nm = s->materials.size(), nl = s->lights.size();
    if (nm && !d_materials) return -1;
    if (threads <= 0) threads = omp_get_max_threads();
    const size_t plane = size_t(w) * h;
    std::vector<std::vector<double>> acc(threads, std::vector<double>(nm * 4 + nl * 3, 0.0));
    #pragma omp parallel for schedule(dynamic) collapse(2) num_threads(threads)
    for (long long y = 0; y < static_cast<long long>(h); ++y) {
        for (long long x = 0; x < static_cast<long long>(w); ++x) {
            size_t p = size_t(y) * w + x;
            const Color delta{d_image[p], d_image[plane + p], d_image[2 * plane + p]};
            if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) continue;
            std::vector<double>& g = acc[omp_get_thread_num()];
            auto add3 = [&](size_t base, Color v) {
                g[base] += double(delta.x) * v.x;
                g[base + 1] += double(delta.y) * v.y;
                g[base + 2] += double(delta.z) * v.z;
            };
            auto add_material = [&](int mi, Color scaled_T, Color le_s, float rc) {
                // scaled_T * df/dtheta * le_s, per parameter of material mi
                const sol_material& m = s->materials[mi];
                // the mix weighs the diffuse lobe by 1 - mix and the Phong lobe by mix
                const bool mixed = m.type == SOL_BSDF_INTERP;
                const float wd = mixed ? 1.0f - m.mix : 1.0f, wp = mixed ? m.mix : 1.0f;
                if (m.type == SOL_BSDF_DIFFUSE || mixed) add3(size_t(mi) * 4, scaled_T * le_s * (kInvPi * wd));
                if ((m.type == SOL_BSDF_PHONG || mixed) && rc > 0.0f) {
                    float dlog = fpow(rc, m.ns) * m.ns * ((m.ns + 2.0f) * std::log(rc) + 1.0f) * (0.5f * kInvPi);
                    Color v = scaled_T * vec(m.ks) * le_s * (dlog * wp);
                    g[size_t(mi) * 4 + 3] += double(delta.x) * v.x + double(delta.y) * v.y + double(delta.z) * v.z;
                }
            };
            Tally tally;
            std::vector<Vertex> path;
            std::vector<Color> suffix;
            for (uint32_t i = 0; i < sample_count; ++i) {
                Pcg sampler(pixel_seed(sample_index + i, x, y));
                Ray ray = camera_ray(*s, sample_pixel(sampler, static_cast<uint32_t>(x), static_cast<uint32_t>(y), w, h));
                path.clear();
                trace_path(*s, *cfg, sampler, ray, tally, &path);
                const size_t K = path.size();
                suffix.assign(K + 1, Color{0, 0, 0});   // suffix[k] = S_k
                for (size_t k = K; k-- > 0;) {
                    const Vertex& vx = path[k];
                    Color S = vx.em_le * vx.em_w + vx.nee_le * vx.nee_f * vx.nee_s;
                    if (vx.cont) S = S + vx.f * vx.a * suffix[k + 1];
                    suffix[k] = S;
                }
                Color T{1, 1, 1};
                for (size_t k = 0; k < K; ++k) {
                    const Vertex& vx = path[k];
                    if (vx.em_light >= 0) add3(nm * 4 + size_t(vx.em_light) * 3, T * vx.em_w);
                    if (vx.nee_light >= 0) {
                        add3(nm * 4 + size_t(vx.nee_light) * 3, T * vx.nee_f * vx.nee_s);
                        add_material(vx.mat, T, vx.nee_le * vx.nee_s, vx.nee_rc);
                    }
                    if (vx.cont) {
                        add_material(vx.mat, T, suffix[k + 1] * vx.a, vx.rc);
                        T = T * vx.f * vx.a;
                    }
                }
            }
        }
    }
    for (const auto& g : acc) {
        for (size_t i = 0; i < nm * 4; ++i) d_materials[i] += static_cast<float>(g[i]);
        for (size_t i = 0; i < nl * 3; ++i) d_lights[i] += static_cast<float>(g[nm * 4 + i]);
    }
    return 0;
}

uint32_t oracle_pixel_seed(uint64_t sample_index, uint64_t x, uint64_t y) { return pixel_seed(sample_index, x, y); }

void oracle_pcg(uint64_t seed, uint32_t n, uint32_t* out_u32, float* out_float) {
    Pcg a(seed), b(seed);
    for (uint32_t i = 0; i < n; ++i) {
        if (out_u32) out_u32[i] = a.next();
        if (out_float) out_float[i] = b();
    }
}

void oracle_sample_cosine_hemisphere(float u, float v, float* dir3, float* pdf) {
    Vec3 d = sample_cosine_hemisphere(u, v, *pdf);
    dir3[0] = d.x; dir3[1] = d.y; dir3[2] = d.z;
}

void oracle_sample_cosine_power_hemisphere(float n, float u, float v, float* dir3, float* pdf) {
    Vec3 d = sample_cosine_power_hemisphere(n, u, v, *pdf);
    dir3[0] = d.x; dir3[1] = d.y; dir3[2] = d.z;
}

void oracle_ortho_basis(const float* n3, float* t3, float* b3) {
    Mat3 m = ortho_basis(vec(n3));
    t3[0] = m.c0.x; t3[1] = m.c0.y; t3[2] = m.c0.z;
    b3[0] = m.c1.x; b3[1] = m.c1.y; b3[2] = m.c1.z;
}

int oracle_ray_triangle(const float* org3, const float* dir3, const float* v0, const float* v1, const float* v2,
                        float* t, float* u, float* v) {
    RayData r = prepare(Ray{vec(org3), vec(dir3), 0.0f, FLT_MAX});
    return ray_triangle(r, vec(v0), vec(v1), vec(v2), *t, *u, *v) ? 1 : 0;
}

int oracle_bsdf_sample(const sol_material* m, const float* n3, const float* out3, int front, const float* randoms,
                       float* in3, float* pdf, float* cos, float* rgb) {
    SurfaceInfo si;
    si.is_front_side = front != 0;
    si.point = {0, 0, 0};
    si.surf_u = si.surf_v = 0;
    si.face_normal = vec(n3);
    si.local = ortho_basis(vec(n3));
    Script rng{randoms};
    BsdfSample s{};
    bool ok = bsdf_sample(*m, rng, si, vec(out3), s);
    in3[0] = s.in_dir.x; in3[1] = s.in_dir.y; in3[2] = s.in_dir.z;
    *pdf = s.pdf;
    *cos = s.cos;
    rgb[0] = s.color.x; rgb[1] = s.color.y; rgb[2] = s.color.z;
    return ok ? 1 : 0;
}

float oracle_fresnel(float eta, float cos_i, float cos_t) { return fresnel_factor(eta, cos_i, cos_t); }

int oracle_intersect_surface(const oracle_scene* s, const sol_ray* rays, uint64_t n, sol_surface_hit* out) {
    if (!s || (n && (!rays || !out))) return -1;
    #pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < static_cast<long long>(n); ++i) {
        const sol_ray& r = rays[i];
        Ray ray{vec(r.org), vec(r.dir), r.tmin, r.tmax};
        Hit hit;
        Counters c;
        sol_surface_hit& o = out[i];
        o = sol_surface_hit{};
        o.material = -1;
        o.light = -1;
        if (intersect_closest(*s, ray, hit, c)) {
            const SurfaceInfo& si = hit.surf_info;
            o.hit = 1;
            o.front = si.is_front_side ? 1 : 0;
            o.material = hit.bsdf;
            o.light = hit.light;
            const Vec3 cols[5] = {si.point, si.face_normal, si.local.c0, si.local.c1, si.local.c2};
            float* dst[5] = {o.point, o.face_normal, o.tangent, o.bitangent, o.normal};
            for (int k = 0; k < 5; ++k) {
                dst[k][0] = cols[k].x;
                dst[k][1] = cols[k].y;
                dst[k][2] = cols[k].z;
            }
            o.surf_coords[0] = si.surf_u;
            o.surf_coords[1] = si.surf_v;
        }
        o.t = ray.tmax;
    }
    return 0;
}

void oracle_sample_pixel(uint64_t sample_index, uint32_t x, uint32_t y, uint32_t w, uint32_t h, float* uv) {
    Pcg sampler(pixel_seed(sample_index, x, y));
    Uv p = sample_pixel(sampler, x, y, w, h);
    uv[0] = p.u;
    uv[1] = p.v;
}

int oracle_camera_ray(const oracle_scene* s, float u, float v, float* org3, float* dir3) {
    if (!s || !org3 || !dir3) return -1;
    Ray r = camera_ray(*s, Uv{u, v});
    org3[0] = r.org.x; org3[1] = r.org.y; org3[2] = r.org.z;
    dir3[0] = r.dir.x; dir3[1] = r.dir.y; dir3[2] = r.dir.z;
    return 0;
}

} // extern "C"
