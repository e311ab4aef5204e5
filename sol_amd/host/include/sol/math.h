// Small fp32 vector / box / ray types for the host side. sol takes these from the external
// `proto` library (not vendored, see src/CMakeLists.txt:27-29 of the reference); this header
// states the operations the host code needs from the mathematics. Operation order is written
// out explicitly (no fused multiply-add: the library is compiled with -ffp-contract=off) so that
// values computed here (camera basis, light normals and areas) are reproducible.
#ifndef SOL_MATH_H
#define SOL_MATH_H

#include <cmath>
#include <cstdint>
#include <limits>
#include <algorithm>

namespace sol {

struct Vec2f {
    float x = 0, y = 0;
    Vec2f() = default;
    constexpr Vec2f(float x, float y) : x(x), y(y) {}
    explicit constexpr Vec2f(float s) : x(s), y(s) {}
    float operator[](int i) const { return i == 0 ? x : y; }
};

struct Vec3f {
    float x = 0, y = 0, z = 0;
    Vec3f() = default;
    constexpr Vec3f(float x, float y, float z) : x(x), y(y), z(z) {}
    explicit constexpr Vec3f(float s) : x(s), y(s), z(s) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Vec3f operator-() const { return Vec3f(-x, -y, -z); }
    Vec3f& operator+=(const Vec3f& o) { x += o.x; y += o.y; z += o.z; return *this; }
    Vec3f& operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
    bool operator==(const Vec3f& o) const { return x == o.x && y == o.y && z == o.z; }
};

inline Vec3f operator+(const Vec3f& a, const Vec3f& b) { return Vec3f(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3f operator-(const Vec3f& a, const Vec3f& b) { return Vec3f(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3f operator*(const Vec3f& a, const Vec3f& b) { return Vec3f(a.x * b.x, a.y * b.y, a.z * b.z); }
inline Vec3f operator*(const Vec3f& a, float s) { return Vec3f(a.x * s, a.y * s, a.z * s); }
inline Vec3f operator*(float s, const Vec3f& a) { return Vec3f(s * a.x, s * a.y, s * a.z); }

inline float dot(const Vec3f& a, const Vec3f& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3f cross(const Vec3f& a, const Vec3f& b) {
    return Vec3f(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(const Vec3f& a) { return std::sqrt(dot(a, a)); }
inline Vec3f normalize(const Vec3f& a) { return a * (1.0f / length(a)); }
inline Vec3f vmin(const Vec3f& a, const Vec3f& b) {
    return Vec3f(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z));
}
inline Vec3f vmax(const Vec3f& a, const Vec3f& b) {
    return Vec3f(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z));
}

struct BBox {
    Vec3f lo, hi;
    static BBox empty() {
        constexpr float inf = std::numeric_limits<float>::infinity();
        return BBox{Vec3f(inf), Vec3f(-inf)};
    }
    BBox& extend(const BBox& o) { lo = vmin(lo, o.lo); hi = vmax(hi, o.hi); return *this; }
    BBox& extend(const Vec3f& p) { lo = vmin(lo, p); hi = vmax(hi, p); return *this; }
    Vec3f diagonal() const { return hi - lo; }
    Vec3f center() const { return (lo + hi) * 0.5f; }
    // Half of the surface area: the SAH only needs ratios.
    float half_area() const {
        Vec3f d = diagonal();
        if (d.x < 0 || d.y < 0 || d.z < 0) return 0.0f;
        return d.x * d.y + d.y * d.z + d.z * d.x;
    }
    int largest_axis() const {
        Vec3f d = diagonal();
        return d.x >= d.y && d.x >= d.z ? 0 : (d.y >= d.z ? 1 : 2);
    }
};

inline BBox merge(BBox a, const BBox& b) { return a.extend(b); }

inline Vec2f operator+(const Vec2f& a, const Vec2f& b) { return Vec2f(a.x + b.x, a.y + b.y); }
inline Vec2f operator*(const Vec2f& a, float s) { return Vec2f(a.x * s, a.y * s); }

// Barycentric interpolation a (1 - u - v) + b u + c v: u weighs the second vertex, v the third
// (the convention of the triangle test's (u, v) and of sol's hit decoding).
inline Vec3f lerp(const Vec3f& a, const Vec3f& b, const Vec3f& c, float u, float v) {
    return a * ((1.0f - u) - v) + b * u + c * v;
}
inline Vec2f lerp(const Vec2f& a, const Vec2f& b, const Vec2f& c, float u, float v) {
    return a * ((1.0f - u) - v) + b * u + c * v;
}

// A ray and its valid interval (proto::Rayf): points org + t dir for tmin < t < tmax; tmax
// defaults to the largest float. intersect_closest shrinks tmax to the hit's distance.
struct Rayf {
    Vec3f org, dir;
    float tmin = 0.0f;
    float tmax = std::numeric_limits<float>::max();
    Rayf() = default;
    Rayf(const Vec3f& org, const Vec3f& dir, float tmin = 0.0f, float tmax = std::numeric_limits<float>::max())
        : org(org), dir(dir), tmin(tmin), tmax(tmax) {}
    Vec3f point_at(float t) const { return org + dir * t; }
};

// 3x3 matrix by columns; a local frame has the tangent, bitangent and normal as its columns.
struct Mat3x3f {
    Vec3f c0, c1, c2;
    const Vec3f& col(int i) const { return i == 0 ? c0 : (i == 1 ? c1 : c2); }
    Vec3f operator*(const Vec3f& v) const { return c0 * v.x + c1 * v.y + c2 * v.z; }
};

// Orthonormal basis around the unit vector n, right-handed, n in the third column (Duff et al.,
// "Building an Orthonormal Basis, Revisited", JCGT 2017).
inline Mat3x3f ortho_basis(const Vec3f& n) {
    const float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    return Mat3x3f{Vec3f(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x), Vec3f(b, sign + n.y * n.y * a, -n.y), n};
}

struct Triangle {
    Vec3f v0, v1, v2;
    BBox bbox() const { BBox b = BBox::empty(); b.extend(v0); b.extend(v1); b.extend(v2); return b; }
    Vec3f center() const { return (v0 + v1 + v2) * (1.0f / 3.0f); }
    // Geometric normal, counter-clockwise winding (front side faces the normal).
    Vec3f raw_normal() const { return cross(v1 - v0, v2 - v0); }
    Vec3f normal() const { return normalize(raw_normal()); }
    float area() const { return 0.5f * length(raw_normal()); }
};

} // namespace sol

#endif
