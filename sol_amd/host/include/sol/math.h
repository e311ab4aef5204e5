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
