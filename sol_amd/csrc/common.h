// Device-side building blocks of the path-tracing hot path (gfx950, wave64).
//
// Arithmetic is fp32, compiled with -ffp-contract=off and correctly rounded division and
// square root, so that every value is reproducible on the CPU. The three transcendental
// functions the path needs (cos and sin of 2*pi*u, and pow) are evaluated in fp64 and rounded
// to fp32 once: the fp64 results of two libraries differ by at most an ulp or two of fp64, so
// the fp32 results agree except in the rare case of a value within 2^-29 of a rounding
// boundary.
//
// Semantics follow sol (paths relative to the sol sources):
//   PCG32 and its float            include/sol/samplers.h:20-65
//   pixel seed (FNV-1a over x,y,i) include/sol/renderer.h:44-46
//   pixel jitter                   include/sol/renderer.h:50-54
//   camera ray                     src/cameras.cpp:23-25
//   BSDFs                          src/bsdfs.cpp:11-165, include/sol/bsdfs.h:89-94
//   lights / triangle sampling     src/lights.cpp:11-111, src/shapes.cpp:6-29
//   hit decoding                   src/triangle_mesh.cpp:64-117
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sol_types.h"

#define SOL_DEV __device__ __forceinline__
#define SOL_HD __host__ __device__ __forceinline__

namespace solhip {

constexpr float kInvPi = 0.318309886183790671538f;
constexpr double kTwoPiD = 6.283185307179586476925;
// 1 + 2 * gamma(3) for fp32 (Ize, "Robust BVH Ray Traversal", JCGT 2013): scaling the far slab
// distance by it makes the box test conservative under rounding.
constexpr float kRobustFar = 1.0f + 3.0f * 0x1p-23f;
// The triangle test computes t with its own rounding. Boxes are therefore culled against
// [0, tmax * (1 + 8 ulps)] rather than [tmin, tmax]: a triangle whose computed t passes the
// tmin < t < tmax test is not skipped because its box's slab distance rounded the other way
// (near tmin - a ray leaving a surface - the triangle's t can be off by far more than an ulp of t).
constexpr float kIntervalHigh = 1.0f + 0x1p-20f;
constexpr float kTinyDir = 1e-20f;

struct V3 {
    float x, y, z;
};

SOL_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
SOL_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SOL_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SOL_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
SOL_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
SOL_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
SOL_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SOL_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
SOL_HD float length(V3 a) { return sqrtf(dot(a, a)); }
SOL_HD V3 normalize(V3 a) { return a * (1.0f / length(a)); }
SOL_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
SOL_HD V3 load3(const float* p) { return V3{p[0], p[1], p[2]}; }
// reflect(d, n) = d - 2 dot(d, n) n
SOL_HD V3 reflect(V3 d, V3 n) { return d - n * (2.0f * dot(d, n)); }
// lerp(a, b, c, u, v) = a (1 - u - v) + b u + c v
SOL_HD V3 lerp3(V3 a, V3 b, V3 c, float u, float v) { return a * ((1.0f - u) - v) + b * u + c * v; }
// lerp(a, b, t) = a (1 - t) + b t, per channel for colours (color.h:66-71)
SOL_HD float lerp1(float a, float b, float t) { return a * (1.0f - t) + b * t; }
SOL_HD V3 lerp2(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
SOL_HD float luminance(V3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
SOL_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
SOL_HD float balance_heuristic(float x, float y) { return 1.0f / (1.0f + y / x); }

// ---- random numbers ----------------------------------------------------------------------------

struct Pcg {
    uint64_t state;
    SOL_DEV uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + 1ull;
        uint32_t xorshifted = static_cast<uint32_t>(((old >> 18) ^ old) >> 27);
        uint32_t rot = static_cast<uint32_t>(old >> 59);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    SOL_DEV float next_float() {
        float f = static_cast<float>(next()) * 0x1p-32f;
        return f < 0x1.fffffep-1f ? f : 0x1.fffffep-1f;
    }
    SOL_DEV void seed(uint64_t s) {
        state = 0;
        next();
        state += s;
        next();
    }
};

// FNV-1a (32 bit) over the eight little-endian bytes of x, then y, then the sample index.
SOL_DEV uint32_t fnv_combine(uint32_t h, uint64_t v) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        h ^= static_cast<uint32_t>((v >> (8 * i)) & 0xffu);
        h *= 0x01000193u;
    }
    return h;
}
SOL_DEV uint32_t pixel_seed(uint64_t sample_index, uint64_t x, uint64_t y) {
    uint32_t h = 0x811c9dc5u;
    h = fnv_combine(h, x);
    h = fnv_combine(h, y);
    h = fnv_combine(h, sample_index);
    return h;
}

// ---- transcendental functions, fp64 rounded once -----------------------------------------------

SOL_DEV float cos_2pi(float u) { return static_cast<float>(cos(kTwoPiD * static_cast<double>(u))); }
SOL_DEV float sin_2pi(float u) { return static_cast<float>(sin(kTwoPiD * static_cast<double>(u))); }
SOL_DEV float pow_det(float x, float y) { return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(y))); }

// ---- sampling ------------------------------------------------------------------------------------

// Cosine-weighted hemisphere around +z: phi = 2 pi u, cos(theta) = sqrt(1 - v); pdf = cos / pi.
SOL_DEV V3 sample_cosine_hemisphere(float u, float v, float& pdf) {
    float c = sqrtf(1.0f - v);
    float s = sqrtf(v);
    pdf = c * kInvPi;
    return V3{cos_2pi(u) * s, sin_2pi(u) * s, c};
}
SOL_DEV float cosine_hemisphere_pdf(float c) { return c * kInvPi; }

// Power-cosine lobe around +z: cos(theta) = v^(1/(n+1)); pdf = (n+1)/(2 pi) cos^n.
SOL_DEV float cosine_power_hemisphere_pdf(float n, float c) {
    return (n + 1.0f) * (0.5f * kInvPi) * pow_det(c, n);
}
SOL_DEV V3 sample_cosine_power_hemisphere(float n, float u, float v, float& pdf) {
    float c = pow_det(v, 1.0f / (n + 1.0f));
    float s = sqrtf(1.0f - c * c);
    pdf = cosine_power_hemisphere_pdf(n, c);
    return V3{cos_2pi(u) * s, sin_2pi(u) * s, c};
}

// Orthonormal basis around a unit vector (Duff et al., "Building an Orthonormal Basis,
// Revisited", JCGT 2017). Columns: t, b, n.
struct Basis {
    V3 t, b, n;
    SOL_DEV V3 to_world(V3 v) const { return t * v.x + b * v.y + n * v.z; }
};
SOL_DEV Basis ortho_basis(V3 n) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    Basis r;
    r.t = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    r.b = V3{b, sign + n.y * n.y * a, -n.y};
    r.n = n;
    return r;
}

// ---- scene ---------------------------------------------------------------------------------------

struct SceneView {
    const float4* nodes;   // 4 x float4 per node (sol_gpu_node)
    const float4* nodes4;  // 8 x float4 per node (sol_gpu_node4), null if not uploaded
    uint32_t num_nodes4;
    uint32_t top_nodes4;
    const float4* tris;    // 3 x float4 per triangle (sol_gpu_tri), leaf order
    const float4* shade;   // 4 x float4 per triangle (sol_gpu_shade), leaf order
    const sol_material* materials;
    const sol_light* lights;
    uint32_t num_lights;
    uint32_t top_nodes;
    uint32_t num_nodes;
    uint32_t num_tris;
    sol_camera camera;
};

// What the hit sorting needs to know of a triangle, its kind: the BSDF kind of its material + 1
// (0: no BSDF) in bits 0-2, bit 3 set if the triangle is a light.
SOL_HD uint8_t tri_kind(int bsdf, int light) {
    return static_cast<uint8_t>((bsdf + 1) | (light >= 0 ? 8 : 0));
}
SOL_HD int kind_bsdf(uint32_t kind) { return static_cast<int>(kind & 7u) - 1; }
SOL_HD bool kind_emissive(uint32_t kind) { return (kind & 8u) != 0; }

// The wavefront renderer's hits carry the hit triangle's kind next to its leaf-order index: index
// | kind << kHitKindShift (scenes hold fewer than 2^28 triangles, and a kind is at most 13, so no
// hit reads as the miss marker 0xFFFFFFFF). The kind comes from the fourth word of the triangle's
// second vertex row (sol_gpu_tri::pad0), which the library keeps equal to tri_kind.
constexpr uint32_t kHitKindShift = 28;
constexpr uint32_t kHitIndexMask = (1u << kHitKindShift) - 1u;

// ---- ray / triangle ------------------------------------------------------------------------------

// Per-ray constants of the watertight test (Woop, Benthin, Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2013) and of the slab test.
struct RayPre {
    V3 org, dir;
    V3 inv;          // 1 / dir, with |dir| < 1e-20 replaced by +-1e-20 (no 0 * inf)
    V3 off_lo, off_hi;  // slab distance of plane p on axis a: fma(p, inv.a, off_lo.a or off_hi.a)
    V3 off_near, off_far;   // the same offsets by role: on axis a the near plane is lo if inv.a >= 0, else hi
    int kx, ky, kz;
    float sx, sy, sz;
};

SOL_HD float safe_dir(float d) { return fabsf(d) < kTinyDir ? copysignf(kTinyDir, d) : d; }

SOL_HD RayPre make_ray_pre(V3 org, V3 dir) {
    RayPre r;
    r.org = org;
    r.dir = dir;
    r.inv = V3{1.0f / safe_dir(dir.x), 1.0f / safe_dir(dir.y), 1.0f / safe_dir(dir.z)};
    // fma(p, inv, -org * inv) = (p - org) * inv up to one rounding and the rounding of org * inv,
    // an absolute error of at most 2^-24 |org * inv|: the near plane's distance is lowered and the
    // far plane's raised by twice that (per axis; which plane is near depends on the sign of inv).
    auto offsets = [](float o, float inv, float& lo, float& hi, float& near_off, float& far_off) {
        float oi = o * inv;
        float e = 0x1p-23f * fabsf(oi);
        near_off = -oi - e;
        far_off = -oi + e;
        lo = inv >= 0.0f ? near_off : far_off;
        hi = inv >= 0.0f ? far_off : near_off;
    };
    offsets(org.x, r.inv.x, r.off_lo.x, r.off_hi.x, r.off_near.x, r.off_far.x);
    offsets(org.y, r.inv.y, r.off_lo.y, r.off_hi.y, r.off_near.y, r.off_far.y);
    offsets(org.z, r.inv.z, r.off_lo.z, r.off_hi.z, r.off_near.z, r.off_far.z);
    float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(dir, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.sx = comp(dir, kx) / dz;
    r.sy = comp(dir, ky) / dz;
    r.sz = 1.0f / dz;
    return r;
}

// Watertight test. On a hit inside the edge functions returns true with t, u (weight of v1)
// and v (weight of v2); the caller checks t against the ray interval.
SOL_HD bool intersect_triangle(const RayPre& r, V3 v0, V3 v1, V3 v2, float& t, float& u, float& v) {
    V3 A = v0 - r.org, B = v1 - r.org, C = v2 - r.org;
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float Ax = comp(A, r.kx) - r.sx * Akz, Ay = comp(A, r.ky) - r.sy * Akz;
    float Bx = comp(B, r.kx) - r.sx * Bkz, By = comp(B, r.ky) - r.sy * Bkz;
    float Cx = comp(C, r.kx) - r.sx * Ckz, Cy = comp(C, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        // Exact signs on the edges: products of two floats are exact in fp64.
        U = static_cast<float>(static_cast<double>(Cx) * By - static_cast<double>(Cy) * Bx);
        V = static_cast<float>(static_cast<double>(Ax) * Cy - static_cast<double>(Ay) * Cx);
        W = static_cast<float>(static_cast<double>(Bx) * Ay - static_cast<double>(By) * Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float inv_det = 1.0f / det;
    t = T * inv_det;
    u = V * inv_det;
    v = W * inv_det;
    return true;
}

// Slab test of one box against [0, tmax] (see kIntervalHigh); returns the entry distance through
// `tnear`. tmin is accepted for symmetry with the triangle test and not used.
// Each slab distance is one fused multiply-add, plane * inv + offset: the offsets widen every
// slab by the absolute error of the rounded org * inv (see make_ray_pre) and kRobustFar covers
// the relative rounding (Ize 2013). The test only decides which nodes are visited (the closest
// hit is the minimum over every triangle the exact test accepts), so it needs to be
// conservative, not identical to any other box test.
SOL_HD bool intersect_box(const RayPre& r, float lox, float hix, float loy, float hiy, float loz, float hiz,
                          float tmin, float tmax, float& tnear) {
    float tx0 = fmaf(lox, r.inv.x, r.off_lo.x), tx1 = fmaf(hix, r.inv.x, r.off_hi.x);
    float ty0 = fmaf(loy, r.inv.y, r.off_lo.y), ty1 = fmaf(hiy, r.inv.y, r.off_hi.y);
    float tz0 = fmaf(loz, r.inv.z, r.off_lo.z), tz1 = fmaf(hiz, r.inv.z, r.off_hi.z);
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * kRobustFar;
    f = fminf(f, tmax * kIntervalHigh);
    tnear = n;
    return n <= f;
}

// Byte offset of the near-plane row of an axis in a sol_gpu_node4 whose lo row is at `lo_row`: the
// lo row if inv >= 0, else the hi row 16 bytes on (inv is never zero). The far row is at near ^ 16.
SOL_HD uint32_t near_row(float inv, uint32_t lo_row) { return lo_row | ((__float_as_uint(inv) >> 27) & 16u); }

// The slab test with the near and far plane of every axis given (the plane rows selected by the
// sign of inv, near_row). For a box with lo <= hi these are the planes intersect_box picks
// with min and max - fma is monotonic in the plane and near_off <= far_off - so the decision and
// the entry distance are the same, for six operations fewer. (They could differ only where a slab
// distance is NaN, which needs |org| near 1e18 and a direction component below 1e-20.)
SOL_HD bool intersect_box_nf(const RayPre& r, float nx, float fx, float ny, float fy, float nz, float fz, float tmax,
                             float& tnear) {
    float tx0 = fmaf(nx, r.inv.x, r.off_near.x), tx1 = fmaf(fx, r.inv.x, r.off_far.x);
    float ty0 = fmaf(ny, r.inv.y, r.off_near.y), ty1 = fmaf(fy, r.inv.y, r.off_far.y);
    float tz0 = fmaf(nz, r.inv.z, r.off_near.z), tz1 = fmaf(fz, r.inv.z, r.off_far.z);
    float n = fmaxf(fmaxf(tx0, ty0), fmaxf(tz0, 0.0f));
    float f = fminf(fminf(tx1, ty1), tz1) * kRobustFar;
    f = fminf(f, tmax * kIntervalHigh);
    tnear = n;
    return n <= f;
}

// Leaf reference decoding: ~ref = (first << 3) | (count - 1).
SOL_HD void decode_leaf(int32_t ref, uint32_t& first, uint32_t& count) {
    uint32_t code = static_cast<uint32_t>(~ref);
    first = code >> 3;
    count = (code & 7u) + 1u;
}

} // namespace solhip
