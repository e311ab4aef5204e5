// Per-bounce pieces of the path tracer, shared by the megakernel, the wavefront kernels and the
// adjoint replay so that all of them perform the same fp32 operations in the same order.
// Reference: PathTracer::render / trace_path (src/algorithms/path_tracer.cpp:9-120 of sol) and
// the BSDF, light and shape code listed in common.h.
#pragma once

#include <float.h>

#include "common.h"

namespace solhip {

struct Surface {
    V3 point;      // hit point, org + t dir
    V3 n;          // shading normal: interpolated, flipped to the ray's side, normalised
    V3 face_n;     // geometric normal cross(v1 - v0, v2 - v0), flipped to the ray's side
    bool front;    // the ray hit the side the geometric normal faces
    int mat;       // material index, -1 = no BSDF
    int light;     // light index, -1 = not emissive
    int bsdf;      // SOL_BSDF_* of the material (from the shading record)
    float su, sv;  // surface coordinates (barycentrics of v1, v2)
};

// Hit decoding (triangle_mesh.cpp:78-104).
// The geometric normal cross(v1 - v0, v2 - v0) comes precomputed (same fp32 operations) in the
// shading record.
SOL_DEV Surface decode_hit(const SceneView& s, V3 org, V3 dir, int slot, float t, float u, float v) {
    const float4* sp = s.shade + slot * 4;
    float4 a = sp[0], b = sp[1], c = sp[2], d = sp[3];
    Surface sf;
    V3 face_n{c.y, c.z, c.w};
    V3 normal = lerp3(V3{a.x, a.y, a.z}, V3{a.w, b.x, b.y}, V3{b.z, b.w, c.x}, u, v);
    sf.front = dot(face_n, dir) < 0.0f;
    if (!sf.front) {
        face_n = -face_n;
        normal = -normal;
    }
    sf.face_n = face_n;
    sf.point = org + dir * t;
    sf.n = normalize(normal);
    sf.mat = __float_as_int(d.x);
    sf.light = __float_as_int(d.y);
    sf.bsdf = __float_as_int(d.z);
    sf.su = u;
    sf.sv = v;
    return sf;
}

SOL_DEV bool is_specular(int type) { return type == SOL_BSDF_GLASS || type == SOL_BSDF_MIRROR; }

// Emission of a front-side hit on a triangle light, weighted against the BSDF sample that
// found it (path_tracer.cpp:49-60; lights.cpp:94-106). Returns T * Le * w, the radiance added
// to the path (0 if none); Le_out and w_out receive the emission and the MIS weight (w_out = 0
// when the light emits nothing towards the ray, so that T * w_out is dradiance / dLe).
SOL_DEV V3 hit_emission(const SceneView& s, uint32_t flags, const Surface& sf, V3 dir, float t, float pdf_prev,
                        float pick_prob, V3 T, V3& Le_out, float& w_out) {
    Le_out = V3{0.0f, 0.0f, 0.0f};
    w_out = 0.0f;
    if (sf.light < 0 || !sf.front) return V3{0.0f, 0.0f, 0.0f};
    V3 out = -dir;
    float pdf_prev_area = pdf_prev * dot(out, sf.n) / (t * t);
    const sol_light& L = s.lights[sf.light];
    V3 Le{0.0f, 0.0f, 0.0f};
    float pdf_from = 1.0f;
    bool emits = false;
    if (L.type == SOL_LIGHT_TRIANGLE) {
        float c = dot(out, load3(L.normal));
        if (c > 0.0f) {
            Le = load3(L.intensity);
            pdf_from = L.inv_area;
            emits = true;
        }
    }
    float w = pdf_prev != 0.0f ? balance_heuristic(pdf_prev_area, pdf_from * pick_prob) : 1.0f;
    if (flags & (SOL_PT_DISABLE_MIS | SOL_PT_DISABLE_NEE)) w = pdf_prev != 0.0f ? 0.0f : 1.0f;
    Le_out = Le;
    w_out = emits ? w : 0.0f;
    return T * Le * w;
}

// ---- BSDFs ---------------------------------------------------------------------------------------

// reflect_cosine (bsdfs.cpp:11-19): -min(dot(in, reflect(out, n)), 0)
SOL_DEV float reflect_cosine(V3 in, V3 n, V3 out) { return -fminf(dot(in, reflect(out, n)), 0.0f); }

SOL_DEV V3 phong_eval(V3 ks, float ns, float rc) {
    return ks * pow_det(rc, ns) * (ns + 2.0f) * (0.5f * kInvPi);
}

// The two lobes, as DiffuseBsdf and PhongBsdf evaluate them (bsdfs.cpp:31-37, 64-70, 87-104).
SOL_DEV V3 diffuse_eval(const sol_material& m) { return load3(m.kd) * kInvPi; }
SOL_DEV float diffuse_pdf(V3 in, const Surface& sf) { return cosine_hemisphere_pdf(fmaxf(dot(in, sf.n), 0.0f)); }
SOL_DEV V3 phong_lobe_eval(const sol_material& m, V3 in, const Surface& sf, V3 out) {
    return phong_eval(load3(m.ks), m.ns, reflect_cosine(in, sf.n, out));
}
SOL_DEV float phong_pdf(const sol_material& m, V3 in, const Surface& sf, V3 out) {
    return cosine_power_hemisphere_pdf(m.ns_sample, reflect_cosine(in, sf.n, out));
}

// InterpBsdf(diffuse, Phong, mix): eval and pdf are lerp(diffuse, Phong, mix) (bsdfs.cpp:185-212).
SOL_DEV V3 bsdf_eval(const sol_material& m, V3 in, const Surface& sf, V3 out) {
    if (m.type == SOL_BSDF_DIFFUSE) return diffuse_eval(m);
    if (m.type == SOL_BSDF_PHONG) return phong_lobe_eval(m, in, sf, out);
    if (m.type == SOL_BSDF_INTERP) return lerp2(diffuse_eval(m), phong_lobe_eval(m, in, sf, out), m.mix);
    return V3{0.0f, 0.0f, 0.0f};
}

SOL_DEV float bsdf_pdf(const sol_material& m, V3 in, const Surface& sf, V3 out) {
    if (m.type == SOL_BSDF_DIFFUSE) return diffuse_pdf(in, sf);
    if (m.type == SOL_BSDF_PHONG) return phong_pdf(m, in, sf, out);
    if (m.type == SOL_BSDF_INTERP) return lerp1(diffuse_pdf(in, sf), phong_pdf(m, in, sf, out), m.mix);
    return 0.0f;
}

struct BsdfSample {
    V3 in;
    float pdf;
    float cos;
    V3 color;
    float rc;    // Phong and mixed lobes: reflect cosine of the sampled direction (for gradients)
};

// validate_sample (bsdfs.h:89-94)
SOL_DEV bool valid_sample(const Surface& sf, const BsdfSample& bs, bool expect_below) {
    bool below = dot(bs.in, sf.face_n) < 0.0f;
    return bs.pdf > 0.0f && below == expect_below;
}

SOL_DEV float fresnel_factor(float eta, float cos_i, float cos_t) {
    float rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return (rs * rs + rp * rp) * 0.5f;
}

SOL_DEV bool sample_diffuse_lobe(const sol_material& m, const Surface& sf, const Basis& basis, Pcg& rng,
                                 BsdfSample& bs) {
    float u = rng.next_float();
    float v = rng.next_float();
    float pdf;
    V3 l = sample_cosine_hemisphere(u, v, pdf);
    bs.in = basis.to_world(l);
    bs.pdf = pdf;
    bs.cos = l.z;
    bs.color = diffuse_eval(m);
    return valid_sample(sf, bs, false);
}

SOL_DEV bool sample_phong_lobe(const sol_material& m, const Surface& sf, V3 out, Pcg& rng, BsdfSample& bs) {
    Basis lobe = ortho_basis(reflect(-out, sf.n));
    float u = rng.next_float();
    float v = rng.next_float();
    float pdf;
    V3 l = sample_cosine_power_hemisphere(m.ns_sample, u, v, pdf);
    bs.in = lobe.to_world(l);
    bs.pdf = pdf;
    bs.cos = fmaxf(dot(bs.in, sf.n), 0.0f);
    bs.rc = reflect_cosine(bs.in, sf.n, out);
    bs.color = phong_eval(load3(m.ks), m.ns, bs.rc);
    return valid_sample(sf, bs, false);
}

// BSDF sampling (bsdfs.cpp:39-48, 72-85, 113-119, 137-164, 192-205). Returns false when the
// sample is rejected; the path then ends.
SOL_DEV bool bsdf_sample(const sol_material& m, const Surface& sf, const Basis& basis, V3 out, Pcg& rng,
                         BsdfSample& bs) {
    bs.rc = 0.0f;
    if (m.type == SOL_BSDF_DIFFUSE) return sample_diffuse_lobe(m, sf, basis, rng, bs);
    if (m.type == SOL_BSDF_PHONG) return sample_phong_lobe(m, sf, out, rng, bs);
    if (m.type == SOL_BSDF_INTERP) {
        // One draw picks the lobe: the Phong lobe (weight k) unless it exceeds k; the sampled
        // lobe's value and density are then mixed with the other lobe's.
        float k = m.mix;
        bool phong = !(rng.next_float() > k);
        if (!phong) k = 1.0f - k;
        bool ok = phong ? sample_phong_lobe(m, sf, out, rng, bs) : sample_diffuse_lobe(m, sf, basis, rng, bs);
        if (!ok) return false;
        if (phong) {
            bs.pdf = lerp1(diffuse_pdf(bs.in, sf), bs.pdf, k);
            bs.color = lerp2(diffuse_eval(m), bs.color, k);
        } else {
            bs.rc = reflect_cosine(bs.in, sf.n, out);
            bs.pdf = lerp1(cosine_power_hemisphere_pdf(m.ns_sample, bs.rc), bs.pdf, k);
            bs.color = lerp2(phong_eval(load3(m.ks), m.ns, bs.rc), bs.color, k);
        }
        return true;
    }
    if (m.type == SOL_BSDF_GLASS) {
        float eta = sf.front ? m.eta : 1.0f / m.eta;
        float cos_i = dot(out, sf.n);
        float cos2_t = 1.0f - eta * eta * (1.0f - cos_i * cos_i);
        if (cos2_t > 0.0f) {
            float cos_t = sqrtf(cos2_t);
            if (rng.next_float() > fresnel_factor(eta, cos_i, cos_t)) {
                bs.in = sf.n * (eta * cos_i - cos_t) - out * eta;
                bs.pdf = 1.0f;
                bs.cos = 1.0f;
                bs.color = load3(m.kt);
                return valid_sample(sf, bs, true);
            }
        }
        bs.in = reflect(-out, sf.n);
        bs.pdf = 1.0f;
        bs.cos = 1.0f;
        bs.color = load3(m.ks);
        return valid_sample(sf, bs, false);
    }
    if (m.type == SOL_BSDF_MIRROR) {
        bs.in = reflect(-out, sf.n);
        bs.pdf = 1.0f;
        bs.cos = 1.0f;
        bs.color = load3(m.ks);
        return valid_sample(sf, bs, false);
    }
    return false;
}

// ---- next-event estimation -----------------------------------------------------------------------

struct NeeSample {
    V3 shadow_dir;   // unnormalised, from the surface point to the light point
    V3 contrib;      // radiance added if the shadow ray is unoccluded
    // Terms kept for gradients.
    int light;
    V3 f;            // BSDF value towards the light
    float rc;        // Phong reflect cosine towards the light
    float scale;     // G cos_surf w / pdf_L: contrib = Le T f scale
};

// Light pick + light sample + everything of path_tracer.cpp:66-97 that does not depend on the
// occlusion test. Returns false if the light sample is rejected (no shadow ray).
SOL_DEV bool nee_sample(const SceneView& s, uint32_t flags, const Surface& sf, const sol_material& m, V3 out,
                        V3 T, Pcg& rng, float pick_prob, NeeSample& ns) {
    // pick_light (path_tracer.cpp:25-29)
    float xi = rng.next_float();
    uint32_t n = s.num_lights;
    uint32_t idx = static_cast<uint32_t>(xi * static_cast<float>(n));
    if (idx > n - 1) idx = n - 1;
    const sol_light& L = s.lights[idx];
    ns.light = static_cast<int>(idx);
    V3 pos, Le = load3(L.intensity);
    float pdf_from, cos_l;
    bool has_area = L.type == SOL_LIGHT_TRIANGLE;
    if (has_area) {
        float u = rng.next_float();
        float v = rng.next_float();
        if (u + v > 1.0f) {
            u = 1.0f - u;
            v = 1.0f - v;
        }
        pos = lerp3(load3(L.v0), load3(L.v1), load3(L.v2), u, v);
        V3 d = normalize(sf.point - pos);
        cos_l = fmaxf(dot(d, load3(L.normal)), 0.0f);
        if (!(cos_l > 0.0f)) return false;
        pdf_from = L.inv_area;
    } else {
        pos = load3(L.v0);
        pdf_from = 1.0f;
        cos_l = 1.0f;
    }
    V3 in = pos - sf.point;
    float cos_surf = dot(in, sf.n);
    ns.shadow_dir = in;
    float inv_d = 1.0f / length(in);
    cos_surf = cos_surf * inv_d;
    in = in * inv_d;
    float pdf_b = has_area ? bsdf_pdf(m, in, sf, out) : 0.0f;
    float pdf_L = pdf_from * pick_prob;
    float G = cos_l * inv_d * inv_d;
    float w = has_area ? balance_heuristic(pdf_L, pdf_b * G) : 1.0f;
    if (flags & SOL_PT_DISABLE_MIS) w = 1.0f;
    V3 f = bsdf_eval(m, in, sf, out);
    ns.f = f;
    ns.rc = (m.type == SOL_BSDF_PHONG || m.type == SOL_BSDF_INTERP) ? reflect_cosine(in, sf.n, out) : 0.0f;
    ns.scale = G * cos_surf * w / pdf_L;
    ns.contrib = Le * T * f * ns.scale;
    return true;
}

// Russian roulette (path_tracer.cpp:100-108). Returns false if the path is terminated;
// `q` receives the survival probability.
SOL_DEV bool russian_roulette(const sol_pt_config& cfg, uint32_t depth, V3 T, Pcg& rng, float& q) {
    q = 1.0f;
    if (!(cfg.flags & SOL_PT_DISABLE_RR) && depth >= cfg.min_rr_path_len) {
        q = clampf(luminance(T), cfg.min_survival_prob, cfg.max_survival_prob);
        if (rng.next_float() >= q) return false;
    }
    return true;
}

// Camera ray of one sample (renderer.h:50-54, cameras.cpp:23-25): two draws, x first.
SOL_DEV void camera_ray(const sol_camera& cam, uint32_t x, uint32_t y, uint32_t w, uint32_t h, Pcg& rng, V3& org,
                        V3& dir) {
    float jx = rng.next_float();
    float jy = rng.next_float();
    float u = (static_cast<float>(x) + jx) * (2.0f / static_cast<float>(w)) - 1.0f;
    float v = 1.0f - (static_cast<float>(y) + jy) * (2.0f / static_cast<float>(h));
    org = load3(cam.eye);
    dir = normalize(load3(cam.dir) + load3(cam.right) * u + load3(cam.up) * v);
}

} // namespace solhip
