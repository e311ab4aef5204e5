// Device code shared by the kernels that run whole paths per lane (the megakernel and the adjoint
// replay): wave-level work fetch, the traversal stack of a lane, and trace_path.
#pragma once

#include "launch.h"
#include "path.h"
#include "traverse.h"

namespace solhip {

SOL_DEV unsigned long long wave_grab(unsigned long long* counter, unsigned chunk) {
    unsigned long long base = 0;
    if (__lane_id() == 0) base = atomicAdd(counter, static_cast<unsigned long long>(chunk));
    return __shfl(base, 0);
}

SOL_DEV void wave_add(unsigned long long* counter, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (__lane_id() == 0 && v) atomicAdd(counter, v);
}

template <int BLOCK, int LDS>
SOL_DEV Stack make_stack(int* lds_stack, const TraversalScratch& sc) {
    int gtid = blockIdx.x * BLOCK + threadIdx.x;
    return Stack::make(lds_stack, threadIdx.x, sc.spill, gtid, sc.max_threads, LDS + sc.spill_entries, sc.overflow);
}

// One sample of PathTracer::trace_path (src/algorithms/path_tracer.cpp:31-120 of sol): returns its
// radiance. n_ext / n_shadow count the traced rays.
// Closest / any hit of a wave-uniform scene over the binary (WIDE false) or 4-wide layout; the top
// levels of that layout are staged in lds_nodes. ANY_NEAR_FIRST: an any-hit query visits the hit
// children of a 4-wide node nearest first, as a closest-hit query does, instead of in slot order.
template <bool ANY, int BLOCK, int LDS, bool WIDE, bool ANY_NEAR_FIRST = false>
SOL_DEV bool trace_ray(const SceneView& s, const float4* lds_nodes, const Stack& st, const RayPre& r, float tmin,
                       float tmax, HitRecord& hit) {
    if (WIDE) return traverse4_with<ANY, BLOCK, LDS, ANY_NEAR_FIRST>(FetchUniform4(s, lds_nodes), st, r, tmin, tmax, hit);
    return traverse<ANY, BLOCK, LDS>(s, lds_nodes, st, r, tmin, tmax, hit);
}

template <int BLOCK, bool WIDE>
SOL_DEV void stage_top_nodes(const SceneView& s, float4* lds_nodes) {
    if (WIDE) load_top_nodes4<BLOCK>(s, lds_nodes);
    else load_top_nodes<BLOCK>(s, lds_nodes);
}

template <int BLOCK, int LDS, bool WIDE = false>
SOL_DEV V3 trace_path(const SceneView& s, const sol_pt_config& cfg, const float4* lds_nodes, const Stack& st,
                      Pcg& rng, V3 org, V3 dir, unsigned& n_ext, unsigned& n_shadow) {
    const float pick = 1.0f / static_cast<float>(s.num_lights);
    float pdf_prev = 0.0f;
    V3 T{1.0f, 1.0f, 1.0f};
    V3 color{0.0f, 0.0f, 0.0f};
    float tmin = 0.0f;
    for (uint32_t depth = 0; depth < cfg.max_path_len; ++depth) {
        RayPre r = make_ray_pre(org, dir);
        HitRecord hit;
        ++n_ext;
        if (!trace_ray<false, BLOCK, LDS, WIDE>(s, lds_nodes, st, r, tmin, FLT_MAX, hit)) break;
        Surface sf = decode_hit(s, org, dir, hit.slot, hit.t, hit.u, hit.v);
        if (sf.light >= 0 && sf.front) {
            V3 Le;
            float w;
            color = color + hit_emission(s, cfg.flags, sf, dir, hit.t, pdf_prev, pick, T, Le, w);
        }
        if (sf.mat < 0) break;
        const sol_material m = s.materials[sf.mat];
        bool skip_nee = (cfg.flags & SOL_PT_DISABLE_NEE) || is_specular(m.type);
        V3 out = -dir;
        if (!skip_nee) {
            NeeSample ns;
            if (nee_sample(s, cfg.flags, sf, m, out, T, rng, pick, ns)) {
                RayPre sr = make_ray_pre(sf.point, ns.shadow_dir);
                HitRecord unused;
                ++n_shadow;
                if (!trace_ray<true, BLOCK, LDS, WIDE>(s, lds_nodes, st, sr, cfg.ray_offset, 1.0f - cfg.ray_offset, unused))
                    color = color + ns.contrib;
            }
        }
        float q;
        if (!russian_roulette(cfg, depth, T, rng, q)) break;
        BsdfSample bs;
        if (!bsdf_sample(m, sf, ortho_basis(sf.n), out, rng, bs)) break;
        T = T * (bs.color * (bs.cos / (bs.pdf * q)));
        org = sf.point;
        dir = bs.in;
        tmin = cfg.ray_offset;
        pdf_prev = skip_nee ? 0.0f : bs.pdf;
    }
    return color;
}

} // namespace solhip
