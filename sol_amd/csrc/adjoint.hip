// Adjoint path replay: gradients of an image loss with respect to the diffuse albedo and the
// Phong exponent (log ns) of every material and the emission of every light.
//
// The sample's radiance C comes from a forward pass of the wavefront renderer that keeps every
// sample's radiance (the same fp32 operations as trace_path, so the same C); the kernel replays
// each path (same PCG stream, same traversal) and attributes C to the parameters (path replay
// backpropagation, Vicini, Speierer and Jakob 2021). With sampling decisions, MIS
// weights and Russian roulette held constant, every term the path adds is a product of BSDF values
// f_j, one emission Le and constants, so with delta = dLoss/dImage at the pixel:
//   emission term T Le w at a light hit      dLe  += delta * T * w
//   NEE term Le T f s at vertex j            dLe  += delta * T * f * s, dtheta_j += delta * term * (df/dtheta) / f
//   continuation through the sampled f_j     dtheta_j += delta * R_j * (df/dtheta) / f
// where R_j is the radiance the path gathers after vertex j: C minus every term added up to and
// including vertex j's NEE. For the diffuse BSDF f = kd / pi, (df/dkd_c) / f_c = 1 / kd_c; for the
// Phong BSDF f = ks rc^ns (ns + 2) / (2 pi), (df/dlog ns) / f = ns (ln rc + 1 / (ns + 2)); for their
// mix f = (1 - k) kd / pi + k Phong, each lobe's share of f times its own ratio.
// Channels whose BSDF value is exactly zero (kd_c = 0, rc = 0) contribute nothing.
//
// Gradients are summed per workgroup in LDS (one float per parameter, LDS float atomics) and
// added to the output with one device atomic per parameter and workgroup; tables too large for the
// LDS budget go straight to device atomics.
#include "megapath.h"

namespace solhip {

namespace {

constexpr int kGradLds = 4096;   // floats of the per-workgroup gradient table

struct GradSink {
    float* lds;                  // per-workgroup table, or nullptr
    float* d_mat;                // num_materials x 4
    float* d_light;              // num_lights x 3
    int mat_floats;              // num_materials * 4

    SOL_DEV void add(int i, float v) const {
        if (v == 0.0f) return;
        if (lds) {
            atomicAdd(lds + i, v);
        } else if (i < mat_floats) {
            atomicAdd(d_mat + i, v);
        } else {
            atomicAdd(d_light + (i - mat_floats), v);
        }
    }

    SOL_DEV void light(int l, V3 v) const {
        int b = mat_floats + 3 * l;
        add(b, v.x);
        add(b + 1, v.y);
        add(b + 2, v.z);
    }

    // v = delta * (the radiance that carries the BSDF value f of material `mi` as a factor).
    SOL_DEV void material(const sol_material& m, int mi, float rc, V3 f, V3 v) const {
        if (m.type == SOL_BSDF_DIFFUSE) {
            if (m.kd[0] > 0.0f) add(4 * mi + 0, v.x / m.kd[0]);
            if (m.kd[1] > 0.0f) add(4 * mi + 1, v.y / m.kd[1]);
            if (m.kd[2] > 0.0f) add(4 * mi + 2, v.z / m.kd[2]);
        } else if (m.type == SOL_BSDF_PHONG) {
            if (rc > 0.0f) add(4 * mi + 3, (v.x + v.y + v.z) * (m.ns * (logf(rc) + 1.0f / (m.ns + 2.0f))));
        } else if (m.type == SOL_BSDF_INTERP) {
            // f = (1 - mix) kd / pi + mix Phong: the lobes' shares of f, channel by channel
            const float wd = (1.0f - m.mix) * kInvPi;
            if (f.x > 0.0f) add(4 * mi + 0, v.x * (wd / f.x));
            if (f.y > 0.0f) add(4 * mi + 1, v.y * (wd / f.y));
            if (f.z > 0.0f) add(4 * mi + 2, v.z * (wd / f.z));
            if (rc > 0.0f) {
                V3 p = phong_eval(load3(m.ks), m.ns, rc) * m.mix;
                float sp = (f.x > 0.0f ? v.x * (p.x / f.x) : 0.0f) + (f.y > 0.0f ? v.y * (p.y / f.y) : 0.0f) +
                           (f.z > 0.0f ? v.z * (p.z / f.z) : 0.0f);
                add(4 * mi + 3, sp * (m.ns * (logf(rc) + 1.0f / (m.ns + 2.0f))));
            }
        }
    }
};

// Replays one sample whose radiance is C, with the same operations as trace_path.
template <int BLOCK, int LDS, bool WIDE>
SOL_DEV void replay_path(const SceneView& s, const sol_pt_config& cfg, const float4* lds_nodes, const Stack& st,
                         Pcg& rng, V3 org, V3 dir, V3 C, V3 delta, const GradSink& g, unsigned& n_ext,
                         unsigned& n_shadow) {
    const float pick = 1.0f / static_cast<float>(s.num_lights);
    float pdf_prev = 0.0f;
    V3 T{1.0f, 1.0f, 1.0f};
    V3 R = C;
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
            V3 term = hit_emission(s, cfg.flags, sf, dir, hit.t, pdf_prev, pick, T, Le, w);
            g.light(sf.light, delta * (T * w));
            R = R - term;
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
                if (!trace_ray<true, BLOCK, LDS, WIDE>(s, lds_nodes, st, sr, cfg.ray_offset, 1.0f - cfg.ray_offset,
                                                       unused)) {
                    g.material(m, sf.mat, ns.rc, ns.f, delta * ns.contrib);
                    g.light(ns.light, delta * (T * ns.f * ns.scale));
                    R = R - ns.contrib;
                }
            }
        }
        float q;
        if (!russian_roulette(cfg, depth, T, rng, q)) break;
        BsdfSample bs;
        if (!bsdf_sample(m, sf, ortho_basis(sf.n), out, rng, bs)) break;
        // R is what the path gathers beyond this vertex; after the last bounce it is only the
        // rounding residue of the subtractions, so nothing is attributed there.
        if (depth + 1 < cfg.max_path_len) g.material(m, sf.mat, bs.rc, bs.color, delta * R);
        T = T * (bs.color * (bs.cos / (bs.pdf * q)));
        org = sf.point;
        dir = bs.in;
        tmin = cfg.ray_offset;
        pdf_prev = skip_nee ? 0.0f : bs.pdf;
    }
}

// One lane per pixel of an 8x8 tile and one sample: a wave takes (tile, sample) tasks, the samples
// of a tile consecutively, so that the work balances at the granularity of one path per lane;
// pixels with a zero adjoint are skipped.
template <int BLOCK, int LDS, bool WIDE>
__global__ __launch_bounds__(BLOCK) void adjoint_kernel(SceneView s, sol_pt_config cfg, uint32_t w, uint32_t h,
                                                        uint64_t sample_index, uint32_t sample_count,
                                                        const float4* __restrict__ radiance,
                                                        const float* __restrict__ d_image, uint32_t num_materials,
                                                        float* d_mat, float* d_light, unsigned long long* counter,
                                                        unsigned long long* stats, TraversalScratch sc) {
    __shared__ float4 lds_nodes[kTopLdsFloat4<WIDE>];
    __shared__ int lds_stack[LDS * BLOCK];
    __shared__ float lds_grad[kGradLds];
    const int mat_floats = static_cast<int>(num_materials) * 4;
    const int n_grad = mat_floats + static_cast<int>(s.num_lights) * 3;
    const bool in_lds = n_grad <= kGradLds;
    if (in_lds)
        for (int i = threadIdx.x; i < n_grad; i += BLOCK) lds_grad[i] = 0.0f;
    stage_top_nodes<BLOCK, WIDE>(s, lds_nodes);
    __syncthreads();
    Stack st = make_stack<BLOCK, LDS>(lds_stack, sc);
    GradSink g{in_lds ? lds_grad : nullptr, d_mat, d_light, mat_floats};
    const uint32_t tiles_x = (w + 7) / 8, tiles_y = (h + 7) / 8;
    const unsigned long long total = static_cast<unsigned long long>(tiles_x) * tiles_y * sample_count * 64;
    const size_t plane = static_cast<size_t>(w) * h;
    unsigned n_ext = 0, n_shadow = 0, n_samples = 0;
    while (true) {
        unsigned long long base = wave_grab(counter, 64);
        if (base >= total) break;
        const unsigned long long task = base / 64;
        const uint32_t tile = static_cast<uint32_t>(task / sample_count);
        const uint32_t i = static_cast<uint32_t>(task % sample_count);
        uint32_t lane = __lane_id();
        uint32_t x = (tile % tiles_x) * 8 + (lane & 7);
        uint32_t y = (tile / tiles_x) * 8 + (lane >> 3);
        if (x >= w || y >= h) continue;
        size_t p = static_cast<size_t>(y) * w + x;
        V3 delta{d_image[p], d_image[plane + p], d_image[2 * plane + p]};
        if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) continue;
        const float4 c = radiance[p * sample_count + i];
        Pcg rng;
        rng.seed(pixel_seed(sample_index + i, x, y));
        V3 org, dir;
        camera_ray(s.camera, x, y, w, h, rng, org, dir);
        replay_path<BLOCK, LDS, WIDE>(s, cfg, lds_nodes, st, rng, org, dir, V3{c.x, c.y, c.z}, delta, g, n_ext, n_shadow);
        n_samples += 1;
    }
    wave_add(stats + 0, n_ext);
    wave_add(stats + 1, n_shadow);
    wave_add(stats + 2, n_samples);
    if (in_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < n_grad; i += BLOCK) {
            float v = lds_grad[i];
            if (v != 0.0f) {
                if (i < mat_floats) atomicAdd(d_mat + i, v);
                else atomicAdd(d_light + (i - mat_floats), v);
            }
        }
    }
}

template <int BLOCK, int LDS>
hipError_t adjoint_impl(const SceneView& s, const sol_pt_config& cfg, uint32_t w, uint32_t h, uint64_t si,
                        uint32_t count, const float4* radiance, const float* d_image, uint32_t num_materials,
                        float* d_mat, float* d_light,
                        unsigned long long* counter, unsigned long long* stats, const TraversalScratch& sc, int grid,
                        bool wide, hipStream_t stream) {
    (void)hipGetLastError();
    if (wide)
        hipLaunchKernelGGL((adjoint_kernel<BLOCK, LDS, true>), dim3(grid), dim3(BLOCK), 0, stream, s, cfg, w, h, si,
                           count, radiance, d_image, num_materials, d_mat, d_light, counter, stats, sc);
    else
        hipLaunchKernelGGL((adjoint_kernel<BLOCK, LDS, false>), dim3(grid), dim3(BLOCK), 0, stream, s, cfg, w, h, si,
                           count, radiance, d_image, num_materials, d_mat, d_light, counter, stats, sc);
    return hipGetLastError();
}

} // namespace

hipError_t launch_adjoint(const SceneView& s, const sol_pt_config& cfg, uint32_t width, uint32_t height,
                          uint64_t sample_index, uint32_t sample_count, const float4* radiance, const float* d_image,
                          uint32_t num_materials,
                          float* d_mat, float* d_light, unsigned long long* work_counter, StatCounters stats,
                          const TraversalScratch& scratch, const LaunchShape& shape, hipStream_t stream) {
    switch (shape.block * 100 + shape.lds_stack) {
        case 6408: return adjoint_impl<64, 8>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 6416: return adjoint_impl<64, 16>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 6432: return adjoint_impl<64, 32>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 12808: return adjoint_impl<128, 8>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 12816: return adjoint_impl<128, 16>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 12832: return adjoint_impl<128, 32>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 25608: return adjoint_impl<256, 8>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 25616: return adjoint_impl<256, 16>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        case 25632: return adjoint_impl<256, 32>(s, cfg, width, height, sample_index, sample_count, radiance, d_image, num_materials, d_mat, d_light, work_counter, stats, scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream);
        default: return hipErrorInvalidValue;
    }
}

} // namespace solhip
