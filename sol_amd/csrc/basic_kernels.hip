// Batched intersection (Geometry::intersect_closest / intersect_any, src/triangle_mesh.cpp:64-117
// of sol) and the plain GPU megakernel (PathTracer::render + trace_path,
// src/algorithms/path_tracer.cpp:9-120), one thread per pixel looping over its samples.
// Both are persistent: the megakernel's waves fetch 64 pixels at a time from an atomic counter,
// the intersection kernel's waves take chunks of rays with the traversal kernel's work fetch.
#include <algorithm>

#include "launch.h"
#include "megapath.h"
#include "work.h"

namespace solhip {

namespace {

// Rays [0, n) of one launch (n < 2^32). Each wave takes a chunk of rays with the traversal
// kernel's work fetch (work.h: guided chunks of 64 to 1024 rays from eight part counters, or from
// one counter above 16M rays; 64 rays by rank, with no atomic, when the launch has no more than 64
// a wave) and traces the chunk 64 rays at a time, one per lane, each batch to its end. One
// counter handing out 64 rays per atomic (65,536 atomics on one address for 4M rays) was the
// bound of this kernel.
template <bool ANY, int BLOCK, int LDS, bool WIDE>
__global__ __launch_bounds__(BLOCK) void intersect_kernel(SceneView s, const sol_ray* __restrict__ rays, unsigned n,
                                                          sol_hit* __restrict__ hits, int32_t* __restrict__ occluded,
                                                          unsigned* counters, TraversalScratch sc) {
    __shared__ float4 lds_nodes[kTopLdsFloat4<WIDE>];
    __shared__ int lds_stack[LDS * BLOCK];
    stage_top_nodes<BLOCK, WIDE>(s, lds_nodes);
    Stack st = make_stack<BLOCK, LDS>(lds_stack, sc);
    const unsigned waves = gridDim.x * (BLOCK / 64);
    TraceWork work(n, BLOCK / 64);
    unsigned base = 0, end = 0;
    while (trace_fetch(counters, waves, work, base, end)) {
        for (unsigned first = base; first < end; first += 64) {
            const unsigned i = first + static_cast<unsigned>(__lane_id());
            if (i >= end) continue;
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            float4 a = rp[0], b = rp[1];
            V3 org{a.x, a.y, a.z}, dir{b.x, b.y, b.z};
            RayPre r = make_ray_pre(org, dir);
            HitRecord hit;
            // any-hit queries nearest child first: 1M-triangle soup 1942 -> 2080 Mrays/s, sphere field
            // 6704 -> 6980 (the megakernel's shadow rays keep the slot order: 5 % slower this way)
            bool found = trace_ray<ANY, BLOCK, LDS, WIDE, true>(s, lds_nodes, st, r, a.w, b.w, hit);
            if (ANY) {
                occluded[i] = found ? 1 : 0;
            } else {
                sol_hit h;
                h.prim = found ? hit.prim : -1;
                h.t = found ? hit.t : b.w;
                h.u = found ? hit.u : 0.0f;
                h.v = found ? hit.v : 0.0f;
                reinterpret_cast<float4*>(hits)[i] = make_float4(__int_as_float(h.prim), h.t, h.u, h.v);
            }
        }
    }
}

// Pixels are handed out in 8x8 tiles (one tile per wave) so that primary rays of a wave are
// coherent.
// WIDE: the 4-wide layout with leaf postponing (the trace kernel's traversal), 808 vs 685
// Msamples/s for the binary one on the 1080p sphere field. With whole paths in one kernel its
// registers bound its occupancy: three waves per SIMD need at most 168 VGPRs and the compiler is
// held to that (left free it takes 170-173 VGPRs and two waves, 20 % slower).
template <int BLOCK, int LDS, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void megakernel(SceneView s, sol_pt_config cfg, uint32_t w, uint32_t h,
                                                    uint64_t sample_index, uint32_t sample_count,
                                                    float* __restrict__ accum, unsigned long long* counter,
                                                    unsigned long long* stats, TraversalScratch sc) {
    __shared__ float4 lds_nodes[kTopLdsFloat4<WIDE>];
    __shared__ int lds_stack[LDS * BLOCK];
    stage_top_nodes<BLOCK, WIDE>(s, lds_nodes);
    Stack st = make_stack<BLOCK, LDS>(lds_stack, sc);
    const uint32_t tiles_x = (w + 7) / 8, tiles_y = (h + 7) / 8;
    const unsigned long long total = static_cast<unsigned long long>(tiles_x) * tiles_y * 64;
    const size_t plane = static_cast<size_t>(w) * h;
    unsigned n_ext = 0, n_shadow = 0, n_samples = 0;
    while (true) {
        unsigned long long base = wave_grab(counter, 64);
        if (base >= total) break;
        uint32_t tile = static_cast<uint32_t>(base / 64);
        uint32_t lane = __lane_id();
        uint32_t x = (tile % tiles_x) * 8 + (lane & 7);
        uint32_t y = (tile / tiles_x) * 8 + (lane >> 3);
        if (x >= w || y >= h) continue;
        V3 color{0.0f, 0.0f, 0.0f};
        for (uint32_t i = 0; i < sample_count; ++i) {
            Pcg rng;
            rng.seed(pixel_seed(sample_index + i, x, y));
            V3 org, dir;
            camera_ray(s.camera, x, y, w, h, rng, org, dir);
            color = color + trace_path<BLOCK, LDS, WIDE>(s, cfg, lds_nodes, st, rng, org, dir, n_ext, n_shadow);
        }
        n_samples += sample_count;
        size_t p = static_cast<size_t>(y) * w + x;
        accum[p] += color.x;
        accum[plane + p] += color.y;
        accum[2 * plane + p] += color.z;
    }
    wave_add(stats + 0, n_ext);
    wave_add(stats + 1, n_shadow);
    wave_add(stats + 2, n_samples);
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(256) void zero_f32_kernel(float* p, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256)
        p[i] = 0.0f;
}

// Launches of at most 2^31 rays, the kTraceParts work counters (kCounterStride words apart)
// cleared before each.
template <int BLOCK, int LDS>
hipError_t intersect_impl(const SceneView& s, const sol_ray* rays, uint64_t n, bool any_hit, sol_hit* hits,
                          int32_t* occluded, unsigned* counters, const TraversalScratch& sc, int grid, bool wide,
                          hipStream_t stream) {
    constexpr uint64_t kMaxRays = 1ull << 31;
    dim3 g(grid), b(BLOCK);
    for (uint64_t first = 0; first < n; first += kMaxRays) {
        const unsigned m = static_cast<unsigned>(std::min<uint64_t>(kMaxRays, n - first));
        const sol_ray* rr = rays + first;
        sol_hit* hh = hits ? hits + first : nullptr;
        int32_t* oo = occluded ? occluded + first : nullptr;
        hipError_t e = launch_zero_u64(reinterpret_cast<unsigned long long*>(counters), kIntersectCounterWords / 2, stream);
        if (e != hipSuccess) return e;
        (void)hipGetLastError();
        if (any_hit) {
            if (wide) hipLaunchKernelGGL((intersect_kernel<true, BLOCK, LDS, true>), g, b, 0, stream, s, rr, m, hh, oo, counters, sc);
            else hipLaunchKernelGGL((intersect_kernel<true, BLOCK, LDS, false>), g, b, 0, stream, s, rr, m, hh, oo, counters, sc);
        } else {
            if (wide) hipLaunchKernelGGL((intersect_kernel<false, BLOCK, LDS, true>), g, b, 0, stream, s, rr, m, hh, oo, counters, sc);
            else hipLaunchKernelGGL((intersect_kernel<false, BLOCK, LDS, false>), g, b, 0, stream, s, rr, m, hh, oo, counters, sc);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int BLOCK, int LDS>
hipError_t megakernel_impl(const SceneView& s, const sol_pt_config& cfg, uint32_t w, uint32_t h, uint64_t si,
                           uint32_t sc_count, float* accum, unsigned long long* counter, unsigned long long* stats,
                           const TraversalScratch& sc, int grid, bool wide, hipStream_t stream) {
    (void)hipGetLastError();
    if (wide)
        hipLaunchKernelGGL((megakernel<BLOCK, LDS, true>), dim3(grid), dim3(BLOCK), 0, stream, s, cfg, w, h, si, sc_count,
                           accum, counter, stats, sc);
    else
        hipLaunchKernelGGL((megakernel<BLOCK, LDS, false>), dim3(grid), dim3(BLOCK), 0, stream, s, cfg, w, h, si,
                           sc_count, accum, counter, stats, sc);
    return hipGetLastError();
}

} // namespace

#define SOLHIP_DISPATCH(FN, ...)                                                                   \
    switch (shape.block * 100 + shape.lds_stack) {                                                 \
        case 6408: return FN<64, 8>(__VA_ARGS__);                                                  \
        case 6416: return FN<64, 16>(__VA_ARGS__);                                                 \
        case 6432: return FN<64, 32>(__VA_ARGS__);                                                 \
        case 12808: return FN<128, 8>(__VA_ARGS__);                                                \
        case 12816: return FN<128, 16>(__VA_ARGS__);                                               \
        case 12832: return FN<128, 32>(__VA_ARGS__);                                               \
        case 25608: return FN<256, 8>(__VA_ARGS__);                                                \
        case 25616: return FN<256, 16>(__VA_ARGS__);                                               \
        case 25632: return FN<256, 32>(__VA_ARGS__);                                               \
        default: return hipErrorInvalidValue;                                                      \
    }

hipError_t launch_intersect(const SceneView& s, const sol_ray* rays, uint64_t n, bool any_hit, sol_hit* hits,
                            int32_t* occluded, unsigned* work_counters, const TraversalScratch& scratch,
                            const LaunchShape& shape, hipStream_t stream) {
    SOLHIP_DISPATCH(intersect_impl, s, rays, n, any_hit, hits, occluded, work_counters, scratch, shape.grid,
                    shape.wide && s.nodes4 != nullptr, stream)
}

hipError_t launch_megakernel(const SceneView& s, const sol_pt_config& cfg, uint32_t width, uint32_t height,
                             uint64_t sample_index, uint32_t sample_count, float* accum,
                             unsigned long long* work_counter, StatCounters stats, const TraversalScratch& scratch,
                             const LaunchShape& shape, hipStream_t stream) {
    SOLHIP_DISPATCH(megakernel_impl, s, cfg, width, height, sample_index, sample_count, accum, work_counter, stats,
                    scratch, shape.grid, shape.wide && s.nodes4 != nullptr, stream)
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t stream) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_zero_f32(float* p, size_t n, hipStream_t stream) {
    (void)hipGetLastError();
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

int persistent_grid(int compute_units, int block) {
    int per_cu = 2048 / block;
    return compute_units * per_cu;
}

} // namespace solhip
