// Wavefront kernels: extend (closest hit + hit decoding + emission + material sort), shade (one
// kernel per material queue: NEE, Russian roulette, BSDF sampling), shadow (any hit), regen
// (retire finished paths, start new ones from the global work queue), reduce (per-pixel sums in
// sample order). Queue appends are compacted per wave: a ballot gives each lane its rank, one
// lane reserves the wave's slots with a single atomic.
//
// The per-bounce arithmetic is the shared code of path.h, in the order of trace_path
// (src/algorithms/path_tracer.cpp:31-120 of sol): the emission of hit k is added in extend(k),
// the NEE contribution of vertex k in shadow(k), before extend(k+1).
#include "launch.h"
#include "path.h"
#include "traverse.h"
#include "wave.h"
#include "wave_launch.h"

namespace solhip {

namespace {

constexpr int kFrontBit = 1 << 30;

SOL_DEV unsigned long long wave_grab64(unsigned* counter, unsigned chunk) {
    unsigned base = 0;
    if (__lane_id() == 0) base = atomicAdd(counter, chunk);
    return __shfl(base, 0);
}

// Wave-compacted append: returns this lane's slot in the queue (or -1 if !pred).
SOL_DEV int wave_append(unsigned* count, bool pred) {
    unsigned long long mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, static_cast<unsigned>(__popcll(mask)));
    base = __shfl(base, leader);
    if (!pred) return -1;
    unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return static_cast<int>(base + static_cast<unsigned>(__popcll(below)));
}

SOL_DEV void wave_add_stat(unsigned long long* counter, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (__lane_id() == 0 && v) atomicAdd(counter, v);
}

template <bool MULTI>
SOL_DEV const SceneView& scene_of(const SceneView& single, const WaveTables& t, const WaveBuf& b, uint32_t slot,
                                  SceneView& tmp) {
    if (!MULTI) return single;
    tmp = t.scenes[b.scene[slot]];
    return tmp;
}

SOL_DEV int queue_of(int type) {
    return type == SOL_BSDF_DIFFUSE ? kQueueDiffuse : (type == SOL_BSDF_PHONG ? kQueuePhong : kQueueSpecular);
}

template <int BLOCK>
SOL_DEV Stack make_stack_w(int* lds_stack, const TraversalScratch& sc) {
    int gtid = blockIdx.x * BLOCK + threadIdx.x;
    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = sc.spill + gtid;
    st.spill_stride = sc.max_threads;
    st.capacity = 0;
    st.overflow = sc.overflow;
    return st;
}

// ---- extend ----------------------------------------------------------------------------------------

template <int BLOCK, int LDS, bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_extend(SceneView single, WaveTables tabs, WaveBuf b, sol_pt_config cfg,
                                                  int cur, TraversalScratch sc) {
    __shared__ float4 lds_nodes[kTopNodesMax * 4];
    __shared__ int lds_stack[LDS * BLOCK];
    if (!MULTI) load_top_nodes<BLOCK>(single, lds_nodes);
    Stack st = make_stack_w<BLOCK>(lds_stack, sc);
    st.capacity = LDS + sc.spill_entries;
    const unsigned count = b.counts[kCntExtend0 + cur];
    const float* const* q = b.eq[cur];
    unsigned processed = 0;
    while (true) {
        unsigned base = static_cast<unsigned>(wave_grab64(&b.counts[kCntWork + 0], 64));
        if (base >= count) break;
        unsigned i = base + __lane_id();
        bool valid = i < count;
        int dest = -1;  // 0..2 shade queue, 3 terminated
        uint32_t slot = 0;
        if (valid) {
            ++processed;
            slot = b.es[cur][i];
            V3 org{q[0][i], q[1][i], q[2][i]};
            V3 dir{q[3][i], q[4][i], q[5][i]};
            float tmin = q[6][i];
            SceneView tmp;
            const SceneView& s = scene_of<MULTI>(single, tabs, b, slot, tmp);
            RayPre r = make_ray_pre(org, dir);
            HitRecord hit;
            bool found = MULTI ? traverse<false, BLOCK, LDS>(s, s.nodes, st, r, tmin, FLT_MAX, hit)
                               : traverse<false, BLOCK, LDS>(s, lds_nodes, st, r, tmin, FLT_MAX, hit);
            if (!found) {
                dest = 3;
            } else {
                Surface sf = decode_hit(s, org, dir, hit.slot, hit.t, hit.u, hit.v);
                if (sf.light >= 0 && sf.front) {
                    V3 T{b.tx[slot], b.ty[slot], b.tz[slot]};
                    V3 C{b.cx[slot], b.cy[slot], b.cz[slot]};
                    V3 Le;
                    float w;
                    C = C + hit_emission(s, cfg.flags, sf, dir, hit.t, b.pdf[slot], 1.0f / static_cast<float>(s.num_lights),
                                         T, Le, w);
                    b.cx[slot] = C.x;
                    b.cy[slot] = C.y;
                    b.cz[slot] = C.z;
                }
                if (sf.mat < 0) {
                    dest = 3;
                } else {
                    b.hpx[slot] = sf.point.x; b.hpy[slot] = sf.point.y; b.hpz[slot] = sf.point.z;
                    b.hnx[slot] = sf.n.x; b.hny[slot] = sf.n.y; b.hnz[slot] = sf.n.z;
                    b.hfx[slot] = sf.face_n.x; b.hfy[slot] = sf.face_n.y; b.hfz[slot] = sf.face_n.z;
                    b.hdx[slot] = dir.x; b.hdy[slot] = dir.y; b.hdz[slot] = dir.z;
                    b.hmat[slot] = sf.mat | (sf.front ? kFrontBit : 0);
                    dest = queue_of(s.materials[sf.mat].type);
                }
            }
        }
        for (int qi = 0; qi < kNumShadeQueues; ++qi) {
            int at = wave_append(&b.counts[kCntShade0 + qi], dest == qi);
            if (at >= 0) b.mq[qi][at] = slot;
        }
        int at = wave_append(&b.counts[kCntRegen], dest == 3);
        if (at >= 0) b.rq[at] = slot;
    }
    wave_add_stat(&b.stats[0], processed);
}

// ---- shade -----------------------------------------------------------------------------------------

template <int QUEUE, bool MULTI>
__global__ __launch_bounds__(256) void k_shade(SceneView single, WaveTables tabs, WaveBuf b, sol_pt_config cfg,
                                               int nxt) {
    const unsigned count = b.counts[kCntShade0 + QUEUE];
    while (true) {
        unsigned base = static_cast<unsigned>(wave_grab64(&b.counts[kCntWork + 1 + QUEUE], 64));
        if (base >= count) break;
        unsigned i = base + __lane_id();
        bool valid = i < count;
        uint32_t slot = 0;
        bool shoot = false, cont = false;
        V3 point{0, 0, 0}, sdir{0, 0, 0}, contrib{0, 0, 0}, ndir{0, 0, 0};
        if (valid) {
            slot = b.mq[QUEUE][i];
            SceneView tmp;
            const SceneView& s = scene_of<MULTI>(single, tabs, b, slot, tmp);
            Surface sf;
            sf.point = V3{b.hpx[slot], b.hpy[slot], b.hpz[slot]};
            sf.n = V3{b.hnx[slot], b.hny[slot], b.hnz[slot]};
            sf.face_n = V3{b.hfx[slot], b.hfy[slot], b.hfz[slot]};
            int hm = b.hmat[slot];
            sf.front = (hm & kFrontBit) != 0;
            sf.mat = hm & ~kFrontBit;
            sf.light = -1;
            point = sf.point;
            V3 dir{b.hdx[slot], b.hdy[slot], b.hdz[slot]};
            V3 out = -dir;
            const sol_material m = s.materials[sf.mat];
            V3 T{b.tx[slot], b.ty[slot], b.tz[slot]};
            Pcg rng;
            rng.state = b.rng[slot];
            uint32_t depth = b.depth[slot];
            bool skip_nee = (cfg.flags & SOL_PT_DISABLE_NEE) || is_specular(m.type);
            if (!skip_nee) {
                NeeSample ns;
                shoot = nee_sample(s, cfg.flags, sf, m, out, T, rng, 1.0f / static_cast<float>(s.num_lights), ns);
                sdir = ns.shadow_dir;
                contrib = ns.contrib;
            }
            float q;
            BsdfSample bs;
            if (russian_roulette(cfg, depth, T, rng, q) && bsdf_sample(m, sf, ortho_basis(sf.n), out, rng, bs)) {
                T = T * (bs.color * (bs.cos / (bs.pdf * q)));
                b.tx[slot] = T.x; b.ty[slot] = T.y; b.tz[slot] = T.z;
                b.pdf[slot] = skip_nee ? 0.0f : bs.pdf;
                b.depth[slot] = depth + 1;
                cont = depth + 1 < cfg.max_path_len;
                ndir = bs.in;
            }
            b.rng[slot] = rng.state;
        }
        int at = wave_append(&b.counts[kCntShadow], shoot);
        if (at >= 0) {
            b.sq[0][at] = point.x; b.sq[1][at] = point.y; b.sq[2][at] = point.z;
            b.sq[3][at] = sdir.x; b.sq[4][at] = sdir.y; b.sq[5][at] = sdir.z;
            b.sq[6][at] = contrib.x; b.sq[7][at] = contrib.y; b.sq[8][at] = contrib.z;
            b.ss[at] = slot;
        }
        at = wave_append(&b.counts[kCntExtend0 + nxt], cont);
        if (at >= 0) {
            float* const* e = b.eq[nxt];
            e[0][at] = point.x; e[1][at] = point.y; e[2][at] = point.z;
            e[3][at] = ndir.x; e[4][at] = ndir.y; e[5][at] = ndir.z;
            e[6][at] = cfg.ray_offset;
            b.es[nxt][at] = slot;
        }
        at = wave_append(&b.counts[kCntRegen], valid && !cont);
        if (at >= 0) b.rq[at] = slot;
    }
}

// ---- shadow ----------------------------------------------------------------------------------------

template <int BLOCK, int LDS, bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_shadow(SceneView single, WaveTables tabs, WaveBuf b, sol_pt_config cfg,
                                                  TraversalScratch sc) {
    __shared__ float4 lds_nodes[kTopNodesMax * 4];
    __shared__ int lds_stack[LDS * BLOCK];
    if (!MULTI) load_top_nodes<BLOCK>(single, lds_nodes);
    Stack st = make_stack_w<BLOCK>(lds_stack, sc);
    st.capacity = LDS + sc.spill_entries;
    const unsigned count = b.counts[kCntShadow];
    unsigned processed = 0;
    while (true) {
        unsigned base = static_cast<unsigned>(wave_grab64(&b.counts[kCntWork + 4], 64));
        if (base >= count) break;
        unsigned i = base + __lane_id();
        if (i >= count) continue;
        ++processed;
        uint32_t slot = b.ss[i];
        SceneView tmp;
        const SceneView& s = scene_of<MULTI>(single, tabs, b, slot, tmp);
        V3 org{b.sq[0][i], b.sq[1][i], b.sq[2][i]};
        V3 dir{b.sq[3][i], b.sq[4][i], b.sq[5][i]};
        RayPre r = make_ray_pre(org, dir);
        HitRecord unused;
        bool occluded = MULTI ? traverse<true, BLOCK, LDS>(s, s.nodes, st, r, cfg.ray_offset, 1.0f - cfg.ray_offset, unused)
                              : traverse<true, BLOCK, LDS>(s, lds_nodes, st, r, cfg.ray_offset, 1.0f - cfg.ray_offset, unused);
        if (!occluded) {
            b.cx[slot] = b.cx[slot] + b.sq[6][i];
            b.cy[slot] = b.cy[slot] + b.sq[7][i];
            b.cz[slot] = b.cz[slot] + b.sq[8][i];
        }
    }
    wave_add_stat(&b.stats[1], processed);
}

// ---- regen -----------------------------------------------------------------------------------------

// Tile owning a global item: the last tile whose first item is <= item.
SOL_DEV uint32_t find_tile(const WaveTables& t, unsigned long long item) {
    uint32_t lo = 0, hi = t.num_tiles;   // tile_item0[lo] <= item < tile_item0[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (t.tile_item0[mid] <= item) lo = mid; else hi = mid;
    }
    return lo;
}

template <bool MULTI>
__global__ __launch_bounds__(256) void k_regen(WaveTables tabs, WaveBuf b, sol_pt_config cfg, int nxt, int initial) {
    const unsigned count = initial ? b.slots : b.counts[kCntRegen];
    const unsigned long long chunk_begin = b.items[2], chunk_end = b.items[1];
    unsigned started = 0;
    while (true) {
        unsigned base = static_cast<unsigned>(wave_grab64(&b.counts[kCntWork + 5], 64));
        if (base >= count) break;
        unsigned i = base + __lane_id();
        bool valid = i < count;
        uint32_t slot = valid ? (initial ? i : b.rq[i]) : 0;
        if (valid && !initial) {
            unsigned long long k = b.item[slot] - chunk_begin;
            b.samples[k] = b.cx[slot];
            b.samples[b.sample_capacity + k] = b.cy[slot];
            b.samples[2 * b.sample_capacity + k] = b.cz[slot];
        }
        // one new item per lane that retired a path
        unsigned long long mask = __ballot(valid);
        unsigned long long first = 0;
        int lane = __lane_id();
        int leader = __ffsll(static_cast<long long>(mask)) - 1;
        if (mask && lane == leader) first = atomicAdd(&b.items[0], static_cast<unsigned long long>(__popcll(mask)));
        first = __shfl(first, leader < 0 ? 0 : leader);
        unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
        unsigned long long item = first + __popcll(below);
        bool start = valid && item < chunk_end;
        V3 org{0, 0, 0}, dir{0, 0, 0};
        if (start) {
            ++started;
            uint32_t tile = find_tile(tabs, item);
            uint32_t tw = tabs.tile_w[tile], th = tabs.tile_h[tile];
            unsigned long long local = item - tabs.tile_item0[tile];
            unsigned long long per_sample = static_cast<unsigned long long>(tw) * th;
            unsigned long long s = local / per_sample;
            uint32_t pix = static_cast<uint32_t>(local - s * per_sample);
            uint32_t job = tabs.tile_job[tile];
            uint32_t x = tabs.tile_x0[tile] + pix % tw;
            uint32_t y = tabs.tile_y0[tile] + pix / tw;
            uint32_t sc = tabs.job_scene[job];
            const sol_camera cam = tabs.scenes[sc].camera;
            Pcg rng;
            rng.seed(pixel_seed(tabs.tile_sample0[tile] + s, x, y));
            camera_ray(cam, x, y, tabs.job_w[job], tabs.job_h[job], rng, org, dir);
            b.rng[slot] = rng.state;
            b.item[slot] = item;
            b.tx[slot] = 1.0f; b.ty[slot] = 1.0f; b.tz[slot] = 1.0f;
            b.cx[slot] = 0.0f; b.cy[slot] = 0.0f; b.cz[slot] = 0.0f;
            b.pdf[slot] = 0.0f;
            b.depth[slot] = 0;
            if (MULTI) b.scene[slot] = sc;
        }
        int at = wave_append(&b.counts[kCntExtend0 + nxt], start);
        if (at >= 0) {
            float* const* e = b.eq[nxt];
            e[0][at] = org.x; e[1][at] = org.y; e[2][at] = org.z;
            e[3][at] = dir.x; e[4][at] = dir.y; e[5][at] = dir.z;
            e[6][at] = 0.0f;
            b.es[nxt][at] = slot;
        }
    }
    wave_add_stat(&b.stats[2], started);
}

// Zeroes the counters the next iteration appends to (kernel, not a memset: it is captured in
// graphs).
__global__ void k_reset(WaveBuf b, int nxt) {
    int i = threadIdx.x;
    if (i < kNumWaveCounters && i != kCntExtend0 + (1 - nxt)) b.counts[i] = 0;
}

__global__ void k_set_chunk(WaveBuf b, unsigned long long begin, unsigned long long end) {
    if (threadIdx.x == 0) {
        b.items[0] = begin;
        b.items[1] = end;
        b.items[2] = begin;
    }
    if (threadIdx.x < kNumWaveCounters) b.counts[threadIdx.x] = 0;
}

// Per pixel of the chunk's tiles: sum the samples in sample order, add to the image once.
__global__ __launch_bounds__(256) void k_reduce(WaveTables tabs, WaveBuf b, uint32_t first_tile,
                                                unsigned long long chunk_begin) {
    uint32_t tile = first_tile + blockIdx.x;
    uint32_t tw = tabs.tile_w[tile], th = tabs.tile_h[tile];
    uint32_t per_sample = tw * th;
    uint32_t pix = blockIdx.y * 256 + threadIdx.x;
    if (pix >= per_sample) return;
    unsigned long long k0 = tabs.tile_item0[tile] - chunk_begin;
    unsigned long long spp = (tabs.tile_item0[tile + 1] - tabs.tile_item0[tile]) / per_sample;
    float r = 0.0f, g = 0.0f, bl = 0.0f;
    for (unsigned long long s = 0; s < spp; ++s) {
        unsigned long long k = k0 + s * per_sample + pix;
        r = r + b.samples[k];
        g = g + b.samples[b.sample_capacity + k];
        bl = bl + b.samples[2 * b.sample_capacity + k];
    }
    uint32_t job = tabs.tile_job[tile];
    uint32_t w = tabs.job_w[job], h = tabs.job_h[job];
    uint32_t x = tabs.tile_x0[tile] + pix % tw, y = tabs.tile_y0[tile] + pix / tw;
    float* accum = tabs.job_accum[job];
    size_t plane = static_cast<size_t>(w) * h, p = static_cast<size_t>(y) * w + x;
    accum[p] += r;
    accum[plane + p] += g;
    accum[2 * plane + p] += bl;
}

template <int BLOCK, int LDS, bool MULTI>
void launch_extend_t(const WaveLaunch& L, int cur) {
    hipLaunchKernelGGL((k_extend<BLOCK, LDS, MULTI>), dim3(L.grid_traverse), dim3(BLOCK), 0, L.stream, L.single, L.tabs,
                       L.buf, L.cfg, cur, L.scratch);
}

template <int BLOCK, int LDS, bool MULTI>
void launch_shadow_t(const WaveLaunch& L) {
    hipLaunchKernelGGL((k_shadow<BLOCK, LDS, MULTI>), dim3(L.grid_traverse), dim3(BLOCK), 0, L.stream, L.single, L.tabs,
                       L.buf, L.cfg, L.scratch);
}

} // namespace

#define SOLHIP_WAVE_DISPATCH(FN, MULTI, ...)                                                         \
    switch (L.block * 100 + L.lds_stack) {                                                           \
        case 6408: FN<64, 8, MULTI>(__VA_ARGS__); break;                                             \
        case 6416: FN<64, 16, MULTI>(__VA_ARGS__); break;                                            \
        case 6432: FN<64, 32, MULTI>(__VA_ARGS__); break;                                            \
        case 12808: FN<128, 8, MULTI>(__VA_ARGS__); break;                                           \
        case 12816: FN<128, 16, MULTI>(__VA_ARGS__); break;                                          \
        case 12832: FN<128, 32, MULTI>(__VA_ARGS__); break;                                          \
        case 25608: FN<256, 8, MULTI>(__VA_ARGS__); break;                                           \
        case 25616: FN<256, 16, MULTI>(__VA_ARGS__); break;                                          \
        case 25632: FN<256, 32, MULTI>(__VA_ARGS__); break;                                          \
        default: return hipErrorInvalidValue;                                                        \
    }

hipError_t wave_launch_extend(const WaveLaunch& L, int cur) {
    if (L.multi) { SOLHIP_WAVE_DISPATCH(launch_extend_t, true, L, cur) }
    else { SOLHIP_WAVE_DISPATCH(launch_extend_t, false, L, cur) }
    return hipGetLastError();
}

hipError_t wave_launch_shadow(const WaveLaunch& L) {
    if (L.multi) { SOLHIP_WAVE_DISPATCH(launch_shadow_t, true, L) }
    else { SOLHIP_WAVE_DISPATCH(launch_shadow_t, false, L) }
    return hipGetLastError();
}

hipError_t wave_launch_shade(const WaveLaunch& L, int queue, int nxt) {
    dim3 g(L.grid_shade), bs(256);
    if (L.multi) {
        if (queue == 0) hipLaunchKernelGGL((k_shade<0, true>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 1) hipLaunchKernelGGL((k_shade<1, true>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 2) hipLaunchKernelGGL((k_shade<2, true>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
    } else {
        if (queue == 0) hipLaunchKernelGGL((k_shade<0, false>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 1) hipLaunchKernelGGL((k_shade<1, false>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 2) hipLaunchKernelGGL((k_shade<2, false>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
    }
    return hipGetLastError();
}

hipError_t wave_launch_regen(const WaveLaunch& L, int nxt, bool initial) {
    if (L.multi)
        hipLaunchKernelGGL((k_regen<true>), dim3(L.grid_shade), dim3(256), 0, L.stream, L.tabs, L.buf, L.cfg, nxt,
                           initial ? 1 : 0);
    else
        hipLaunchKernelGGL((k_regen<false>), dim3(L.grid_shade), dim3(256), 0, L.stream, L.tabs, L.buf, L.cfg, nxt,
                           initial ? 1 : 0);
    return hipGetLastError();
}

hipError_t wave_launch_reset(const WaveLaunch& L, int nxt) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, L.stream, L.buf, nxt);
    return hipGetLastError();
}

hipError_t wave_launch_set_chunk(const WaveLaunch& L, unsigned long long begin, unsigned long long end) {
    hipLaunchKernelGGL(k_set_chunk, dim3(1), dim3(64), 0, L.stream, L.buf, begin, end);
    return hipGetLastError();
}

hipError_t wave_launch_reduce(const WaveLaunch& L, uint32_t first_tile, uint32_t num_tiles, uint32_t max_tile_pixels,
                              unsigned long long chunk_begin) {
    dim3 g(num_tiles, (max_tile_pixels + 255) / 256);
    hipLaunchKernelGGL(k_reduce, g, dim3(256), 0, L.stream, L.tabs, L.buf, first_tile, chunk_begin);
    return hipGetLastError();
}

} // namespace solhip
