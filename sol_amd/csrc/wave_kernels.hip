// Wavefront kernels.
//   k_trace  closest hit of the extension rays (hit written in place) and any hit of the previous
//            bounce's shadow rays (unoccluded contribution added to the path or its sample);
//            traversal only, work fetched by waves in shrinking chunks (guided self-scheduling)
//            from one counter, or from eight part counters or by rank in smaller launches.
//   k_sort   material queues from each hit triangle's kind (packed into the hit by the traversal
//            kernel); a path that missed ends there (its radiance goes to its sample entry, unless
//            the chunk runs in place) and its slot starts the next work item.
//   k_shade  one kernel per material queue: hit decoding, emission, NEE, Russian roulette, BSDF
//            sampling.
//   k_start  fills every slot at the beginning of a chunk of work.
//   k_reduce per-pixel sums of the chunk's samples in sample order.
// Appends are compacted per workgroup: ballots give each lane its rank in its wave, the wave
// counts are scanned in LDS and one lane per queue reserves the workgroup's entries with a single
// device atomic (device-scope atomics are slow on a multi-XCD part: few of them, never per lane).
//
// The per-bounce arithmetic is the shared code of path.h, in the order of trace_path
// (src/algorithms/path_tracer.cpp:31-120 of sol): NEE of vertex k is added by trace(k+1), before
// the emission of hit k+1 is added by shade(k+1).
#include <algorithm>

#include "launch.h"
#include "path.h"
#include "traverse.h"
#include "wave.h"
#include "wave_launch.h"
#include "work.h"

namespace solhip {

namespace {

constexpr uint32_t kFrontBit = 1u << 30;
constexpr uint32_t kSampleTarget = 0xFFFFFFFFu;
constexpr int kBlock = kWaveBlock;
constexpr uint32_t kHole = 0xFFFFFFFFu;     // slot field of an unused extension entry
constexpr uint32_t kMissWord = 0xFFFFFFFFu;  // first word of a miss's hit
constexpr int kSortItems = kWaveSortItems;  // entries per lane and round of the sort
constexpr int kWaves = kBlock / 64;

SOL_DEV uint32_t as_u(float f) { return __float_as_uint(f); }
SOL_DEV float as_f(uint32_t u) { return __uint_as_float(u); }
SOL_DEV float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
SOL_DEV V3 xyz(float4 v) { return V3{v.x, v.y, v.z}; }
SOL_DEV unsigned long long item_of(float4 p2) {
    return static_cast<unsigned long long>(as_u(p2.z)) | (static_cast<unsigned long long>(as_u(p2.w)) << 32);
}

// Adds the wave's v to statistic `which` (the workgroup's copy of it, see kStatSlots).
SOL_DEV void wave_add_stat(unsigned long long* stats, int which, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (__lane_id() == 0 && v) atomicAdd(&stats[(blockIdx.x % kStatSlots) * kStatStride + which], v);
}

// Workgroup-wide queue reservations of one round (one entry per thread). Each thread belongs to
// at most one of NQ plain queues (`dest`), may continue its path (`cont`: one extension entry) or
// retire it (`retire`). Retired paths take the next work items in thread order; every retired
// path also gets an extension entry after the round's `cont` entries, used for the new path if
// an item was left and marked as a hole (slot kHole) otherwise - holes only appear when the
// chunk runs out of items. Work items come from the workgroup's segment first; when it runs dry
// the other segments are tried in turn, so no item is stranded while any path retires. All
// reservations are issued at once by different lanes: two barriers and one device atomic per
// non-empty queue per round. The LDS block is double-buffered by round parity, so no third
// barrier is needed before the next round overwrites it.
template <int NQ>
struct RoundLds {
    unsigned wc[NQ + 2][kWaves];    // per-wave counts: plain queues, cont, retire
    unsigned pre[NQ + 2][kWaves];   // their exclusive prefix sums over the waves
    unsigned base[NQ];
    unsigned ext_base;
    unsigned n_cont;
    unsigned n_pieces;              // work items granted: n_pieces ranges [piece_base, + piece_count)
    unsigned piece_count[kMaxSegments];
    unsigned long long piece_base[kMaxSegments];
};

struct RoundPlan {
    int pos;                     // position in the plain queue (in the segment), -1 if none
    int ext_pos;                 // position in the extension queue (in the segment), -1 if none
    bool start;                  // the retired path got a new work item
    unsigned long long item;     // that item
};

// Reserves up to `need` work items, own segment first (items[item_next(s)] counts up to
// items[item_end(s)]), then the others of the nseg segments. Counters only grow, so a relaxed
// read at or past the end means the segment is exhausted and it is skipped without an atomic.
template <typename Lds>
SOL_DEV void grant_items(Lds& lds, unsigned long long* items, int seg, int nseg, unsigned need) {
    unsigned n = 0;
    for (int k = 0; k < nseg && need > 0; ++k) {
        int sk = (seg + k) & (nseg - 1);
        const unsigned long long end = items[item_end(sk)];
        if (k > 0 && __hip_atomic_load(&items[item_next(sk)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= end)
            continue;
        unsigned long long base = atomicAdd(&items[item_next(sk)], static_cast<unsigned long long>(need));
        if (base < end) {
            unsigned got = static_cast<unsigned>(min(static_cast<unsigned long long>(need), end - base));
            lds.piece_base[n] = base;
            lds.piece_count[n] = got;
            ++n;
            need -= got;
        }
    }
    lds.n_pieces = n;
}

template <int NQ>
SOL_DEV RoundPlan plan_round(RoundLds<NQ>& lds, unsigned* const* qcounters, unsigned* ext_counter,
                             unsigned long long* items, int seg, int nseg, int dest, bool cont, bool retire) {
    const int wave = threadIdx.x >> 6;
    const bool lane0 = __lane_id() == 0;
    const unsigned long long below = lanes_below();
    unsigned rank = 0;
#pragma unroll
    for (int d = 0; d < NQ; ++d) {
        unsigned long long m = __ballot(dest == d);
        if (lane0) lds.wc[d][wave] = static_cast<unsigned>(__popcll(m));
        if (dest == d) rank = static_cast<unsigned>(__popcll(m & below));
    }
    unsigned long long mc = __ballot(cont), mr = __ballot(retire);
    if (lane0) {
        lds.wc[NQ][wave] = static_cast<unsigned>(__popcll(mc));
        lds.wc[NQ + 1][wave] = static_cast<unsigned>(__popcll(mr));
    }
    unsigned rank_c = static_cast<unsigned>(__popcll(mc & below));
    unsigned rank_r = static_cast<unsigned>(__popcll(mr & below));
    __syncthreads();
    if (threadIdx.x < NQ + 2) {
        const int d = threadIdx.x;
        unsigned run = 0;
        for (int w = 0; w < kWaves; ++w) {
            lds.pre[d][w] = run;
            run += lds.wc[d][w];
        }
        if (d < NQ) {
            lds.base[d] = run ? atomicAdd(qcounters[d], run) : 0u;
        } else if (d == NQ) {
            // extension entries: the cont lanes, then one per retired lane
            unsigned n_ret = 0;
            for (int w = 0; w < kWaves; ++w) n_ret += lds.wc[NQ + 1][w];
            unsigned n = run + n_ret;
            lds.ext_base = n ? atomicAdd(ext_counter, n) : 0u;
            lds.n_cont = run;
        } else {
            grant_items(lds, items, seg, nseg, run);
        }
    }
    __syncthreads();
    RoundPlan p;
    p.pos = dest >= 0 ? static_cast<int>(lds.base[dest] + lds.pre[dest][wave] + rank) : -1;
    const unsigned ridx = lds.pre[NQ + 1][wave] + rank_r;
    p.start = false;
    p.item = 0;
    if (retire) {
        unsigned off = ridx;
        const unsigned np = lds.n_pieces;
        for (unsigned k = 0; k < np; ++k) {
            unsigned c = lds.piece_count[k];
            if (off < c) {
                p.start = true;
                p.item = lds.piece_base[k] + off;
                break;
            }
            off -= c;
        }
    }
    p.ext_pos = cont ? static_cast<int>(lds.ext_base + lds.pre[NQ][wave] + rank_c)
                     : (retire ? static_cast<int>(lds.ext_base + lds.n_cont + ridx) : -1);
    return p;
}

// A round of K entries per lane (K x kBlock entries, entry k of lane t at k x kBlock + t): the
// same reservations as plan_round - plain queues by `dest`, an extension entry and a work item per
// retiring entry - with one device atomic per non-empty queue for all K x kBlock entries, so a
// latency-bound pass pays its barriers and atomics once per K entries. Positions follow the order
// (k, wave, lane). No entry continues a path (the sort's case).
template <int NQ, int K>
struct RoundLdsK {
    unsigned wc[K][NQ + 1][kWaves];    // per-wave counts: plain queues, retire
    unsigned pre[K][NQ + 1][kWaves];   // their exclusive prefix sums over (k, wave)
    unsigned base[NQ];
    unsigned ext_base;
    unsigned n_pieces;
    unsigned piece_count[kMaxSegments];
    unsigned long long piece_base[kMaxSegments];
};

template <int NQ, int K>
SOL_DEV void plan_round_k(RoundLdsK<NQ, K>& lds, unsigned* const* qcounters, unsigned* ext_counter,
                          unsigned long long* items, int seg, int nseg, const int (&dest)[K],
                          const bool (&retire)[K], RoundPlan (&plan)[K]) {
    const int wave = threadIdx.x >> 6;
    const bool lane0 = __lane_id() == 0;
    const unsigned long long below = lanes_below();
    unsigned rank[K], rank_r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        rank[k] = 0;
#pragma unroll
        for (int d = 0; d < NQ; ++d) {
            const unsigned long long m = __ballot(dest[k] == d);
            if (lane0) lds.wc[k][d][wave] = static_cast<unsigned>(__popcll(m));
            if (dest[k] == d) rank[k] = static_cast<unsigned>(__popcll(m & below));
        }
        const unsigned long long mr = __ballot(retire[k]);
        if (lane0) lds.wc[k][NQ][wave] = static_cast<unsigned>(__popcll(mr));
        rank_r[k] = static_cast<unsigned>(__popcll(mr & below));
    }
    __syncthreads();
    if (threadIdx.x < NQ + 1) {
        const int d = threadIdx.x;
        unsigned run = 0;
        for (int k = 0; k < K; ++k)
            for (int w = 0; w < kWaves; ++w) {
                lds.pre[k][d][w] = run;
                run += lds.wc[k][d][w];
            }
        if (d < NQ) {
            lds.base[d] = run ? atomicAdd(qcounters[d], run) : 0u;
        } else {
            lds.ext_base = run ? atomicAdd(ext_counter, run) : 0u;
            grant_items(lds, items, seg, nseg, run);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        RoundPlan& p = plan[k];
        p.pos = dest[k] >= 0 ? static_cast<int>(lds.base[dest[k]] + lds.pre[k][dest[k]][wave] + rank[k]) : -1;
        p.start = false;
        p.item = 0;
        p.ext_pos = -1;
        if (retire[k]) {
            unsigned off = lds.pre[k][NQ][wave] + rank_r[k];
            p.ext_pos = static_cast<int>(lds.ext_base + off);
            const unsigned np = lds.n_pieces;
            for (unsigned q = 0; q < np; ++q) {
                const unsigned c = lds.piece_count[q];
                if (off < c) {
                    p.start = true;
                    p.item = lds.piece_base[q] + off;
                    break;
                }
                off -= c;
            }
        }
    }
}

// Whether any segment still has work items to hand out. Item counters only grow, so once this is
// false in a launch it stays false for the rest of it: paths that end then reserve no extension
// entry (they would only become holes).
SOL_DEV bool items_left(const WaveBuf& b) {
    bool left = false;
    for (unsigned k = 0; k < b.nseg; ++k)
        left = left || __hip_atomic_load(&b.items[item_next(static_cast<int>(k))], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) < b.items[item_end(static_cast<int>(k))];
    return left;
}

// A retired lane whose extension entry stays unused marks it as a hole (trace and sort skip it).
SOL_DEV void write_hole(const WaveBuf& b, int nxt, unsigned entry) {
    *b.ext(nxt, entry, 1) = make_float4(0.0f, 0.0f, 0.0f, as_f(kHole));
}

// The segment this workgroup appends to (grids are multiples of kMaxSegments, so of nseg).
SOL_DEV int my_segment(const WaveBuf& b) { return static_cast<int>(blockIdx.x & (b.nseg - 1)); }

// A queue read as the concatenation of its nseg segments: entry(i) is the storage index of the
// i-th entry. Segments past nseg count as empty.
struct SegQueue {
    unsigned pre[kMaxSegments + 1];
    unsigned cap;

    SOL_DEV SegQueue(const WaveBuf& b, int queue, unsigned capacity) : cap(capacity) {
        pre[0] = 0;
#pragma unroll
        for (int k = 0; k < kMaxSegments; ++k)
            pre[k + 1] = pre[k] + (k < static_cast<int>(b.nseg) ? b.counts[qcount(queue, k)] : 0u);
    }
    SOL_DEV unsigned size() const { return pre[kMaxSegments]; }
    SOL_DEV unsigned entry(unsigned i) const {
        unsigned seg = 0, start = 0;
#pragma unroll
        for (int k = 1; k < kMaxSegments; ++k) {
            bool ge = i >= pre[k];
            seg = ge ? static_cast<unsigned>(k) : seg;
            start = ge ? pre[k] : start;
        }
        return seg * cap + (i - start);
    }
};

template <bool MULTI>
SOL_DEV const SceneView& scene_of(const SceneView& single, const WaveTables& t, const WaveBuf& b, uint32_t slot,
                                  SceneView& tmp) {
    if (!MULTI) return single;
    tmp = t.scenes[as_u(b.rec(slot)->w)];
    return tmp;
}


// ---- work items ------------------------------------------------------------------------------------

SOL_DEV uint32_t find_job(const WaveTables& t, unsigned long long item) {
    uint32_t lo = 0, hi = t.num_jobs;   // job_item0[lo] <= item < job_item0[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (t.job_item0[mid] <= item) lo = mid; else hi = mid;
    }
    return lo;
}

// Position (dx, dy) in a tw x th tile of its pixel number pix: pixels are numbered in bands of 8
// rows (the last band may be lower), column by column within a band, so that 64 consecutive
// pixels - the primary rays of one wave - cover an 8 x 8 block rather than two rows of a 32-wide
// tile (1080p frame 52.9 -> 52.6 ms, trace 41.1 -> 40.7 ms, images identical).
SOL_DEV void tile_pixel(uint32_t pix, uint32_t tw, uint32_t th, uint32_t& dx, uint32_t& dy) {
    const uint32_t band_pixels = 8u * tw;
    const uint32_t band = pix / band_pixels;
    const uint32_t q = pix - band * band_pixels;
    const uint32_t hb = min(8u, th - band * 8u);
    dx = q / hb;
    dy = band * 8u + (q - dx * hb);
}

// Pixel and sample of a work item (the order of WaveTables). A job holds fewer than 2^32 items.
SOL_DEV void decode_item(const WaveTables& t, unsigned long long item, uint32_t& job, uint32_t& x, uint32_t& y,
                         unsigned long long& sample) {
    job = find_job(t, item);
    uint32_t local = static_cast<uint32_t>(item - t.job_item0[job]);
    uint32_t w = t.job_w[job], h = t.job_h[job], tw = t.job_tw[job], th = t.job_th[job], spp = t.job_spp[job];
    uint32_t bands = (h + th - 1) / th;
    uint32_t band_items = w * th * spp;
    uint32_t band = min(local / band_items, bands - 1);
    uint32_t q = local - band * band_items;
    uint32_t bh = min(th, h - band * th);
    uint32_t tiles_x = (w + tw - 1) / tw;
    uint32_t tile_items = tw * bh * spp;
    uint32_t tile = min(q / tile_items, tiles_x - 1);
    uint32_t r = q - tile * tile_items;
    uint32_t tw2 = min(tw, w - tile * tw);
    uint32_t per_sample = tw2 * bh;
    uint32_t s = r / per_sample;
    uint32_t pix = r - s * per_sample;
    uint32_t dx, dy;
    tile_pixel(pix, tw2, bh, dx, dy);
    x = tile * tw + dx;
    y = band * th + dy;
    sample = t.job_sample0[job] + s;
}

// Path start of one work item: seed, camera ray, the PCG state after the camera's draws.
struct PathStart {
    V3 org, dir;
    unsigned long long rng;
    uint32_t scene;
};

SOL_DEV PathStart start_path(const WaveTables& t, unsigned long long item) {
    uint32_t job, x, y;
    unsigned long long sample;
    decode_item(t, item, job, x, y, sample);
    PathStart ps;
    ps.scene = t.job_scene[job];
    const sol_camera cam = t.scenes[ps.scene].camera;
    Pcg rng;
    rng.seed(pixel_seed(sample, x, y));
    camera_ray(cam, x, y, t.job_w[job], t.job_h[job], rng, ps.org, ps.dir);
    ps.rng = rng.state;
    return ps;
}

SOL_DEV float4 rng_item_word(unsigned long long rng, unsigned long long item) {
    return make_float4(as_f(static_cast<uint32_t>(rng)), as_f(static_cast<uint32_t>(rng >> 32)),
                       as_f(static_cast<uint32_t>(item)), as_f(static_cast<uint32_t>(item >> 32)));
}
SOL_DEV unsigned long long rng_of(float4 w) {
    return static_cast<unsigned long long>(as_u(w.x)) | (static_cast<unsigned long long>(as_u(w.y)) << 32);
}

// A path continues with the extension entry `a` of queue q: origin and depth (depth 0, a new
// path, is traced from tmin 0, deeper rays from the configured offset), direction and slot,
// throughput and the pdf of the bounce that chose the direction, PCG state and work item. At
// depth 0 the throughput is 1 and the pdf 0, so that word is neither written nor read.
SOL_DEV void write_ext(const WaveBuf& b, int q, uint32_t a, V3 org, uint32_t depth, V3 dir, uint32_t slot, V3 T,
                       float pdf, unsigned long long rng, unsigned long long item) {
    *b.ext(q, a, 0) = f4(org, as_f(depth));
    *b.ext(q, a, 1) = f4(dir, as_f(slot));
    if (depth != 0) *b.ext(q, a, 2) = f4(T, pdf);
    *b.ext(q, a, 3) = rng_item_word(rng, item);
}

// Throughput and previous pdf of the path of extension entry `a` of queue q at `depth`.
SOL_DEV float4 ext_throughput(const WaveBuf& b, int q, uint32_t a, uint32_t depth) {
    return depth == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : *b.ext(q, a, 2);
}

// A new path for work item `item` in `slot`: its slot record (no radiance yet, its scene) and its
// first extension entry.
SOL_DEV void init_path(const WaveTables& t, const WaveBuf& b, unsigned long long item, uint32_t slot, int q,
                       uint32_t a) {
    const PathStart ps = start_path(t, item);
    *b.rec(slot) = make_float4(0.0f, 0.0f, 0.0f, as_f(ps.scene));
    write_ext(b, q, a, ps.org, 0u, ps.dir, slot, V3{1.0f, 1.0f, 1.0f}, 0.0f, ps.rng, item);
}

// The radiance of an ended path goes to its sample entry.
SOL_DEV void retire(const WaveBuf& b, float4 radiance, unsigned long long item) {
    b.samples[item - b.items[kItemChunkBegin]] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
}

template <int BLOCK>
SOL_DEV Stack make_stack_w(int* lds_stack, const TraversalScratch& sc, int lds_entries) {
    int gtid = blockIdx.x * BLOCK + threadIdx.x;
    return Stack::make(lds_stack, threadIdx.x, sc.spill, gtid, sc.max_threads, lds_entries + sc.spill_entries,
                       sc.overflow);
}

// ---- trace -----------------------------------------------------------------------------------------

// Counters are cleared by the first threads of workgroup 0 of a kernel that does not use them,
// between their last reader and their first writer (no kernel of its own: a launch costs about
// 5 µs in the bounce loop, and no memset, which graph replays do not take well).
//   trace(k): the shade queues and the next extension queue, which sort(k) and shade(k) append to
//   (their last readers were shade(k-1) and trace(k-1) / sort(k-1)).
//   sort(k): the shadow queue, which trace(k) has read and shade(k) appends to, and the work
//   counters of the traversal kernel, which trace(k) has used and trace(k+1) uses. When cur is 1,
//   extension queue 0 has been traced and shaded since the last chunk started, so it no longer holds
//   a start fill: the flag is cleared too.
SOL_DEV void clear_for_sort(const WaveBuf& b, int cur) {
    constexpr int n = (kNumShadeQueues + 1) * kMaxSegments;
    if (blockIdx.x == 0 && threadIdx.x < n) {
        const int q = static_cast<int>(threadIdx.x) / kMaxSegments, g = static_cast<int>(threadIdx.x) % kMaxSegments;
        b.counts[qcount(q < kNumShadeQueues ? kQShade0 + q : kQExtend0 + (1 - cur), g)] = 0;
    }
}
SOL_DEV void clear_for_shade(const WaveBuf& b, int cur) {
    if (blockIdx.x == 0 && threadIdx.x < kMaxSegments + kTraceParts) {
        const int i = static_cast<int>(threadIdx.x);
        if (i < kMaxSegments) b.counts[qcount(kQShadow, i)] = 0;
        else b.counts[kCntTraceWork + (i - kMaxSegments) * kCounterStride] = 0;
        if (i == 0 && cur == 1) b.counts[kCntStartFill] = 0;
    }
}


// Persistent traversal with per-lane refill: every lane runs its own ray one traversal step at a
// time; when kRefillLanes lanes are idle the wave hands them the next entries of its chunk
// (fetching a new chunk when the current one runs out). Entries [0, n_ext) are the closest-hit
// queries of extension queue `cur` (hit written in place), [n_ext, n_ext + n_shadow) the any-hit
// queries of the shadow queue (unoccluded contribution added to the path or its sample).
// WIDE: traverse the 4-wide layout (nodes4) instead of the binary one.
template <int BLOCK, int LDS, bool MULTI, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MULTI ? 5 : 6))) void k_trace(SceneView single, WaveTables tabs, WaveBuf b, sol_pt_config cfg,
                                                 int cur, TraversalScratch sc) {
    __shared__ float4 lds_nodes[kTopLdsFloat4<WIDE>];
    __shared__ int lds_stack[LDS * BLOCK];
    clear_for_sort(b, cur);
    const SegQueue eqv(b, kQExtend0 + cur, b.cap_e), sqv(b, kQShadow, b.cap_s);
    const unsigned n_ext = eqv.size();
    const unsigned n_shadow = sqv.size();
    const unsigned total = n_ext + n_shadow;
    if (total == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&b.stats[0], static_cast<unsigned long long>(n_ext));
        atomicAdd(&b.stats[1], static_cast<unsigned long long>(n_shadow));
    }
    // extension queue 0 holding a chunk's start fill: origins are the camera's eye
    const bool start_fill = cur == 0 && b.counts[kCntStartFill] != 0;
    if (!MULTI) {
        if (WIDE) load_top_nodes4<BLOCK>(single, lds_nodes);
        else load_top_nodes<BLOCK>(single, lds_nodes);
    }
    const Stack st = make_stack_w<BLOCK>(lds_stack, sc, LDS);
    // Several scenes (4-wide): each wave keeps the first kWaveNodes nodes of one scene in its share
    // of lds_nodes, the scene of its first busy lane, staged again when that changes.
    constexpr int kWaveNodes = kTopNodes4Max * 64 / BLOCK;
    float4* const wave_nodes = lds_nodes + (threadIdx.x / 64) * (kWaveNodes * 8);
    uint32_t wave_scene = 0xFFFFFFFFu;   // wave-uniform
    int wave_top = 0;
    uint32_t lane_scene = 0;             // the scene of this lane's ray
    const unsigned waves = gridDim.x * (BLOCK / 64);
    const FetchUniform fu(single, lds_nodes);
    const FetchUniform4 fu4(single, lds_nodes);
    FetchGlobal fg(single);
    FetchGlobal4 fg4(single);
    int idx = -1;        // entry this lane traces, -1 = idle
    bool any = false;
    float tmin = 0.0f;
    RayPre r;
    TraversalState ts;
    unsigned c_next = 0, c_end = 0;   // the wave's current chunk [c_next, c_end)
    TraceWork work(total, BLOCK / 64);
    bool exhausted = false;
    unsigned n_holes = 0;   // extension entries without a ray
    while (true) {
        const unsigned long long idle = __ballot(idx < 0);
        const unsigned n_idle = static_cast<unsigned>(__popcll(idle));
        if (!exhausted && n_idle >= static_cast<unsigned>(kRefillLanes)) {
            const unsigned rank = static_cast<unsigned>(__popcll(idle & lanes_below()));
            const unsigned avail = c_end - c_next;
            unsigned base2 = 0, end2 = 0;
            if (avail < n_idle && !trace_fetch(&b.counts[kCntTraceWork], waves, work, base2, end2)) exhausted = true;
            if (idx < 0) {
                unsigned e = 0xFFFFFFFFu;
                if (rank < avail) e = c_next + rank;
                else if (base2 + (rank - avail) < end2) e = base2 + (rank - avail);
                if (e < n_ext) {
                    const unsigned a = eqv.entry(e);
                    float4 e1 = *b.ext(cur, a, 1);
                    uint32_t slot = as_u(e1.w);
                    // a hole's hit reads as a miss (the sort of an in-place chunk reads only hits)
                    if (slot == kHole) {
                        b.hits[e] = make_float4(as_f(0xFFFFFFFFu), 0.0f, 0.0f, 0.0f);
                        ++n_holes;
                    }
                    if (slot != kHole) {
                        idx = static_cast<int>(e);
                        any = false;
                        float4 e0;
                        if (start_fill) {
                            const SceneView& sv = MULTI ? tabs.scenes[as_u(b.rec(slot)->w)] : single;
                            e0 = make_float4(sv.camera.eye[0], sv.camera.eye[1], sv.camera.eye[2], as_f(0u));
                        } else {
                            e0 = *b.ext(cur, a, 0);
                        }
                        tmin = as_u(e0.w) == 0 ? 0.0f : cfg.ray_offset;   // depth 0: a new path
                        r = make_ray_pre(xyz(e0), xyz(e1));
                        ts.start(FLT_MAX);
                        if (MULTI) {
                            lane_scene = as_u(b.rec(slot)->w);
                            const SceneView& sv = tabs.scenes[lane_scene];
                            fg = FetchGlobal(sv);
                            fg4 = FetchGlobal4(sv);
                        }
                    }
                } else if (e != 0xFFFFFFFFu) {
                    const unsigned sa = sqv.entry(e - n_ext);
                    float4 e0 = *b.shadow(sa, 0), e1 = *b.shadow(sa, 1);
                    idx = static_cast<int>(e);
                    any = true;
                    tmin = cfg.ray_offset;
                    r = make_ray_pre(xyz(e0), xyz(e1));
                    ts.start(1.0f - cfg.ray_offset);
                    if (MULTI) {
                        uint32_t target = as_u(e0.w);
                        uint32_t sid;
                        if (target == kSampleTarget) {
                            unsigned long long item = (static_cast<unsigned long long>(as_u(e1.w)) << 32) |
                                                      as_u(b.shadow(sa, 2)->w);
                            sid = tabs.job_scene[find_job(tabs, item)];
                        } else {
                            sid = as_u(b.rec(target)->w);
                        }
                        lane_scene = sid;
                        fg = FetchGlobal(tabs.scenes[sid]);
                        fg4 = FetchGlobal4(tabs.scenes[sid]);
                    }
                }
            }
            if (avail >= n_idle) {
                c_next += n_idle;
            } else if (end2 > base2) {
                c_next = min(base2 + (n_idle - avail), end2);
                c_end = end2;
            } else {
                c_next = c_end;
            }
        }
        const unsigned long long busy = __ballot(idx >= 0);
        if (busy == 0) {
            if (exhausted) break;
            continue;
        }
        if (MULTI && WIDE) {
            const uint32_t lead_scene = static_cast<uint32_t>(
                __builtin_amdgcn_readlane(static_cast<int>(lane_scene), static_cast<int>(__builtin_ctzll(busy))));
            if (lead_scene != wave_scene) {   // uniform: every lane stages its part of the copy
                wave_scene = lead_scene;
                const SceneView& sv = tabs.scenes[lead_scene];
                wave_top = static_cast<int>(min(sv.num_nodes4, static_cast<uint32_t>(kWaveNodes)));
                for (int i = __lane_id(); i < wave_top * 8; i += 64) wave_nodes[i] = sv.nodes4[i];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        if (idx >= 0) {
            int status;
            if (WIDE) status = MULTI ? traverse4_step<BLOCK, LDS, kLeafWaitLanes, true>(
                                           FetchMulti4{fg4, wave_nodes, lane_scene == wave_scene ? wave_top : 0}, st, r, tmin, any, ts)
                                     : traverse4_step<BLOCK, LDS, kLeafWaitLanes, true>(fu4, st, r, tmin, any, ts);
            else status = MULTI ? traverse_step<BLOCK, LDS, true>(fg, st, r, tmin, any, ts)
                                : traverse_step<BLOCK, LDS, true>(fu, st, r, tmin, any, ts);
            if (status != kStepContinue) {
                const unsigned e = static_cast<unsigned>(idx);
                if (!any) {
                    b.hits[e] = ts.hit.prim >= 0
                                    ? make_float4(as_f(static_cast<uint32_t>(ts.hit.slot)), ts.hit.t, ts.hit.u, ts.hit.v)
                                    : make_float4(as_f(0xFFFFFFFFu), 0.0f, 0.0f, 0.0f);
                } else if (status == kStepDone) {
                    // unoccluded: add the light sample's contribution
                    const unsigned sa = sqv.entry(e - n_ext);
                    float4 e0 = *b.shadow(sa, 0), e1 = *b.shadow(sa, 1), ct = *b.shadow(sa, 2);
                    uint32_t target = as_u(e0.w);
                    if (target == kSampleTarget) {
                        unsigned long long item = (static_cast<unsigned long long>(as_u(e1.w)) << 32) | as_u(ct.w);
                        float4* d = b.samples + (item - b.items[kItemChunkBegin]);
                        float4 v = *d;
                        *d = make_float4(v.x + ct.x, v.y + ct.y, v.z + ct.z, 0.0f);
                    } else {
                        float4 p = *b.rec(target);
                        *b.rec(target) = make_float4(p.x + ct.x, p.y + ct.y, p.z + ct.z, p.w);
                    }
                }
                idx = -1;
            }
        }
    }
    // rays traced: the launch's entries, counted once, less the holes (added modulo 2^64)
    wave_add_stat(b.stats, 0, 0ull - static_cast<unsigned long long>(n_holes));
}

// ---- sort ------------------------------------------------------------------------------------------

// Diffuse / glossy (Phong, and the diffuse-Phong mix) / specular: the queues follow the BSDF's
// shading work; a light without BSDF goes to the diffuse queue, whose kernel adds its emission and
// ends the path.
SOL_DEV int queue_of_kind(uint32_t kind) {
    const int bsdf = kind_bsdf(kind);
    return bsdf < 0 ? kQueueDiffuse : queue_of(bsdf);
}

// Material queues of the extension queue's hits: each hit's entry index goes to the queue of its
// triangle's kind (packed into the hit by the traversal kernel, so the sort reads nothing but the
// hit, in one scene as in many). A path that missed, or hit a triangle that neither scatters nor
// emits, ends here; its slot starts the next work item.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_sort(WaveTables tabs, WaveBuf b, int cur) {
    constexpr int K = kSortItems;
    __shared__ RoundLdsK<kNumShadeQueues, K> lds[2];
    clear_for_shade(b, cur);
    const SegQueue eqv(b, kQExtend0 + cur, b.cap_e);
    const unsigned count = eqv.size();
    const int nxt = 1 - cur;
    const int seg = my_segment(b), nseg = static_cast<int>(b.nseg);
    unsigned* qcnt[kNumShadeQueues] = {&b.counts[qcount(kQShade0, seg)], &b.counts[qcount(kQShade0 + 1, seg)],
                                       &b.counts[qcount(kQShade0 + 2, seg)]};
    unsigned* ecnt = &b.counts[qcount(kQExtend0 + nxt, seg)];
    const bool restart = items_left(b);   // uniform
    // the slot word: to retire an ending path (an in-place chunk's paths end where they are)
    const bool need_slot = !b.in_place;
    unsigned started = 0, parity = 0;
    for (unsigned round = blockIdx.x; round * (K * kBlock) < count; round += gridDim.x, parity ^= 1) {
        int dest[K];
        bool dead[K], ret[K];
        uint32_t slot[K];
        unsigned ea[K];
        uint32_t tri[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {   // every load of the round first
            const unsigned i = (round * K + k) * kBlock + threadIdx.x;
            const bool valid = i < count;
            ea[k] = valid ? eqv.entry(i) : 0u;
            tri[k] = valid ? as_u(b.hits[i].x) : kMissWord;   // index | kind; a hole's reads as a miss
            slot[k] = valid && need_slot ? as_u(b.ext(cur, ea[k], 1)->w) : 0u;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const unsigned i = (round * K + k) * kBlock + threadIdx.x;
            const bool valid = i < count && !(need_slot && slot[k] == kHole);
            dest[k] = -1;
            dead[k] = false;
            if (valid) {
                const uint32_t kind = tri[k] == kMissWord ? 0u : tri[k] >> kHitKindShift;
                if (kind == 0) dead[k] = true;   // a miss, or no BSDF and no light
                else dest[k] = b.merged ? 0 : queue_of_kind(kind);
                if (dead[k] && !b.in_place) retire(b, *b.rec(slot[k]), item_of(*b.ext(cur, ea[k], 3)));
            }
            ret[k] = dead[k] && restart;
        }
        RoundPlan plan[K];
        plan_round_k<kNumShadeQueues, K>(lds[parity], qcnt, ecnt, b.items, seg, nseg, dest, ret, plan);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const unsigned i = (round * K + k) * kBlock + threadIdx.x;
            if (plan[k].pos >= 0) *b.shade_entry(dest[k], seg * b.cap_m + static_cast<unsigned>(plan[k].pos)) = i;
            if (plan[k].start) {
                ++started;
                init_path(tabs, b, plan[k].item, slot[k], nxt, seg * b.cap_e + static_cast<unsigned>(plan[k].ext_pos));
            } else if (plan[k].ext_pos >= 0) {
                write_hole(b, nxt, seg * b.cap_e + static_cast<unsigned>(plan[k].ext_pos));
            }
        }
    }
    wave_add_stat(b.stats, 2, started);
}

// ---- shade -----------------------------------------------------------------------------------------

// One kernel per queue; the material type of the queue is a compile-time constant (glass and
// mirror share the specular queue), so each kernel carries only its BSDF's code. The glossy
// kernel (Phong and the diffuse-Phong mix) is held to three waves per SIMD (at most 168 VGPRs).
// QUEUE == kQueueAll: the merged queue (stored as queue 0), every BSDF chosen at run time, held
// to three waves per SIMD like the glossy kernel.
// Per entry: the ray and path state from the extension entry of queue 1 - nxt, the hit decoded
// (triangle_mesh.cpp:78-104), the emission of a front-side light hit added to the path's radiance
// (path_tracer.cpp:49-60), then - if the surface has a BSDF - NEE, Russian roulette and the BSDF
// sample of the next ray (path_tracer.cpp:62-118).
template <int QUEUE, bool MULTI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(QUEUE == kQueuePhong || QUEUE == kQueueAll ? 3 : 1))) void k_shade(SceneView single, WaveTables tabs, WaveBuf b, sol_pt_config cfg,
                                                  int nxt) {
    constexpr int SQ = QUEUE == kQueueAll ? 0 : QUEUE;   // storage queue
    __shared__ RoundLds<1> lds[2];
    const int cur = 1 - nxt;
    const SegQueue mqv(b, kQShade0 + SQ, b.cap_m), eqv(b, kQExtend0 + cur, b.cap_e);
    const unsigned count = mqv.size();
    // extension queue 0 holding a chunk's start fill: origins are the camera's eye (see k_trace)
    const bool start_fill = cur == 0 && b.counts[kCntStartFill] != 0;
    const int seg = my_segment(b), nseg = static_cast<int>(b.nseg);
    unsigned* scnt[1] = {&b.counts[qcount(kQShadow, seg)]};
    unsigned* ecnt = &b.counts[qcount(kQExtend0 + nxt, seg)];
    const bool restart = items_left(b);   // uniform
    unsigned started = 0, parity = 0;
    for (unsigned round = blockIdx.x; round * kBlock < count; round += gridDim.x, parity ^= 1) {
        unsigned j = round * kBlock + threadIdx.x;
        bool valid = j < count;
        uint32_t slot = 0;
        bool shoot = false, cont = false;
        V3 point{0, 0, 0}, sdir{0, 0, 0}, contrib{0, 0, 0}, ndir{0, 0, 0}, T_next{0, 0, 0};
        float pdf_next = 0.0f;
        uint32_t depth_next = 0;
        unsigned long long item = 0, rng_next = 0;
        if (valid) {
            const unsigned i = *b.shade_entry(SQ, mqv.entry(j));
            const unsigned ea = eqv.entry(i);
            const float4 h = b.hits[i], e1 = *b.ext(cur, ea, 1), e3 = *b.ext(cur, ea, 3);
            slot = as_u(e1.w);
            SceneView tmp;
            const SceneView& s = scene_of<MULTI>(single, tabs, b, slot, tmp);
            const float4 e0 = start_fill ? make_float4(s.camera.eye[0], s.camera.eye[1], s.camera.eye[2], as_f(0u))
                                         : *b.ext(cur, ea, 0);
            const uint32_t depth0 = as_u(e0.w);
            const float4 e2 = ext_throughput(b, cur, ea, depth0);
            const V3 dir = xyz(e1);
            const Surface sf = decode_hit(s, xyz(e0), dir, static_cast<int>(as_u(h.x) & kHitIndexMask), h.y, h.z, h.w);
            const float pick = 1.0f / static_cast<float>(s.num_lights);
            item = item_of(e3);
            V3 T = xyz(e2);
            if (sf.light >= 0 && sf.front) {
                const float4 rad = *b.rec(slot);
                V3 Le;
                float w;
                const V3 C = xyz(rad) + hit_emission(s, cfg.flags, sf, dir, h.y, e2.w, pick, T, Le, w);
                *b.rec(slot) = f4(C, rad.w);
            }
            if (sf.mat >= 0) {
                point = sf.point;
                const V3 out = -dir;
                sol_material m = s.materials[sf.mat];
                if (QUEUE == kQueueDiffuse) m.type = SOL_BSDF_DIFFUSE;
                if (QUEUE == kQueuePhong && m.type != SOL_BSDF_INTERP) m.type = SOL_BSDF_PHONG;
                if (QUEUE == kQueueSpecular && m.type != SOL_BSDF_GLASS) m.type = SOL_BSDF_MIRROR;
                Pcg rng;
                rng.state = rng_of(e3);
                uint32_t depth = depth0;
                bool skip_nee = QUEUE == kQueueSpecular || (QUEUE == kQueueAll && is_specular(m.type)) ||
                                (cfg.flags & SOL_PT_DISABLE_NEE);
                if (!skip_nee) {
                    NeeSample ns;
                    shoot = nee_sample(s, cfg.flags, sf, m, out, T, rng, pick, ns);
                    sdir = ns.shadow_dir;
                    contrib = ns.contrib;
                }
                float q;
                BsdfSample bs;
                if (russian_roulette(cfg, depth, T, rng, q) && bsdf_sample(m, sf, ortho_basis(sf.n), out, rng, bs)) {
                    T = T * (bs.color * (bs.cos / (bs.pdf * q)));
                    depth += 1;
                    cont = depth < cfg.max_path_len;
                    ndir = bs.in;
                    T_next = T;
                    pdf_next = skip_nee ? 0.0f : bs.pdf;
                    depth_next = depth;
                    rng_next = rng.state;
                }
            }
            if (!cont && !b.in_place) retire(b, *b.rec(slot), item);
        }
        bool retired = valid && !cont && restart;
        RoundPlan plan = plan_round<1>(lds[parity], scnt, ecnt, b.items, seg, nseg, shoot ? 0 : -1, cont, retired);
        if (plan.pos >= 0) {
            const unsigned a = seg * b.cap_s + static_cast<unsigned>(plan.pos);
            *b.shadow(a, 0) = f4(point, as_f(cont ? slot : kSampleTarget));
            *b.shadow(a, 1) = f4(sdir, as_f(static_cast<uint32_t>(item >> 32)));
            *b.shadow(a, 2) = f4(contrib, as_f(static_cast<uint32_t>(item)));
        }
        if (plan.ext_pos >= 0) {
            const unsigned a = seg * b.cap_e + static_cast<unsigned>(plan.ext_pos);
            if (cont) {
                write_ext(b, nxt, a, point, depth_next, ndir, slot, T_next, pdf_next, rng_next, item);
            } else if (plan.start) {
                ++started;
                init_path(tabs, b, plan.item, slot, nxt, a);
            } else {
                write_hole(b, nxt, a);
            }
        }
    }
    wave_add_stat(b.stats, 2, started);
}

// ---- start / chunk / reduce ------------------------------------------------------------------------

// The start fill lays the slots out as the rounds of the other passes would: slot s is in round
// s / kBlock, round r in segment r % nseg, and a slot's place in its segment follows the rounds in
// order. Every slot gets one extension entry (a path, or a hole if no item is left).
SOL_HD unsigned start_segment(unsigned slot, unsigned nseg) { return (slot / kBlock) % nseg; }
SOL_HD unsigned start_position(unsigned slot, unsigned nseg) {
    return (slot / kBlock / nseg) * kBlock + slot % kBlock;
}
SOL_HD unsigned start_slots(unsigned slots, unsigned seg, unsigned nseg) {   // slots of segment seg
    const unsigned full = slots / kBlock, rem = slots % kBlock;
    unsigned c = full > seg ? ((full - 1 - seg) / nseg + 1) * kBlock : 0u;
    if (rem && full % nseg == seg) c += rem;
    return c;
}

// Starts a path in every slot (beginning of a chunk of work), in the layout above and without
// atomics: k_set_chunk gave segment g the items from items[item_next(g) + 1], and the slot at
// position p of segment g takes item items[item_next(g) + 1] + p, if below items[item_next(g)].
// A wave fills 64 consecutive slots at a time, so its slot records are one contiguous 1 KB run. Of
// the extension entries only the direction and slot and the PCG state and item are written: while
// the start flag is set, the origin is the camera's eye, the depth 0 and the throughput 1.
__global__ __launch_bounds__(kBlock) void k_start(WaveTables tabs, WaveBuf b, int nxt) {
    const unsigned wave_step = gridDim.x * kBlock;
    unsigned started = 0;
    for (unsigned base = blockIdx.x * kBlock + (threadIdx.x & ~63u); base < b.slots; base += wave_step) {
        const unsigned slot = base + __lane_id();
        if (slot < b.slots) {
            const unsigned seg = start_segment(slot, b.nseg), pos = start_position(slot, b.nseg);
            const unsigned long long item = b.items[item_next(static_cast<int>(seg)) + 1] + pos;
            const unsigned entry = seg * b.cap_e + pos;
            if (item < b.items[item_next(static_cast<int>(seg))]) {
                ++started;
                const PathStart ps = start_path(tabs, item);
                // an in-place chunk names its paths by their items: the record is the sample entry
                const unsigned id = b.in_place ? static_cast<unsigned>(item - b.items[kItemChunkBegin]) : slot;
                *b.rec(id) = make_float4(0.0f, 0.0f, 0.0f, as_f(ps.scene));
                *b.ext(nxt, entry, 1) = f4(ps.dir, as_f(id));
                *b.ext(nxt, entry, 3) = rng_item_word(ps.rng, item);
            } else {
                write_hole(b, nxt, entry);
            }
        }
    }
    wave_add_stat(b.stats, 2, started);
}

// Work items of a chunk [begin, end) by segment, consecutive ranges in segment order: segment g
// first gets as many items as the start fill has slots in it (k_start takes them, so its counter
// starts past them and the range start is kept in the next word), then an even share of the rest
// for the paths that restart there. The start fill's extension entries are counted in advance.
__global__ void k_set_chunk(WaveBuf b, unsigned long long begin, unsigned long long end, int nxt) {
    if (threadIdx.x < kNumWaveCounters) b.counts[threadIdx.x * kCounterStride] = 0;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const unsigned nseg = b.nseg;
    unsigned long long left = end - begin, filled[kMaxSegments];
    for (unsigned g = 0; g < nseg; ++g) {
        filled[g] = min(static_cast<unsigned long long>(start_slots(b.slots, g, nseg)), left);
        left -= filled[g];
    }
    unsigned long long base = begin;
    for (unsigned g = 0; g < nseg; ++g) {
        const unsigned long long extra = left * (g + 1) / nseg - left * g / nseg;
        b.items[item_next(g) + 1] = base;
        b.items[item_next(g)] = base + filled[g];
        b.items[item_end(g)] = base + filled[g] + extra;
        b.counts[qcount(kQExtend0 + nxt, g)] = start_slots(b.slots, g, nseg);
        base += filled[g] + extra;
    }
    b.items[kItemChunkBegin] = begin;
    b.counts[kCntStartFill] = 1;
}

// After an even number of bounce iterations of a chunk that should have ended in them: adds the
// entries left in extension queue 0 and the shadow queue (zero when it has ended) to statistic
// kStatUndrained, which the host reads with the others.
__global__ void k_check_drained(WaveBuf b) {
    if (threadIdx.x != 0) return;
    unsigned long long left = 0;
    for (int k = 0; k < kMaxSegments; ++k) left += b.counts[qcount(kQExtend0, k)] + b.counts[qcount(kQShadow, k)];
    if (left) atomicAdd(&b.stats[kStatUndrained], left);
}

// Per pixel of the chunk's tiles of jobs in reduction layer `layer`: sum the samples in sample
// order, add to the image once.
__global__ __launch_bounds__(256) void k_reduce(WaveTables tabs, WaveBuf b, uint32_t first_tile,
                                                unsigned long long chunk_begin, uint32_t layer) {
    uint32_t tile = first_tile + blockIdx.x;
    if (tabs.job_layer[tabs.tile_job[tile]] != layer) return;
    uint32_t tw = tabs.tile_w[tile], th = tabs.tile_h[tile];
    uint32_t per_sample = tw * th;
    uint32_t pix = blockIdx.y * 256 + threadIdx.x;
    if (pix >= per_sample) return;
    unsigned long long k0 = tabs.tile_item0[tile] - chunk_begin;
    unsigned long long spp = (tabs.tile_item0[tile + 1] - tabs.tile_item0[tile]) / per_sample;
    uint32_t job = tabs.tile_job[tile];
    uint32_t w = tabs.job_w[job], h = tabs.job_h[job];
    uint32_t dx, dy;
    tile_pixel(pix, tw, th, dx, dy);
    const uint32_t x = tabs.tile_x0[tile] + dx, y = tabs.tile_y0[tile] + dy;
    size_t plane = static_cast<size_t>(w) * h, p = static_cast<size_t>(y) * w + x;
    float4* sout = tabs.job_sout[job];
    if (sout) sout += p * tabs.job_sstride[job] + tabs.job_soff[job];
    float r = 0.0f, g = 0.0f, bl = 0.0f;
    for (unsigned long long s = 0; s < spp; ++s) {
        float4 v = b.samples[k0 + s * per_sample + pix];
        if (sout) sout[s] = make_float4(v.x, v.y, v.z, 0.0f);   // (an in-place record holds its scene in w)
        r = r + v.x;
        g = g + v.y;
        bl = bl + v.z;
    }
    float* accum = tabs.job_accum[job];
    accum[p] += r;
    accum[plane + p] += g;
    accum[2 * plane + p] += bl;
}

template <int BLOCK, int LDS, bool MULTI>
void launch_trace_t(const WaveLaunch& L, int cur) {
    if (L.wide)
        hipLaunchKernelGGL((k_trace<BLOCK, LDS, MULTI, true>), dim3(L.grid_traverse), dim3(BLOCK), 0, L.stream, L.single,
                           L.tabs, L.buf, L.cfg, cur, L.scratch);
    else
        hipLaunchKernelGGL((k_trace<BLOCK, LDS, MULTI, false>), dim3(L.grid_traverse), dim3(BLOCK), 0, L.stream,
                           L.single, L.tabs, L.buf, L.cfg, cur, L.scratch);
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

namespace {
template <typename K>
hipError_t resident_grid(K kernel, int compute_units, int cap, int& grid) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel),
                                                               kBlock, 0);
    if (e != hipSuccess) return e;
    // a multiple of kMaxSegments: round r goes to workgroup r % grid, whose segment is then r % nseg
    grid = std::min(cap, compute_units * std::max(1, per_cu)) / kMaxSegments * kMaxSegments;
    grid = std::max(grid, kMaxSegments);
    return hipSuccess;
}
} // namespace

hipError_t wave_set_grids(WaveLaunch& L, int compute_units) {
    hipError_t e;
    const int cap = L.grid_shade;
    if ((e = resident_grid(k_sort, compute_units, cap, L.grid_sort)) != hipSuccess) return e;
    if (L.multi) {
        if ((e = resident_grid(k_shade<0, true>, compute_units, cap, L.grid_kshade[0])) != hipSuccess) return e;
        if ((e = resident_grid(k_shade<1, true>, compute_units, cap, L.grid_kshade[1])) != hipSuccess) return e;
        if ((e = resident_grid(k_shade<2, true>, compute_units, cap, L.grid_kshade[2])) != hipSuccess) return e;
        if ((e = resident_grid(k_shade<kQueueAll, true>, compute_units, cap, L.grid_kshade[3])) != hipSuccess) return e;
    } else {
        if ((e = resident_grid(k_shade<0, false>, compute_units, cap, L.grid_kshade[0])) != hipSuccess) return e;
        if ((e = resident_grid(k_shade<1, false>, compute_units, cap, L.grid_kshade[1])) != hipSuccess) return e;
        if ((e = resident_grid(k_shade<2, false>, compute_units, cap, L.grid_kshade[2])) != hipSuccess) return e;
        if ((e = resident_grid(k_shade<kQueueAll, false>, compute_units, cap, L.grid_kshade[3])) != hipSuccess) return e;
    }
    return resident_grid(k_start, compute_units, cap, L.grid_start);
}

hipError_t wave_launch_trace(const WaveLaunch& L, int cur) {
    (void)hipGetLastError();
    if (L.multi) { SOLHIP_WAVE_DISPATCH(launch_trace_t, true, L, cur) }
    else { SOLHIP_WAVE_DISPATCH(launch_trace_t, false, L, cur) }
    return hipGetLastError();
}

hipError_t wave_launch_sort(const WaveLaunch& L, int cur) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_sort, dim3(L.grid_sort), dim3(kBlock), 0, L.stream, L.tabs, L.buf, cur);
    return hipGetLastError();
}

hipError_t wave_launch_shade(const WaveLaunch& L, int queue, int nxt) {
    dim3 g(L.grid_kshade[queue]), bs(kBlock);
    (void)hipGetLastError();
    if (L.multi) {
        if (queue == 0) hipLaunchKernelGGL((k_shade<0, true>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 1) hipLaunchKernelGGL((k_shade<1, true>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 2) hipLaunchKernelGGL((k_shade<2, true>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == kQueueAll)
            hipLaunchKernelGGL((k_shade<kQueueAll, true>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
    } else {
        if (queue == 0) hipLaunchKernelGGL((k_shade<0, false>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 1) hipLaunchKernelGGL((k_shade<1, false>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == 2) hipLaunchKernelGGL((k_shade<2, false>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
        if (queue == kQueueAll)
            hipLaunchKernelGGL((k_shade<kQueueAll, false>), g, bs, 0, L.stream, L.single, L.tabs, L.buf, L.cfg, nxt);
    }
    return hipGetLastError();
}

hipError_t wave_launch_start(const WaveLaunch& L, int nxt) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_start, dim3(L.grid_start), dim3(kBlock), 0, L.stream, L.tabs, L.buf, nxt);
    return hipGetLastError();
}

hipError_t wave_launch_set_chunk(const WaveLaunch& L, unsigned long long begin, unsigned long long end) {
    (void)hipGetLastError();
    // the start fill that follows writes extension queue 0
    hipLaunchKernelGGL(k_set_chunk, dim3(1), dim3(64), 0, L.stream, L.buf, begin, end, 0);
    return hipGetLastError();
}

hipError_t wave_launch_check_drained(const WaveLaunch& L) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_check_drained, dim3(1), dim3(64), 0, L.stream, L.buf);
    return hipGetLastError();
}

hipError_t wave_launch_reduce(const WaveLaunch& L, uint32_t first_tile, uint32_t num_tiles, uint32_t max_tile_pixels,
                              unsigned long long chunk_begin, uint32_t layer) {
    dim3 g(num_tiles, (max_tile_pixels + 255) / 256);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_reduce, g, dim3(256), 0, L.stream, L.tabs, L.buf, first_tile, chunk_begin, layer);
    return hipGetLastError();
}

} // namespace solhip
