// Host side of the wavefront renderer: buffers, the tile work queue, the graph-captured bounce
// loop, and the single-scene and batched entry points.
#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "state.h"
#include "wave_launch.h"

namespace solhip {

namespace {

#define HIP_OK(expr)                                                                             \
    do {                                                                                         \
        hipError_t e__ = (expr);                                                                 \
        if (e__ != hipSuccess) {                                                                 \
            (void)hipGetLastError();                                                             \
            return set_error(SOLHIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e__));   \
        }                                                                                        \
    } while (0)

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Longest graph (bounce iterations) captured to finish a chunk that fits in the paths in flight.
constexpr uint32_t kMaxFillIters = 64;

struct Job {
    uint32_t scene;
    uint32_t w, h;
    uint64_t sample_index;
    uint32_t sample_count;
    float* accum;
    float4* samples = nullptr;   // per-sample radiance, pixel-major (sample_count per pixel), or null
};

} // namespace

// The tables of a render call (scenes, jobs, tiles) in one device buffer, written by one copy from
// pinned host memory (twenty copies from pageable memory took about 0.3 ms a call). Each table
// keeps a grow-only slot, so its device address changes only when it outgrows the slot.
struct TableSet {
    static constexpr int kMax = 24;
    size_t cap[kMax] = {}, off[kMax] = {};
    size_t used = 0;               // bytes of the laid-out slots
    DeviceBuffer dev;
    void* host = nullptr;          // pinned staging of `used` bytes
    hipEvent_t copied = nullptr;   // the last copy out of the staging memory

    TableSet() = default;
    TableSet(const TableSet&) = delete;
    TableSet& operator=(const TableSet&) = delete;
    ~TableSet() {
        if (copied) {
            (void)hipEventSynchronize(copied);
            (void)hipEventDestroy(copied);
        }
        if (host) (void)hipHostFree(host);
    }

    // Copies table i (bytes[i] bytes at data[i]) to the device and gives its address in dst[i];
    // bumps `generation` when the layout, and so any address, changes.
    int upload(int n, const void* const* data, const size_t* bytes, void** dst, hipStream_t s, uint64_t& generation) {
        if (n > kMax) return set_error(SOLHIP_E_INVALID, "too many tables");
        if (!copied) HIP_OK(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
        else HIP_OK(hipEventSynchronize(copied));   // the staging memory is free again
        bool relayout = !dev.ptr || !host;   // also after an allocation that failed
        for (int i = 0; i < n; ++i) {
            if (bytes[i] > cap[i]) {
                cap[i] = std::max<size_t>(256, 2 * bytes[i]);
                relayout = true;
            }
        }
        if (relayout) {
            size_t o = 0;
            for (int i = 0; i < n; ++i) {
                off[i] = o;
                o += align_up(cap[i]);
            }
            used = o;
            HIP_OK(dev.alloc(used));
            if (host) HIP_OK(hipHostFree(host));
            host = nullptr;
            HIP_OK(hipHostMalloc(&host, used, hipHostMallocDefault));
            ++generation;
        }
        size_t end = 0;
        for (int i = 0; i < n; ++i) {
            if (bytes[i]) std::memcpy(static_cast<char*>(host) + off[i], data[i], bytes[i]);
            dst[i] = dev.as<char>() + off[i];
            end = std::max(end, off[i] + bytes[i]);
        }
        if (end) HIP_OK(hipMemcpyAsync(dev.ptr, host, end, hipMemcpyHostToDevice, s));
        HIP_OK(hipEventRecord(copied, s));
        return SOLHIP_OK;
    }
};

struct WavefrontState {
    uint32_t slots = 0;
    DeviceBuffer arena, counters, samples;
    unsigned long long sample_capacity = 0;
    WaveBuf buf{};
    TableSet tables;
    uint64_t generation = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;
    std::vector<hipEvent_t> events;
    std::vector<unsigned> host_counts = std::vector<unsigned>(kCounterWords);
    std::vector<unsigned long long> host_stats = std::vector<unsigned long long>(kStatWords);

    ~WavefrontState() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }

    // Segment capacities (see WaveBuf) of n slots in nseg segments: extension queues
    // 2 n / nseg + 8 R, shade queues n / nseg + 2 R, shadow queue n / nseg + 6 R, R = kWaveRound,
    // rounded up to 64 entries. The bounds hold for the material-split queues (three shade
    // launches) and so for the merged queue (one).
    struct Layout {
        size_t cap_e, cap_m, cap_s;
        // slot record and hit 16 B each per slot; queue entries 64 B (extension), 4 B (shade),
        // 48 B (shadow) per segment entry; each array 256-byte aligned
        size_t bytes(size_t n, size_t nseg) const {
            return 2 * align_up(sizeof(float4) * n) + 2 * align_up(sizeof(float4) * 4 * nseg * cap_e) +
                   kNumShadeQueues * align_up(sizeof(uint32_t) * nseg * cap_m) + align_up(sizeof(float4) * 3 * nseg * cap_s);
        }
    };
    static Layout layout(size_t n, size_t nseg) {
        const size_t R = kWaveRound;
        auto round64 = [](size_t v) { return (v + 63) & ~size_t(63); };
        return {round64(2 * n / nseg + 8 * R), round64(n / nseg + 2 * R), round64(n / nseg + 6 * R)};
    }

    // Grows the per-slot arena to at least n slots (never shrinks it; sized for every segment
    // count) and lays the queues out in it for nseg segments and the shading queue split.
    int ensure_slots(uint32_t n, uint32_t nseg, bool merged) {
        if (!arena.ptr || n > slots) {
            size_t total = 0;
            for (size_t g = 1; g <= kMaxSegments; g *= 2) total = std::max(total, layout(n, g).bytes(n, g));
            HIP_OK(arena.alloc(total));
            HIP_OK(counters.alloc(sizeof(unsigned) * kCounterWords + align_up(sizeof(unsigned long long) * kItemWords) +
                                  sizeof(unsigned long long) * kStatWords));
            slots = n;
            buf.nseg = 0;   // lay out again
        }
        WaveBuf& b = buf;
        if (b.nseg == nseg && (b.merged != 0) == merged) return SOLHIP_OK;
        const Layout lay = layout(slots, nseg);
        char* p = arena.as<char>();
        auto f4s = [&](size_t count) { float4* r = reinterpret_cast<float4*>(p); p += align_up(sizeof(float4) * count); return r; };
        auto u32s = [&](size_t count) { uint32_t* r = reinterpret_cast<uint32_t*>(p); p += align_up(sizeof(uint32_t) * count); return r; };
        b.slots = slots;
        b.nseg = nseg;
        b.merged = merged ? 1u : 0u;
        b.cap_e = static_cast<uint32_t>(lay.cap_e);
        b.cap_m = static_cast<uint32_t>(lay.cap_m);
        b.cap_s = static_cast<uint32_t>(lay.cap_s);
        b.stride_e = static_cast<uint32_t>(nseg * lay.cap_e);
        b.stride_s = static_cast<uint32_t>(nseg * lay.cap_s);
        b.path = f4s(slots);
        b.hits = f4s(slots);
        for (int q = 0; q < 2; ++q) b.eq[q] = f4s(nseg * lay.cap_e * 4);
        for (int q = 0; q < kNumShadeQueues; ++q) b.mq[q] = u32s(nseg * lay.cap_m);
        b.sq = f4s(nseg * lay.cap_s * 3);
        char* c = counters.as<char>();
        b.counts = reinterpret_cast<unsigned*>(c);
        b.items = reinterpret_cast<unsigned long long*>(c + sizeof(unsigned) * kCounterWords);
        b.stats = reinterpret_cast<unsigned long long*>(c + sizeof(unsigned) * kCounterWords +
                                                        align_up(sizeof(unsigned long long) * kItemWords));
        ++generation;
        return SOLHIP_OK;
    }

    int ensure_samples(unsigned long long items) {
        if (items <= sample_capacity && samples.ptr) return SOLHIP_OK;
        HIP_OK(samples.alloc(sizeof(float4) * items));
        sample_capacity = items;
        buf.samples = samples.as<float4>();
        buf.sample_capacity = items;
        ++generation;
        return SOLHIP_OK;
    }

    int ensure_events(size_t n) {
        while (events.size() < n) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            events.push_back(e);
        }
        return SOLHIP_OK;
    }
};

void Context::release_wavefront() {
    delete wave;
    wave = nullptr;
}

namespace {

// One bounce of every path in flight: trace (closest hits of queue `cur`, any hits of the shadow
// queue), sort, shade (the material queues in use, or the merged one). Paths that end in sort or
// shade start their slot's next work item there.
hipError_t enqueue_iteration(const WaveLaunch& L, int cur, hipEvent_t* ev) {
    int nxt = 1 - cur;
    hipError_t e;
    if (ev && (e = hipEventRecord(ev[0], L.stream)) != hipSuccess) return e;
    if ((e = wave_launch_trace(L, cur)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[1], L.stream)) != hipSuccess) return e;
    if ((e = wave_launch_sort(L, cur)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[2], L.stream)) != hipSuccess) return e;
    if (L.buf.merged) {
        if ((e = wave_launch_shade(L, kQueueAll, nxt)) != hipSuccess) return e;
    } else {
        for (int q = 0; q < kNumShadeQueues; ++q)
            if ((L.shade_mask >> q & 1u) && (e = wave_launch_shade(L, q, nxt)) != hipSuccess) return e;
    }
    if (ev && (e = hipEventRecord(ev[3], L.stream)) != hipSuccess) return e;
    return hipSuccess;
}

std::string graph_key(const WaveLaunch& L, const WavefrontState& st, int iters) {
    std::string k(reinterpret_cast<const char*>(&L.cfg), sizeof(L.cfg));
    k.append(reinterpret_cast<const char*>(&L.single), sizeof(L.single));
    k.append(reinterpret_cast<const char*>(&L.tabs), sizeof(L.tabs));
    k.append(reinterpret_cast<const char*>(&L.buf), sizeof(L.buf));
    int meta[14] = {L.block,          L.lds_stack,      L.grid_traverse,  L.grid_shade,
                    L.grid_sort,      L.grid_kshade[0], L.grid_kshade[1], L.grid_kshade[2],
                    L.grid_kshade[3], L.grid_start,     L.multi ? 1 : 0,  L.wide ? 1 : 0,
                    static_cast<int>(L.shade_mask), iters};
    k.append(reinterpret_cast<const char*>(meta), sizeof(meta));
    k.append(reinterpret_cast<const char*>(&st.generation), sizeof(st.generation));
    k.append(reinterpret_cast<const char*>(&L.scratch), sizeof(L.scratch));
    return k;
}

// Reduction layer of each job (planar RGB images of 3 w h floats at accum): 0 unless the job's
// image overlaps the image of an earlier job, else one more than the largest layer of the earlier
// jobs it overlaps. Layers are reduced in increasing order, so overlapping jobs add in job order
// (as successive renders would) and jobs of one layer never share a pixel. Intervals are swept in
// address order; only jobs in a run of overlapping intervals are compared pairwise.
std::vector<uint32_t> reduction_layers(const std::vector<float*>& accum, const std::vector<uint32_t>& w,
                                       const std::vector<uint32_t>& h, uint32_t& num_layers) {
    const size_t n = accum.size();
    std::vector<uint32_t> layer(n, 0);
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    auto lo = [&](size_t i) { return reinterpret_cast<uintptr_t>(accum[i]); };
    auto hi = [&](size_t i) { return lo(i) + sizeof(float) * 3 * static_cast<uintptr_t>(w[i]) * h[i]; };
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return lo(a) < lo(b) || (lo(a) == lo(b) && a < b); });
    num_layers = 1;
    for (size_t r = 0; r < n;) {
        size_t e = r + 1;
        uintptr_t end = hi(order[r]);
        while (e < n && lo(order[e]) < end) end = std::max(end, hi(order[e++]));
        if (e - r > 1) {
            std::vector<size_t> run(order.begin() + r, order.begin() + e);
            std::sort(run.begin(), run.end());   // job order
            for (size_t a = 0; a < run.size(); ++a)
                for (size_t b = 0; b < a; ++b)
                    if (lo(run[b]) < hi(run[a]) && lo(run[a]) < hi(run[b]))
                        layer[run[a]] = std::max(layer[run[a]], layer[run[b]] + 1);
            for (size_t a : run) num_layers = std::max(num_layers, layer[a] + 1);
        }
        r = e;
    }
    return layer;
}

// Runs the jobs to completion. Jobs index `views` through Job::scene.
int run_wavefront(Context& ctx, const std::vector<SceneView>& views, const std::vector<Job>& jobs,
                  const sol_pt_config& cfg, uint32_t max_depth, uint32_t max_depth4, uint32_t shade_mask) {
    if (!ctx.wave) ctx.wave = new WavefrontState();
    WavefrontState& st = *ctx.wave;
    const solhip_tuning& tu = ctx.tuning;
    int rc;

    // Jobs in passes of samples: a pass keeps a job under 2^32 items and one tile's samples
    // under the sample-buffer capacity (far above any ordinary sample count); each pass is
    // accumulated into the image on its own.
    const unsigned long long cap = tu.chunk_items;
    std::vector<unsigned long long> job_item0, job_sample0;
    std::vector<uint32_t> job_scene, job_w, job_h, job_spp, job_tw, job_th;
    std::vector<float*> job_accum;
    std::vector<float4*> job_sout;
    std::vector<uint32_t> job_sstride, job_soff;
    std::vector<uint32_t> tile_job, tile_x0, tile_y0, tile_w, tile_h;
    std::vector<unsigned long long> tile_item0;
    unsigned long long total = 0;
    uint32_t max_tile_pixels = 0;
    for (const Job& j : jobs) {
        uint32_t tw = std::min(tu.tile_w, j.w), th = std::min(tu.tile_h, j.h);
        unsigned long long px = static_cast<unsigned long long>(j.w) * j.h;
        unsigned long long max_pass = std::min<unsigned long long>(cap / (static_cast<unsigned long long>(tw) * th),
                                                                   0xFFFFFFFFull / px);
        if (max_pass == 0) return set_error(SOLHIP_E_INVALID, "image too large for one pass");
        for (uint32_t s0 = 0; s0 < j.sample_count;) {
            uint32_t spp = static_cast<uint32_t>(std::min<unsigned long long>(max_pass, j.sample_count - s0));
            uint32_t jid = static_cast<uint32_t>(job_item0.size());
            job_item0.push_back(total);
            job_sample0.push_back(j.sample_index + s0);
            job_scene.push_back(j.scene);
            job_w.push_back(j.w);
            job_h.push_back(j.h);
            job_spp.push_back(spp);
            job_tw.push_back(tw);
            job_th.push_back(th);
            job_accum.push_back(j.accum);
            job_sout.push_back(j.samples);
            job_sstride.push_back(j.sample_count);
            job_soff.push_back(s0);
            // tiles in item order: bands of th rows, tiles of tw columns
            for (uint32_t y0 = 0; y0 < j.h; y0 += th) {
                uint32_t bh = std::min(th, j.h - y0);
                for (uint32_t x0 = 0; x0 < j.w; x0 += tw) {
                    uint32_t w2 = std::min(tw, j.w - x0);
                    tile_job.push_back(jid);
                    tile_x0.push_back(x0);
                    tile_y0.push_back(y0);
                    tile_w.push_back(w2);
                    tile_h.push_back(bh);
                    tile_item0.push_back(total);
                    total += static_cast<unsigned long long>(w2) * bh * spp;
                    max_tile_pixels = std::max(max_tile_pixels, w2 * bh);
                }
            }
            s0 += spp;
        }
    }
    job_item0.push_back(total);
    tile_item0.push_back(total);
    const uint32_t num_tiles = static_cast<uint32_t>(tile_job.size());
    const uint32_t num_jobs = static_cast<uint32_t>(job_scene.size());
    uint32_t num_layers = 1;
    const std::vector<uint32_t> job_layer = reduction_layers(job_accum, job_w, job_h, num_layers);

    hipStream_t s = ctx.stream;
    void* tab[20];
    {
        auto entry = [](const auto& v) {
            return std::make_pair(static_cast<const void*>(v.data()), sizeof(v[0]) * v.size());
        };
        const std::pair<const void*, size_t> t[20] = {
            entry(views),      entry(job_item0), entry(job_sample0), entry(job_scene),  entry(job_accum),
            entry(job_w),      entry(job_sout),  entry(job_sstride),  entry(job_soff),   entry(job_layer),
            entry(job_h),      entry(job_spp),   entry(job_tw),       entry(job_th),     entry(tile_job),
            entry(tile_x0),    entry(tile_y0),   entry(tile_w),       entry(tile_h),     entry(tile_item0)};
        const void* data[20];
        size_t bytes[20];
        for (int i = 0; i < 20; ++i) {
            data[i] = t[i].first;
            bytes[i] = t[i].second;
        }
        if ((rc = st.tables.upload(20, data, bytes, tab, s, st.generation))) return rc;
    }

    // Chunks of whole tiles whose items fit the sample buffer.
    struct Chunk { uint32_t first, count; unsigned long long begin, end; };
    std::vector<Chunk> chunks;
    unsigned long long biggest = 0;
    for (uint32_t t = 0; t < num_tiles;) {
        Chunk c{t, 0, tile_item0[t], tile_item0[t]};
        while (t < num_tiles && (c.count == 0 || tile_item0[t + 1] - c.begin <= cap)) {
            c.end = tile_item0[t + 1];
            ++c.count;
            ++t;
        }
        biggest = std::max(biggest, c.end - c.begin);
        chunks.push_back(c);
    }
    if ((rc = st.ensure_samples(std::max<unsigned long long>(biggest, 1)))) return rc;
    // no more slots than the largest chunk has items: a small render does not walk idle slots
    const uint32_t want_slots =
        static_cast<uint32_t>(std::min<unsigned long long>(tu.path_slots, std::max<unsigned long long>(256, (biggest + 255) & ~255ull)));
    if ((rc = st.ensure_slots(want_slots, tu.queue_segments, tu.shade_queues == 1))) return rc;

    const bool multi = views.size() > 1;
    WaveLaunch L{};
    L.single = views[0];
    L.tabs.scenes = static_cast<SceneView*>(tab[0]);
    L.tabs.num_jobs = num_jobs;
    L.tabs.job_item0 = static_cast<unsigned long long*>(tab[1]);
    L.tabs.job_sample0 = static_cast<unsigned long long*>(tab[2]);
    L.tabs.job_scene = static_cast<uint32_t*>(tab[3]);
    L.tabs.job_accum = static_cast<float**>(tab[4]);
    L.tabs.job_w = static_cast<uint32_t*>(tab[5]);
    L.tabs.job_sout = static_cast<float4**>(tab[6]);
    L.tabs.job_sstride = static_cast<uint32_t*>(tab[7]);
    L.tabs.job_soff = static_cast<uint32_t*>(tab[8]);
    L.tabs.job_layer = static_cast<uint32_t*>(tab[9]);
    L.tabs.job_h = static_cast<uint32_t*>(tab[10]);
    L.tabs.job_spp = static_cast<uint32_t*>(tab[11]);
    L.tabs.job_tw = static_cast<uint32_t*>(tab[12]);
    L.tabs.job_th = static_cast<uint32_t*>(tab[13]);
    L.tabs.tile_job = static_cast<uint32_t*>(tab[14]);
    L.tabs.tile_x0 = static_cast<uint32_t*>(tab[15]);
    L.tabs.tile_y0 = static_cast<uint32_t*>(tab[16]);
    L.tabs.tile_w = static_cast<uint32_t*>(tab[17]);
    L.tabs.tile_h = static_cast<uint32_t*>(tab[18]);
    L.tabs.tile_item0 = static_cast<unsigned long long*>(tab[19]);
    L.tabs.num_tiles = num_tiles;
    L.buf = st.buf;
    L.buf.slots = want_slots;
    L.cfg = cfg;
    L.block = static_cast<int>(tu.block_size);
    L.lds_stack = static_cast<int>(tu.lds_stack);
    L.grid_traverse = persistent_grid(ctx.info.compute_units, L.block);
    L.grid_shade = persistent_grid(ctx.info.compute_units, kWaveBlock);
    L.multi = multi;
    L.shade_mask = shade_mask;
    L.wide = tu.bvh_width == 4 &&
             std::all_of(views.begin(), views.end(), [](const SceneView& v) { return v.nodes4 != nullptr; });
    HIP_OK(wave_set_grids(L, ctx.info.compute_units));
    L.stream = s;
    if ((rc = ctx.ensure_scratch(L.grid_traverse * L.block))) return rc;
    L.scratch = ctx.scratch;
    if (max_depth > static_cast<uint32_t>(L.lds_stack + ctx.scratch.spill_entries + 1))
        return set_error(SOLHIP_E_TOO_DEEP, "BVH deeper than the traversal stack");
    // a 4-wide level pushes up to three nodes
    if (L.wide && 3 * (max_depth4 > 0 ? max_depth4 - 1 : 0) > static_cast<uint32_t>(L.lds_stack + ctx.scratch.spill_entries))
        return set_error(SOLHIP_E_TOO_DEEP, "4-wide BVH deeper than the traversal stack");

    int iters = static_cast<int>(std::max<uint32_t>(2, tu.graph_iters & ~1u));
    // When every chunk fits in the paths in flight, its start fill takes all of its items and no
    // path restarts: each path ends within max_path_len bounces and the shadow rays of its last
    // bounce are traced one iteration later, so max_path_len + 1 iterations (rounded up to even)
    // finish a chunk in one replay, with no empty iterations after it.
    const bool one_fill = std::all_of(chunks.begin(), chunks.end(),
                                      [&](const Chunk& c) { return c.end - c.begin <= want_slots; });
    const uint64_t need = (static_cast<uint64_t>(cfg.max_path_len) + 2) & ~1ull;
    if (one_fill && need <= kMaxFillIters) iters = static_cast<int>(std::max<uint64_t>(2, need));
    // Such a chunk also runs in place: each path keeps one work item from start to end, so its
    // record (radiance, scene) is its item's entry of the sample buffer, named by the item's
    // offset in the chunk, and a path that ends needs no copy.
    if (one_fill) {
        L.buf.in_place = 1;
        L.buf.path = L.buf.samples;
    }
    // Profiling records HIP events between the kernels of every iteration. Events recorded by a
    // graph cannot be timed on this runtime, so a profiled run launches the same iterations
    // eagerly; otherwise the loop replays the captured graph.
    const bool profile = ctx.profiling;
    if (profile && (rc = st.ensure_events(static_cast<size_t>(iters) * 4))) return rc;
    const bool use_graph = !profile;
    if (use_graph) {
        std::string key = graph_key(L, st, iters);
        if (key != st.graph_key || !st.graph_exec) {
            if (st.graph_exec) HIP_OK(hipGraphExecDestroy(st.graph_exec));
            if (st.graph) HIP_OK(hipGraphDestroy(st.graph));
            st.graph_exec = nullptr;
            st.graph = nullptr;
            st.graph_key.clear();
            HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (int it = 0; it < iters && e == hipSuccess; ++it) e = enqueue_iteration(L, it & 1, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(s, &g);
            HIP_OK(e);
            HIP_OK(e2);
            st.graph = g;
            HIP_OK(hipGraphInstantiate(&st.graph_exec, g, nullptr, nullptr, 0));
            st.graph_key = key;
        }
    }

    // A chunk that fits in the paths in flight ends within one replay of `need` iterations (see
    // above): rather than a round trip to the host to see its queues empty, k_check_drained counts
    // what is left in them into a statistic read with the others at the end, which must be zero.
    // (A profiled run reads its events after each replay, so it keeps the round trip.)
    const bool drains_in_one = one_fill && iters == static_cast<int>(need) && !profile;
    // trace, sort and the shading kernels launched
    const uint64_t per_iteration = 2 + (L.buf.merged ? 1 : static_cast<uint64_t>(__builtin_popcount(L.shade_mask & 7u)));
    HIP_OK(hipMemsetAsync(st.buf.stats, 0, sizeof(unsigned long long) * kStatWords, s));
    HIP_OK(hipEventRecord(ctx.ev_begin, s));
    uint64_t replays = 0, launches = 0;
    for (const Chunk& c : chunks) {
        HIP_OK(wave_launch_set_chunk(L, c.begin, c.end));
        HIP_OK(wave_launch_start(L, 0));
        launches += 2;
        while (true) {
            if (use_graph) {
                HIP_OK(hipGraphLaunch(st.graph_exec, s));
            } else {
                for (int it = 0; it < iters; ++it)
                    HIP_OK(enqueue_iteration(L, it & 1, profile ? &st.events[it * 4] : nullptr));
            }
            ++replays;
            launches += static_cast<uint64_t>(iters) * per_iteration;
            if (drains_in_one) {
                HIP_OK(wave_launch_check_drained(L));
                ++launches;
                break;
            }
            // After an even number of iterations the next extension queue is queue 0; shadow rays
            // of the last shade still need a trace.
            std::vector<unsigned>& counts = st.host_counts;
            HIP_OK(hipMemcpyAsync(counts.data(), st.buf.counts, sizeof(unsigned) * kCounterWords, hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            if (profile) {
                for (int it = 0; it < iters; ++it) {
                    float ms[3];
                    for (int k = 0; k < 3; ++k)
                        HIP_OK(hipEventElapsedTime(&ms[k], st.events[it * 4 + k], st.events[it * 4 + k + 1]));
                    ctx.stats.trace_ms += ms[0];
                    ctx.stats.sort_ms += ms[1];
                    ctx.stats.shade_ms += ms[2];
                    ctx.stats.trace_launches += 1;
                }
            }
            unsigned pending = 0;
            for (int k = 0; k < kMaxSegments; ++k)
                pending += counts[qcount(kQExtend0, k)] + counts[qcount(kQShadow, k)];
            if (pending == 0) break;
        }
        for (uint32_t layer = 0; layer < num_layers; ++layer)
            HIP_OK(wave_launch_reduce(L, c.first, c.count, max_tile_pixels, c.begin, layer));
        launches += num_layers;
    }
    HIP_OK(hipEventRecord(ctx.ev_end, s));
    HIP_OK(hipMemcpyAsync(st.host_stats.data(), st.buf.stats, sizeof(unsigned long long) * kStatWords,
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    unsigned long long stats[3] = {0, 0, 0};
    for (int k = 0; k < kStatSlots; ++k)
        for (int i = 0; i < 3; ++i) stats[i] += st.host_stats[static_cast<size_t>(k) * kStatStride + i];
    if (st.host_stats[kStatUndrained] != 0)
        return set_error(SOLHIP_E_HIP, "a chunk of work did not end within its bounce iterations (" +
                                           std::to_string(st.host_stats[kStatUndrained]) + " queue entries left)");
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, ctx.ev_begin, ctx.ev_end));
    ctx.stats.render_ms += ms;
    ctx.stats.extension_rays += stats[0];
    ctx.stats.shadow_rays += stats[1];
    ctx.stats.samples += stats[2];
    ctx.stats.graph_replays += use_graph ? replays : 0;
    ctx.stats.launches += launches;
    return SOLHIP_OK;
}

// Material queues a scene's hits can enter: the queue of each of its materials' BSDFs, and the
// diffuse queue, which also takes the hits of lights without a BSDF. The kernels of the others
// are not launched (a scene of diffuse materials shades with one kernel a bounce).
uint32_t shade_mask_of(const solhip_scene& s) {
    uint32_t mask = 1u << kQueueDiffuse;
    for (const sol_material& m : s.materials_host)
        if (m.type != SOL_BSDF_NONE) mask |= 1u << queue_of(m.type);
    return mask;
}

} // namespace

int render_wavefront(solhip_scene& s, const sol_pt_config& cfg, uint32_t w, uint32_t h, uint64_t si, uint32_t sc,
                     float* d_accum, float4* d_samples) {
    if (cfg.max_path_len == 0) return SOLHIP_OK;
    std::vector<SceneView> views{s.view};
    std::vector<Job> jobs{Job{0, w, h, si, sc, d_accum, d_samples}};
    return run_wavefront(*s.ctx, views, jobs, cfg, s.bvh_depth, s.bvh4_depth, shade_mask_of(s));
}

} // namespace solhip

using namespace solhip;

extern "C" {

int solhip_render_batch(solhip_ctx* ctx, const solhip_batch_job* jobs, uint32_t num_jobs, const sol_pt_config* cfg) {
    if (!ctx || !cfg || (num_jobs && !jobs)) return set_error(SOLHIP_E_INVALID, "null argument");
    if (cfg->min_survival_prob < 0 || cfg->max_survival_prob > 1 || cfg->min_survival_prob > cfg->max_survival_prob)
        return set_error(SOLHIP_E_INVALID, "survival probabilities must satisfy 0 <= min <= max <= 1");
    std::vector<SceneView> views;
    std::unordered_map<const solhip_scene*, uint32_t> index;
    std::vector<Job> list;
    uint32_t depth = 0, depth4 = 0, shade_mask = 0;
    for (uint32_t i = 0; i < num_jobs; ++i) {
        const solhip_batch_job& j = jobs[i];
        if (!j.scene || !j.accum || j.width == 0 || j.height == 0) return set_error(SOLHIP_E_INVALID, "invalid batch job");
        if (j.scene->ctx != ctx) return set_error(SOLHIP_E_CONTEXT, "scene belongs to another context");
        if (!j.scene->num_lights) return set_error(SOLHIP_E_NO_LIGHTS, "scene without light cannot be rendered");
        if (j.sample_count == 0) continue;
        auto it = index.find(j.scene);
        uint32_t si;
        if (it == index.end()) {
            si = static_cast<uint32_t>(views.size());
            index[j.scene] = si;
            views.push_back(j.scene->view);
            depth = std::max(depth, j.scene->bvh_depth);
            depth4 = std::max(depth4, j.scene->bvh4_depth);
            shade_mask |= shade_mask_of(*j.scene);
        } else {
            si = it->second;
        }
        list.push_back(Job{si, j.width, j.height, j.sample_index, j.sample_count, j.accum});
    }
    if (list.empty() || cfg->max_path_len == 0) return SOLHIP_OK;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return set_error(SOLHIP_E_HIP, hipGetErrorString(e));
    int rc = run_wavefront(*ctx, views, list, *cfg, depth, depth4, shade_mask);
    if (rc) return rc;
    return ctx->check_overflow();
}

} // extern "C"

namespace {

// Coordinate descent over the wavefront launch parameters from the context's current ones: paths in
// flight (candidates above `items`, the work of one trial, only add idle slots and are left out),
// traversal workgroup size, LDS stack depth, bounce iterations per graph, BVH node width, the queue
// split (segments per queue, material-split or merged shading queues) and tile shape. `trial` renders with ctx.tuning and returns its duration in ms (the first call of a
// candidate builds the slot arena and captures the graph, so each candidate runs twice and the
// second run counts). A candidate replaces the current best if it is at least 1 % faster. The
// statistics and the profiling switch of the context are restored; ctx.tuning ends as the best.
template <typename Trial>
int tune(Context& ctx, unsigned long long items, Trial trial, solhip_tuning* out) {
    const solhip_stats saved_stats = ctx.stats;
    const bool saved_profiling = ctx.profiling;
    ctx.profiling = false;
    const solhip_tuning start = ctx.tuning;
    auto timed = [&](const solhip_tuning& t, float& ms) -> int {
        ctx.tuning = t;
        int rc = trial(ms);
        if (rc) return rc;
        return trial(ms);
    };
    const uint32_t max_slots = default_path_slots(ctx.info.total_memory);
    std::vector<uint32_t> slot_cands;
    for (uint32_t c : {1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26, 1u << 27}) {
        if (c > max_slots) continue;
        uint32_t v = static_cast<uint32_t>(std::min<unsigned long long>(c, std::max<unsigned long long>(4096, items)));
        if (std::find(slot_cands.begin(), slot_cands.end(), v) == slot_cands.end()) slot_cands.push_back(v);
    }
    solhip_tuning best = start;
    float best_ms = 0.0f;
    int rc = timed(best, best_ms);
    for (int dim = 0; rc == SOLHIP_OK && dim < 8; ++dim) {
        std::vector<solhip_tuning> cands;
        solhip_tuning t = best;
        switch (dim) {
            case 0:   // with chunks of as many items (every chunk then runs in place)
                for (uint32_t v : slot_cands) {
                    t.path_slots = v;
                    t.chunk_items = default_chunk_items(v);
                    cands.push_back(t);
                }
                break;
            case 1: for (uint32_t v : {128u, 256u}) { t.block_size = v; cands.push_back(t); } break;
            case 2: for (uint32_t v : {8u, 16u}) { t.lds_stack = v; cands.push_back(t); } break;
            case 3: for (uint32_t v : {8u, 16u, 32u}) { t.graph_iters = v; cands.push_back(t); } break;
            case 4: for (uint32_t v : {2u, 4u}) { t.bvh_width = v; cands.push_back(t); } break;
            case 5: for (uint32_t v : {1u, 2u, 4u, 8u}) { t.queue_segments = v; cands.push_back(t); } break;
            case 6: for (uint32_t v : {1u, 3u}) { t.shade_queues = v; cands.push_back(t); } break;
            default:
                for (auto wh : {std::pair<uint32_t, uint32_t>{16, 16}, {32, 32}, {64, 16}, {128, 8}}) {
                    t.tile_w = wh.first;
                    t.tile_h = wh.second;
                    cands.push_back(t);
                }
        }
        for (const solhip_tuning& c : cands) {
            if (std::memcmp(&c, &best, sizeof(c)) == 0) continue;
            float ms = 0.0f;
            if ((rc = timed(c, ms)) != SOLHIP_OK) break;
            if (ms < 0.99f * best_ms) {
                best = c;
                best_ms = ms;
            }
        }
    }
    ctx.tuning = rc == SOLHIP_OK ? best : start;
    ctx.stats = saved_stats;
    ctx.profiling = saved_profiling;
    if (rc) return rc;
    if (out) *out = best;
    return SOLHIP_OK;
}

// Duration in ms of fn(), measured with the context's HIP events on its stream.
template <typename Fn>
int event_time(Context& ctx, Fn fn, float& ms) {
    HIP_OK(hipEventRecord(ctx.ev_begin, ctx.stream));
    int rc = fn();
    if (rc) return rc;
    HIP_OK(hipEventRecord(ctx.ev_end, ctx.stream));
    HIP_OK(hipEventSynchronize(ctx.ev_end));
    HIP_OK(hipEventElapsedTime(&ms, ctx.ev_begin, ctx.ev_end));
    return SOLHIP_OK;
}

bool valid_survival(const sol_pt_config& c) {
    return c.min_survival_prob >= 0 && c.max_survival_prob <= 1 && c.min_survival_prob <= c.max_survival_prob;
}

} // namespace

extern "C" {

// Launch-parameter search on one scene and image size: every trial renders budget_spp samples per
// pixel of the image into a scratch image.
int solhip_autotune(solhip_scene* s, const sol_pt_config* cfg, uint32_t width, uint32_t height, uint32_t budget_spp,
                    solhip_tuning* out) {
    if (!s || !cfg) return set_error(SOLHIP_E_INVALID, "null argument");
    if (width == 0 || height == 0) return set_error(SOLHIP_E_INVALID, "empty image");
    if (!valid_survival(*cfg)) return set_error(SOLHIP_E_INVALID, "survival probabilities must satisfy 0 <= min <= max <= 1");
    if (!s->num_lights) return set_error(SOLHIP_E_NO_LIGHTS, "scene without light cannot be rendered");
    if (budget_spp == 0) budget_spp = 1;
    Context& ctx = *s->ctx;
    HIP_OK(hipSetDevice(ctx.device));
    const size_t plane = static_cast<size_t>(width) * height;
    DeviceBuffer img;
    HIP_OK(img.alloc(sizeof(float) * 3 * plane));
    auto trial = [&](float& ms) {
        return event_time(ctx, [&] { return render_wavefront(*s, *cfg, width, height, 0, budget_spp, img.as<float>()); },
                          ms);
    };
    return tune(ctx, static_cast<unsigned long long>(plane) * budget_spp, trial, out);
}

// Launch-parameter search on a batch: every trial is the whole solhip_render_batch of the jobs, with
// their images replaced by scratch images (the jobs' own images are not touched).
int solhip_autotune_batch(solhip_ctx* ctx, const solhip_batch_job* jobs, uint32_t num_jobs, const sol_pt_config* cfg,
                          solhip_tuning* out) {
    if (!ctx || !cfg || !jobs || num_jobs == 0) return set_error(SOLHIP_E_INVALID, "null argument or no job");
    if (!valid_survival(*cfg)) return set_error(SOLHIP_E_INVALID, "survival probabilities must satisfy 0 <= min <= max <= 1");
    size_t floats = 0;
    unsigned long long items = 0;
    for (uint32_t i = 0; i < num_jobs; ++i) {
        const solhip_batch_job& j = jobs[i];
        if (!j.scene || j.width == 0 || j.height == 0) return set_error(SOLHIP_E_INVALID, "invalid batch job");
        if (j.scene->ctx != ctx) return set_error(SOLHIP_E_CONTEXT, "scene belongs to another context");
        floats += 3 * static_cast<size_t>(j.width) * j.height;
        items += static_cast<unsigned long long>(j.width) * j.height * j.sample_count;
    }
    HIP_OK(hipSetDevice(ctx->device));
    DeviceBuffer img;
    HIP_OK(img.alloc(sizeof(float) * floats));
    std::vector<solhip_batch_job> local(jobs, jobs + num_jobs);
    float* p = img.as<float>();
    for (solhip_batch_job& j : local) {
        j.accum = p;
        p += 3 * static_cast<size_t>(j.width) * j.height;
    }
    auto trial = [&](float& ms) {
        return event_time(*ctx, [&] { return solhip_render_batch(ctx, local.data(), num_jobs, cfg); }, ms);
    };
    return tune(*ctx, items, trial, out);
}

} // extern "C"
