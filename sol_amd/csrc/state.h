// Host-side state behind the opaque handles of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "solhip.h"
#include "launch.h"

namespace solhip {

// Spill entries per lane behind the LDS short stack.
constexpr int kSpillEntries = 88;
constexpr int kNumCounters = 16;

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { if (ptr) (void)hipFree(ptr); }
    hipError_t alloc(size_t n) {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, n ? n : 1);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    // Grows to at least n bytes (contents are not preserved).
    hipError_t reserve(size_t n) { return n <= bytes && ptr ? hipSuccess : alloc(n); }
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

struct WavefrontState;  // defined in wavefront.hip

// Device memory of the wavefront state per path slot: the slot record and the hit (32 B) and the
// queue segments (two extension queues 256 B, three shade queues 12 B, the shadow queue 48 B).
constexpr unsigned long long kPathSlotBytes = 348;

// Paths in flight by default: the largest power of two up to 2^27 whose state takes at most a
// quarter of the device's memory (2^27 slots, 47 GB, on a 288 GB MI355X: a 1080p 64-spp frame is
// then in flight at once; 2^24 slots where the memory size is not known).
inline uint32_t default_path_slots(unsigned long long device_memory) {
    if (device_memory == 0) return 1u << 24;
    uint32_t s = 1u << 27;
    while (s > 4096u && static_cast<unsigned long long>(s) * kPathSlotBytes > device_memory / 4) s >>= 1;
    return s;
}

// Work items per chunk at most by default (the per-sample radiance buffer, 16 B an item: 3 GiB).
constexpr uint32_t kDefaultChunkItems = 192u << 20;

// Default chunk size for `path_slots` paths in flight: as many items as paths (within the bounds
// of chunk_items), so that every chunk fits in the start fill and runs in place - its paths keep
// one item each and their records are the sample entries.
inline uint32_t default_chunk_items(uint32_t path_slots) {
    return std::min<uint32_t>(kDefaultChunkItems, std::max<uint32_t>(path_slots, 1u << 16));
}

inline solhip_tuning default_tuning(unsigned long long device_memory) {
    solhip_tuning t{};
    t.block_size = 256;
    t.lds_stack = 16;
    t.path_slots = default_path_slots(device_memory);
    t.tile_w = 32;
    t.tile_h = 32;
    t.graph_iters = 16;
    t.bvh_width = 4;
    t.chunk_items = default_chunk_items(t.path_slots);
    t.queue_segments = 1;
    t.shade_queues = 3;
    return t;
}

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    solhip_device_info info{};
    solhip_tuning tuning{};
    bool profiling = false;
    solhip_stats stats{};
    unsigned long long* counters = nullptr;  // kNumCounters device counters
    TraversalScratch scratch{};
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    WavefrontState* wave = nullptr;
    DeviceBuffer adj_radiance, adj_image;   // forward pass of solhip_backward (grow-only)
    DeviceBuffer isect_rays, isect_out;     // host-pointer staging of solhip_intersect (grow-only)
    DeviceBuffer isect_counters;            // work counters of the intersection kernel (launch_intersect)
    DeviceBuffer host_image;                // device image of solhip_render for host images (grow-only)
    DeviceBuffer build_ws;                  // temporaries of solhip_scene_build (grow-only)
    void* pinned = nullptr;                 // pinned host staging (grow-only)
    size_t pinned_bytes = 0;
    int live_scenes = 0;             // scenes uploaded and not yet destroyed
    bool destroy_requested = false;  // solhip_destroy was called while scenes were alive

    int ensure_scratch(int threads);
    LaunchShape shape(int block, int lds) const;
    int check_overflow();
    int reserve_pinned(size_t bytes);
    int read_counters(unsigned long long* host, int n);
    void release_wavefront();
};

int set_error(int code, const std::string& msg);
// Looks up the device address of solhip_render_noalloc's scratch array on `device` (current).
int noalloc_setup(int device);

} // namespace solhip

struct solhip_ctx : solhip::Context {};

struct solhip_scene {
    solhip::Context* ctx = nullptr;
    sol_gpu_node* nodes = nullptr;
    sol_gpu_node4* nodes4 = nullptr;   // may be null (the wavefront traversal then uses nodes)
    sol_gpu_tri* tris = nullptr;
    sol_gpu_shade* shade = nullptr;
    sol_material* materials = nullptr;
    sol_light* lights = nullptr;
    uint32_t num_nodes = 0, num_tris = 0, num_materials = 0, num_lights = 0, bvh_depth = 0, bvh4_depth = 0;
    std::vector<sol_material> materials_host;
    std::vector<sol_light> lights_host;
    solhip::SceneView view{};
};

namespace solhip {
// GPU BVH construction (gpu_build.hip); the mesh and configuration are validated by the caller.
int build_scene(Context& ctx, const solhip_mesh_desc& mesh, const solhip_build_config& cfg, solhip_scene** out,
                solhip_build_stats* stats);
int render_megakernel(solhip_scene& s, const sol_pt_config& cfg, uint32_t w, uint32_t h, uint64_t si, uint32_t sc,
                      float* d_accum);
// d_samples (optional): each sample's radiance, pixel-major, sc per pixel.
int render_wavefront(solhip_scene& s, const sol_pt_config& cfg, uint32_t w, uint32_t h, uint64_t si, uint32_t sc,
                     float* d_accum, float4* d_samples = nullptr);
} // namespace solhip
