// C ABI of libsolhip.so (include/solhip.h): contexts, scene upload, and the entry points that
// launch the kernels.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "solhip.h"
#include "launch.h"
#include "state.h"

using namespace solhip;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e__ = (expr);                                                                 \
        if (e__ != hipSuccess) {                                                                 \
            (void)hipGetLastError();                                                             \
            return fail(SOLHIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e__));        \
        }                                                                                        \
    } while (0)

// BSDF kind of every shading record, and every triangle's kind word (which the traversal kernel
// packs into its hits), from the current material table.
__global__ void k_shade_kinds(sol_gpu_shade* shade, sol_gpu_tri* tris, uint32_t n, const sol_material* mats) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int m = shade[i].material;
    const int bsdf = m >= 0 ? mats[m].type : SOL_BSDF_NONE;
    shade[i].bsdf = bsdf;
    tris[i].pad0 = tri_kind(bsdf, shade[i].light);
}

int refresh_kinds(solhip_scene* s) {
    hipLaunchKernelGGL(k_shade_kinds, dim3((s->num_tris + 255) / 256), dim3(256), 0, s->ctx->stream, s->shade,
                       s->tris, s->num_tris, s->materials);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    return SOLHIP_OK;
}

bool valid_shape_values(uint32_t block, uint32_t lds) {
    return (block == 64 || block == 128 || block == 256) && (lds == 8 || lds == 16 || lds == 32);
}

} // namespace

namespace solhip {

int set_error(int code, const std::string& msg) { return fail(code, msg); }

int Context::ensure_scratch(int threads) {
    if (threads <= scratch.max_threads && scratch.spill) return SOLHIP_OK;
    if (scratch.spill) HIP_TRY(hipFree(scratch.spill));
    scratch.spill = nullptr;
    size_t bytes = static_cast<size_t>(threads) * kSpillEntries * sizeof(int);
    HIP_TRY(hipMalloc(&scratch.spill, bytes));
    scratch.max_threads = threads;
    scratch.spill_entries = kSpillEntries;
    return SOLHIP_OK;
}

LaunchShape Context::shape(int block, int lds) const {
    LaunchShape s;
    s.block = block;
    s.lds_stack = lds;
    s.grid = persistent_grid(info.compute_units, block);
    s.wide = tuning.bvh_width == 4;
    return s;
}

int Context::reserve_pinned(size_t bytes) {
    if (pinned && bytes <= pinned_bytes) return SOLHIP_OK;
    if (pinned) HIP_TRY(hipHostFree(pinned));
    pinned = nullptr;
    pinned_bytes = 0;
    HIP_TRY(hipHostMalloc(&pinned, bytes, hipHostMallocDefault));
    pinned_bytes = bytes;
    return SOLHIP_OK;
}

int Context::check_overflow() {
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, scratch.overflow, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (flag) {
        int zero = 0;
        HIP_TRY(hipMemcpy(scratch.overflow, &zero, sizeof(int), hipMemcpyHostToDevice));
        return fail(SOLHIP_E_TOO_DEEP, "traversal stack overflow");
    }
    return SOLHIP_OK;
}

int Context::read_counters(unsigned long long* host, int n) {
    HIP_TRY(hipMemcpyAsync(host, counters, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SOLHIP_OK;
}

} // namespace solhip

extern "C" {

const char* solhip_last_error(void) { return g_error.c_str(); }

int solhip_create(int device, solhip_ctx** out) {
    if (!out) return fail(SOLHIP_E_INVALID, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(SOLHIP_E_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(SOLHIP_E_NO_DEVICE, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    if (int st = noalloc_setup(device)) return st;
    auto ctx = std::make_unique<solhip_ctx>();
    ctx->device = device;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::snprintf(ctx->info.name, sizeof(ctx->info.name), "%s", prop.name);
    std::snprintf(ctx->info.arch, sizeof(ctx->info.arch), "%s", prop.gcnArchName);
    ctx->info.compute_units = prop.multiProcessorCount;
    ctx->info.wavefront_size = prop.warpSize;
    ctx->info.total_memory = prop.totalGlobalMem;
    ctx->info.lds_per_block = static_cast<int32_t>(prop.sharedMemPerBlock);
    if (prop.warpSize != 64) return fail(SOLHIP_E_NO_DEVICE, "wave64 device required");
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&ctx->counters, sizeof(unsigned long long) * kNumCounters));
    HIP_TRY(hipMalloc(&ctx->scratch.overflow, sizeof(int)));
    HIP_TRY(hipMemset(ctx->scratch.overflow, 0, sizeof(int)));
    HIP_TRY(hipEventCreate(&ctx->ev_begin));
    HIP_TRY(hipEventCreate(&ctx->ev_end));
    ctx->tuning = default_tuning(ctx->info.total_memory);
    int st = ctx->ensure_scratch(persistent_grid(ctx->info.compute_units, 64) * 64);
    if (st) return st;
    *out = ctx.release();
    return SOLHIP_OK;
}

namespace {
void destroy_context(solhip_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->release_wavefront();
    if (ctx->scratch.spill) (void)hipFree(ctx->scratch.spill);
    if (ctx->scratch.overflow) (void)hipFree(ctx->scratch.overflow);
    if (ctx->counters) (void)hipFree(ctx->counters);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    (void)hipEventDestroy(ctx->ev_begin);
    (void)hipEventDestroy(ctx->ev_end);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}
} // namespace

// A context outlives its scenes: destroying it while scenes are alive defers the release to the
// destruction of the last scene.
void solhip_destroy(solhip_ctx* ctx) {
    if (!ctx) return;
    if (ctx->live_scenes > 0) {
        ctx->destroy_requested = true;
        return;
    }
    destroy_context(ctx);
}

int solhip_device_query(solhip_ctx* ctx, solhip_device_info* out) {
    if (!ctx || !out) return fail(SOLHIP_E_INVALID, "null argument");
    *out = ctx->info;
    return SOLHIP_OK;
}

int solhip_set_tuning(solhip_ctx* ctx, const solhip_tuning* t) {
    if (!ctx || !t) return fail(SOLHIP_E_INVALID, "null argument");
    solhip_tuning d = default_tuning(ctx->info.total_memory);
    solhip_tuning n = *t;
    if (!n.block_size) n.block_size = d.block_size;
    if (!n.lds_stack) n.lds_stack = d.lds_stack;
    if (!n.path_slots) n.path_slots = d.path_slots;
    if (!n.tile_w) n.tile_w = d.tile_w;
    if (!n.tile_h) n.tile_h = d.tile_h;
    if (!n.graph_iters) n.graph_iters = d.graph_iters;
    if (!n.bvh_width) n.bvh_width = d.bvh_width;
    if (!n.chunk_items) n.chunk_items = default_chunk_items(n.path_slots);
    if (!n.queue_segments) n.queue_segments = d.queue_segments;
    if (!n.shade_queues) n.shade_queues = d.shade_queues;
    if (n.chunk_items < (1u << 16) || n.chunk_items > (1u << 28)) return fail(SOLHIP_E_INVALID, "chunk_items out of range");
    if (n.queue_segments > 8 || (n.queue_segments & (n.queue_segments - 1)))
        return fail(SOLHIP_E_INVALID, "queue_segments must be 1, 2, 4 or 8");
    if (n.shade_queues != 1 && n.shade_queues != 3) return fail(SOLHIP_E_INVALID, "shade_queues must be 1 or 3");
    if (n.bvh_width != 2 && n.bvh_width != 4) return fail(SOLHIP_E_INVALID, "bvh_width must be 2 or 4");
    if (!valid_shape_values(n.block_size, n.lds_stack)) return fail(SOLHIP_E_INVALID, "unsupported block size or LDS stack");
    if (n.path_slots < 4096 || n.path_slots > (1u << 27)) return fail(SOLHIP_E_INVALID, "path_slots out of range");
    if (n.graph_iters > 256) return fail(SOLHIP_E_INVALID, "graph_iters out of range");
    ctx->tuning = n;
    return SOLHIP_OK;
}

int solhip_get_tuning(solhip_ctx* ctx, solhip_tuning* out) {
    if (!ctx || !out) return fail(SOLHIP_E_INVALID, "null argument");
    *out = ctx->tuning;
    return SOLHIP_OK;
}

int solhip_set_profiling(solhip_ctx* ctx, int profiling) {
    if (!ctx) return fail(SOLHIP_E_INVALID, "null context");
    ctx->profiling = profiling != 0;
    return SOLHIP_OK;
}

int solhip_get_stats(solhip_ctx* ctx, solhip_stats* out) {
    if (!ctx || !out) return fail(SOLHIP_E_INVALID, "null argument");
    *out = ctx->stats;
    return SOLHIP_OK;
}

int solhip_reset_stats(solhip_ctx* ctx) {
    if (!ctx) return fail(SOLHIP_E_INVALID, "null context");
    ctx->stats = solhip_stats{};
    return SOLHIP_OK;
}

// Known BSDF kinds; a mix weight outside [0, 1] would make the lobe pick and the lerp meaningless.
static const char* check_materials(const sol_material* m, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        if (m[i].type < SOL_BSDF_DIFFUSE || m[i].type > SOL_BSDF_INTERP) return "unknown material type";
        if (m[i].type == SOL_BSDF_INTERP && !(m[i].mix >= 0.0f && m[i].mix <= 1.0f))
            return "mixed material weight outside [0, 1]";
    }
    return nullptr;
}

int solhip_scene_upload(solhip_ctx* ctx, const solhip_scene_desc* d, solhip_scene** out) {
    if (!ctx || !d || !out) return fail(SOLHIP_E_INVALID, "null argument");
    *out = nullptr;
    if (!d->num_nodes || !d->nodes || !d->num_tris || !d->tris || !d->shade)
        return fail(SOLHIP_E_INVALID, "scene without geometry");
    // a scene without light serves intersection queries; rendering it is refused (SOLHIP_E_NO_LIGHTS)
    if (d->num_lights && !d->lights) return fail(SOLHIP_E_INVALID, "null light table");
    if (d->num_materials && !d->materials) return fail(SOLHIP_E_INVALID, "null material table");
    if (const char* e = check_materials(d->materials, d->num_materials)) return fail(SOLHIP_E_INVALID, e);
    // a hit carries the triangle index below kHitKindShift bits (and the kind above)
    if (d->num_tris >= (1u << kHitKindShift)) return fail(SOLHIP_E_INVALID, "too many triangles");
    // traversal addresses nodes and triangles with 32-bit byte offsets
    if (static_cast<uint64_t>(d->num_tris) * sizeof(sol_gpu_tri) > 0xFFFFFFFFull ||
        static_cast<uint64_t>(d->num_nodes) * sizeof(sol_gpu_node) > 0xFFFFFFFFull)
        return fail(SOLHIP_E_INVALID, "scene larger than 4 GiB of nodes or triangles");
    if (d->bvh_depth > static_cast<uint32_t>(8 + kSpillEntries))
        return fail(SOLHIP_E_TOO_DEEP, "BVH depth " + std::to_string(d->bvh_depth) + " exceeds the traversal stack");
    for (uint32_t i = 0; i < d->num_tris; ++i) {
        int m = d->shade[i].material, l = d->shade[i].light;
        if (m >= static_cast<int>(d->num_materials) || m < -1 || l >= static_cast<int>(d->num_lights) || l < -1)
            return fail(SOLHIP_E_INVALID, "material or light index out of range");
    }
    if (d->num_nodes4 && !d->nodes4) return fail(SOLHIP_E_INVALID, "null 4-wide node array");
    if (static_cast<uint64_t>(d->num_nodes4) * sizeof(sol_gpu_node4) > 0xFFFFFFFFull)
        return fail(SOLHIP_E_INVALID, "scene larger than 4 GiB of nodes");
    for (uint32_t i = 0; i < d->num_nodes4; ++i) {
        int used = 0;
        for (int c = 0; c < 4; ++c) {
            int32_t ref = d->nodes4[i].child[c];
            if (ref == SOL_GPU_EMPTY) continue;
            ++used;
            if (ref >= 0) {
                if (static_cast<uint32_t>(ref) >= d->num_nodes4) return fail(SOLHIP_E_INVALID, "4-wide node index out of range");
            } else {
                uint32_t code = static_cast<uint32_t>(~ref);
                if ((code >> 3) + (code & 7u) + 1u > d->num_tris) return fail(SOLHIP_E_INVALID, "leaf range out of range");
            }
        }
        if (used == 0) return fail(SOLHIP_E_INVALID, "4-wide node without children");
    }
    // a 4-wide level pushes at most three nodes on the traversal stack
    if (d->num_nodes4 && 3 * (d->bvh4_depth > 0 ? d->bvh4_depth - 1 : 0) > static_cast<uint32_t>(8 + kSpillEntries))
        return fail(SOLHIP_E_TOO_DEEP, "4-wide BVH depth " + std::to_string(d->bvh4_depth) + " exceeds the traversal stack");
    for (uint32_t i = 0; i < d->num_nodes; ++i) {
        for (int c = 0; c < 2; ++c) {
            int32_t ref = d->nodes[i].child[c];
            if (ref >= 0) {
                if (static_cast<uint32_t>(ref) >= d->num_nodes) return fail(SOLHIP_E_INVALID, "node index out of range");
            } else {
                uint32_t code = static_cast<uint32_t>(~ref);
                if ((code >> 3) + (code & 7u) + 1u > d->num_tris) return fail(SOLHIP_E_INVALID, "leaf range out of range");
            }
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    auto s = std::make_unique<solhip_scene>();
    s->ctx = ctx;
    s->num_nodes = d->num_nodes;
    s->num_tris = d->num_tris;
    s->num_materials = d->num_materials;
    s->num_lights = d->num_lights;
    s->bvh_depth = d->bvh_depth;
    s->bvh4_depth = d->num_nodes4 ? d->bvh4_depth : 0;
    s->materials_host.assign(d->materials, d->materials + d->num_materials);
    s->lights_host.assign(d->lights, d->lights + d->num_lights);
    HIP_TRY(hipMalloc(&s->nodes, sizeof(sol_gpu_node) * d->num_nodes));
    if (d->num_nodes4) {
        // unused slots get the inverted box, which every ray misses (the traversal tests it)
        std::vector<sol_gpu_node4> n4(d->nodes4, d->nodes4 + d->num_nodes4);
        const float inf = std::numeric_limits<float>::infinity();
        for (sol_gpu_node4& g : n4)
            for (int c = 0; c < 4; ++c)
                if (g.child[c] == SOL_GPU_EMPTY) {
                    g.lo_x[c] = g.lo_y[c] = g.lo_z[c] = inf;
                    g.hi_x[c] = g.hi_y[c] = g.hi_z[c] = -inf;
                }
        HIP_TRY(hipMalloc(&s->nodes4, sizeof(sol_gpu_node4) * d->num_nodes4));
        HIP_TRY(hipMemcpy(s->nodes4, n4.data(), sizeof(sol_gpu_node4) * d->num_nodes4, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&s->tris, sizeof(sol_gpu_tri) * d->num_tris));
    HIP_TRY(hipMalloc(&s->shade, sizeof(sol_gpu_shade) * d->num_tris));
    HIP_TRY(hipMalloc(&s->materials, sizeof(sol_material) * std::max<uint32_t>(1, d->num_materials)));
    HIP_TRY(hipMalloc(&s->lights, sizeof(sol_light) * std::max<uint32_t>(1, d->num_lights)));
    HIP_TRY(hipMemcpy(s->nodes, d->nodes, sizeof(sol_gpu_node) * d->num_nodes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->tris, d->tris, sizeof(sol_gpu_tri) * d->num_tris, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->shade, d->shade, sizeof(sol_gpu_shade) * d->num_tris, hipMemcpyHostToDevice));
    if (d->num_materials)
        HIP_TRY(hipMemcpy(s->materials, d->materials, sizeof(sol_material) * d->num_materials, hipMemcpyHostToDevice));
    if (d->num_lights)
        HIP_TRY(hipMemcpy(s->lights, d->lights, sizeof(sol_light) * d->num_lights, hipMemcpyHostToDevice));
    SceneView& v = s->view;
    v.nodes = reinterpret_cast<const float4*>(s->nodes);
    v.nodes4 = reinterpret_cast<const float4*>(s->nodes4);
    v.num_nodes4 = d->num_nodes4;
    v.top_nodes4 = std::min<uint32_t>(d->top_nodes4, d->num_nodes4);
    v.tris = reinterpret_cast<const float4*>(s->tris);
    v.shade = reinterpret_cast<const float4*>(s->shade);
    v.materials = s->materials;
    v.lights = s->lights;
    v.num_lights = d->num_lights;
    v.top_nodes = std::min<uint32_t>(d->top_nodes, d->num_nodes);
    v.num_nodes = d->num_nodes;
    v.num_tris = d->num_tris;
    v.camera = d->camera;
    if (int st = refresh_kinds(s.get())) return st;
    ctx->live_scenes += 1;
    *out = s.release();
    return SOLHIP_OK;
}

void solhip_scene_destroy(solhip_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(s->nodes);
    if (s->nodes4) (void)hipFree(s->nodes4);
    (void)hipFree(s->tris);
    (void)hipFree(s->shade);
    (void)hipFree(s->materials);
    (void)hipFree(s->lights);
    solhip_ctx* ctx = static_cast<solhip_ctx*>(s->ctx);
    delete s;
    ctx->live_scenes -= 1;
    if (ctx->live_scenes == 0 && ctx->destroy_requested) destroy_context(ctx);
}

int solhip_scene_build(solhip_ctx* ctx, const solhip_mesh_desc* m, const solhip_build_config* cfg, solhip_scene** out,
                       solhip_build_stats* stats) {
    if (!ctx || !m || !out) return fail(SOLHIP_E_INVALID, "null argument");
    *out = nullptr;
    if (!m->num_tris || !m->num_vertices || !m->positions || !m->normals || !m->indices || !m->tri_material ||
        !m->tri_light)
        return fail(SOLHIP_E_INVALID, "mesh without geometry");
    if (!m->num_lights || !m->lights) return fail(SOLHIP_E_NO_LIGHTS, "scene without light");
    if (m->num_materials && !m->materials) return fail(SOLHIP_E_INVALID, "null material table");
    if (const char* e = check_materials(m->materials, m->num_materials)) return fail(SOLHIP_E_INVALID, e);
    if (m->num_tris >= (1u << kHitKindShift)) return fail(SOLHIP_E_INVALID, "too many triangles");
    for (uint64_t i = 0; i < 3ull * m->num_tris; ++i)
        if (m->indices[i] >= m->num_vertices) return fail(SOLHIP_E_INVALID, "vertex index out of range");
    for (uint32_t t = 0; t < m->num_tris; ++t) {
        int mi = m->tri_material[t], li = m->tri_light[t];
        if (mi >= static_cast<int>(m->num_materials) || mi < -1 || li >= static_cast<int>(m->num_lights) || li < -1)
            return fail(SOLHIP_E_INVALID, "material or light index out of range");
    }
    solhip_build_config c{8, 1.0f, 1.0f, 16};
    if (cfg) {
        if (cfg->max_leaf_size) c.max_leaf_size = cfg->max_leaf_size;
        if (cfg->traversal_cost > 0.0f) c.traversal_cost = cfg->traversal_cost;
        if (cfg->intersection_cost > 0.0f) c.intersection_cost = cfg->intersection_cost;
        if (cfg->search_radius) c.search_radius = cfg->search_radius;
    }
    if (c.max_leaf_size > 8 || c.search_radius > 32) return fail(SOLHIP_E_INVALID, "build configuration out of range");
    if (stats) *stats = solhip_build_stats{};
    return build_scene(*ctx, *m, c, out, stats);
}

int solhip_scene_info_get(const solhip_scene* s, solhip_scene_info* out) {
    if (!s || !out) return fail(SOLHIP_E_INVALID, "null argument");
    out->num_nodes = s->num_nodes;
    out->num_nodes4 = s->view.num_nodes4;
    out->num_tris = s->num_tris;
    out->num_materials = s->num_materials;
    out->num_lights = s->num_lights;
    out->bvh_depth = s->bvh_depth;
    out->bvh4_depth = s->bvh4_depth;
    out->top_nodes = s->view.top_nodes;
    out->top_nodes4 = s->view.top_nodes4;
    return SOLHIP_OK;
}

int solhip_scene_download(const solhip_scene* s, sol_gpu_node* nodes, sol_gpu_node4* nodes4, sol_gpu_tri* tris,
                          sol_gpu_shade* shade) {
    if (!s) return fail(SOLHIP_E_INVALID, "null scene");
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    if (nodes) HIP_TRY(hipMemcpy(nodes, s->nodes, sizeof(sol_gpu_node) * s->num_nodes, hipMemcpyDeviceToHost));
    if (nodes4 && s->nodes4)
        HIP_TRY(hipMemcpy(nodes4, s->nodes4, sizeof(sol_gpu_node4) * s->view.num_nodes4, hipMemcpyDeviceToHost));
    if (tris) HIP_TRY(hipMemcpy(tris, s->tris, sizeof(sol_gpu_tri) * s->num_tris, hipMemcpyDeviceToHost));
    if (shade) HIP_TRY(hipMemcpy(shade, s->shade, sizeof(sol_gpu_shade) * s->num_tris, hipMemcpyDeviceToHost));
    return SOLHIP_OK;
}

int solhip_scene_set_materials(solhip_scene* s, const sol_material* m, uint32_t n) {
    if (!s || !m || n != s->num_materials) return fail(SOLHIP_E_INVALID, "material count mismatch");
    if (const char* e = check_materials(m, n)) return fail(SOLHIP_E_INVALID, e);
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    HIP_TRY(hipMemcpy(s->materials, m, sizeof(sol_material) * n, hipMemcpyHostToDevice));
    bool kinds_changed = false;
    for (uint32_t i = 0; i < n; ++i) kinds_changed = kinds_changed || m[i].type != s->materials_host[i].type;
    s->materials_host.assign(m, m + n);
    if (kinds_changed) return refresh_kinds(s);
    return SOLHIP_OK;
}

int solhip_scene_set_lights(solhip_scene* s, const sol_light* l, uint32_t n) {
    if (!s || !l || n != s->num_lights) return fail(SOLHIP_E_INVALID, "light count mismatch");
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    HIP_TRY(hipMemcpy(s->lights, l, sizeof(sol_light) * n, hipMemcpyHostToDevice));
    s->lights_host.assign(l, l + n);
    return SOLHIP_OK;
}

int solhip_intersect(solhip_scene* s, const sol_ray* rays, uint64_t n, int any_hit, void* out, int device_ptrs) {
    if (!s || (n && (!rays || !out))) return fail(SOLHIP_E_INVALID, "null argument");
    if (n == 0) return SOLHIP_OK;
    Context* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const sol_ray* d_rays = rays;
    void* d_out = out;
    size_t out_bytes = any_hit ? sizeof(int32_t) * n : sizeof(sol_hit) * n;
    if (!device_ptrs) {
        // staging buffers of the context, grown on demand: many small calls (one ray per call
        // through sol::Geometry) allocate nothing
        HIP_TRY(ctx->isect_rays.reserve(sizeof(sol_ray) * n));
        HIP_TRY(ctx->isect_out.reserve(out_bytes));
        HIP_TRY(hipMemcpyAsync(ctx->isect_rays.ptr, rays, sizeof(sol_ray) * n, hipMemcpyHostToDevice, ctx->stream));
        d_rays = ctx->isect_rays.as<const sol_ray>();
        d_out = ctx->isect_out.ptr;
    }
    LaunchShape shape = ctx->shape(ctx->tuning.block_size, ctx->tuning.lds_stack);
    int st = ctx->ensure_scratch(shape.grid * shape.block);
    if (st) return st;
    HIP_TRY(ctx->isect_counters.reserve(sizeof(unsigned) * kIntersectCounterWords));
    if (ctx->profiling) HIP_TRY(hipEventRecord(ctx->ev_begin, ctx->stream));
    HIP_TRY(launch_intersect(s->view, d_rays, n, any_hit != 0, any_hit ? nullptr : static_cast<sol_hit*>(d_out),
                             any_hit ? static_cast<int32_t*>(d_out) : nullptr, ctx->isect_counters.as<unsigned>(),
                             ctx->scratch, shape, ctx->stream));
    if (ctx->profiling) HIP_TRY(hipEventRecord(ctx->ev_end, ctx->stream));
    if (!device_ptrs) HIP_TRY(hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->profiling) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
        ctx->stats.trace_ms += ms;
        ctx->stats.trace_launches += 1;
    }
    (any_hit ? ctx->stats.shadow_rays : ctx->stats.extension_rays) += n;
    ctx->stats.launches += 1;
    return ctx->check_overflow();
}

int solhip_render(solhip_scene* s, const sol_pt_config* cfg, int renderer, uint32_t width, uint32_t height,
                  uint64_t sample_index, uint32_t sample_count, float* accum, int device_ptr) {
    if (!s || !cfg || !accum) return fail(SOLHIP_E_INVALID, "null argument");
    if (width == 0 || height == 0) return fail(SOLHIP_E_INVALID, "empty image");
    if (renderer != SOLHIP_RENDERER_MEGAKERNEL && renderer != SOLHIP_RENDERER_WAVEFRONT)
        return fail(SOLHIP_E_INVALID, "unknown renderer");
    if (cfg->min_survival_prob < 0 || cfg->max_survival_prob > 1 || cfg->min_survival_prob > cfg->max_survival_prob)
        return fail(SOLHIP_E_INVALID, "survival probabilities must satisfy 0 <= min <= max <= 1");
    if (!s->num_lights) return fail(SOLHIP_E_NO_LIGHTS, "scene without light cannot be rendered");
    if (sample_count == 0) return SOLHIP_OK;
    Context* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = 3 * static_cast<size_t>(width) * height;
    float* d_accum = accum;
    if (!device_ptr) {
        // A host image: the samples are summed into a cleared device staging image, which comes
        // back once through pinned staging and is added on the host. The image is not sent to the
        // device: accum + (0 + s) is accum + s bit for bit (a pixel's sum s is never -0).
        HIP_TRY(ctx->host_image.reserve(sizeof(float) * n));
        d_accum = ctx->host_image.as<float>();
        HIP_TRY(launch_zero_f32(d_accum, n, ctx->stream));
    }
    int st;
    if (renderer == SOLHIP_RENDERER_MEGAKERNEL)
        st = render_megakernel(*s, *cfg, width, height, sample_index, sample_count, d_accum);
    else
        st = render_wavefront(*s, *cfg, width, height, sample_index, sample_count, d_accum);
    if (st) return st;
    if (!device_ptr) {
        if ((st = ctx->reserve_pinned(sizeof(float) * n))) return st;
        HIP_TRY(hipMemcpyAsync(ctx->pinned, d_accum, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        const float* sum = static_cast<const float*>(ctx->pinned);
        for (size_t i = 0; i < n; ++i) accum[i] += sum[i];
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ctx->check_overflow();
}

int solhip_backward(solhip_scene* s, const sol_pt_config* cfg, uint32_t width, uint32_t height, uint64_t sample_index,
                    uint32_t sample_count, const float* d_image, float* d_materials, float* d_lights, int device_ptrs) {
    if (!s || !cfg || !d_image || !d_lights || (s->num_materials && !d_materials))
        return fail(SOLHIP_E_INVALID, "null argument");
    if (width == 0 || height == 0) return fail(SOLHIP_E_INVALID, "empty image");
    if (cfg->min_survival_prob < 0 || cfg->max_survival_prob > 1 || cfg->min_survival_prob > cfg->max_survival_prob)
        return fail(SOLHIP_E_INVALID, "survival probabilities must satisfy 0 <= min <= max <= 1");
    if (!s->num_lights) return fail(SOLHIP_E_NO_LIGHTS, "scene without light cannot be rendered");
    if (sample_count == 0 || cfg->max_path_len == 0) return SOLHIP_OK;   // no path, no gradient
    Context* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t plane = static_cast<size_t>(width) * height;
    const size_t n_mat = static_cast<size_t>(s->num_materials) * 4, n_light = static_cast<size_t>(s->num_lights) * 3;
    const float* img = d_image;
    float* gm = d_materials;
    float* gl = d_lights;
    DeviceBuffer tmp_img, tmp_grad;
    if (!device_ptrs) {
        HIP_TRY(tmp_img.alloc(sizeof(float) * 3 * plane));
        HIP_TRY(tmp_grad.alloc(sizeof(float) * (n_mat + n_light + 1)));
        HIP_TRY(hipMemcpyAsync(tmp_img.ptr, d_image, sizeof(float) * 3 * plane, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemsetAsync(tmp_grad.ptr, 0, sizeof(float) * (n_mat + n_light + 1), ctx->stream));
        img = static_cast<const float*>(tmp_img.ptr);
        gm = static_cast<float*>(tmp_grad.ptr);
        gl = gm + n_mat;
    }
    LaunchShape shape = ctx->shape(ctx->tuning.block_size, ctx->tuning.lds_stack);
    int st = ctx->ensure_scratch(shape.grid * shape.block);
    if (st) return st;
    // Samples in batches whose per-sample radiance fits the forward buffer (kAdjointBytes).
    constexpr size_t kAdjointBytes = size_t(4) << 30;
    const uint32_t batch = static_cast<uint32_t>(
        std::max<size_t>(1, std::min<size_t>(sample_count, kAdjointBytes / (sizeof(float4) * plane))));
    HIP_TRY(ctx->adj_radiance.reserve(sizeof(float4) * plane * batch));
    HIP_TRY(ctx->adj_image.reserve(sizeof(float) * 3 * plane));
    for (uint32_t first = 0; first < sample_count; first += batch) {
        const uint32_t n = std::min(batch, sample_count - first);
        // forward: every sample's radiance (the image it accumulates is discarded)
        st = render_wavefront(*s, *cfg, width, height, sample_index + first, n, ctx->adj_image.as<float>(),
                              ctx->adj_radiance.as<float4>());
        if (st) return st;
        HIP_TRY(launch_zero_u64(ctx->counters, 4, ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev_begin, ctx->stream));
        HIP_TRY(launch_adjoint(s->view, *cfg, width, height, sample_index + first, n, ctx->adj_radiance.as<float4>(),
                               img, s->num_materials, gm, gl, ctx->counters, ctx->counters + 1, ctx->scratch, shape,
                               ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev_end, ctx->stream));
        unsigned long long c[4];
        st = ctx->read_counters(c, 4);
        if (st) return st;
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
        ctx->stats.render_ms += ms;
        ctx->stats.extension_rays += c[1];
        ctx->stats.shadow_rays += c[2];
        ctx->stats.launches += 1;
    }
    if (!device_ptrs) {
        std::vector<float> g(n_mat + n_light);
        HIP_TRY(hipMemcpy(g.data(), gm, sizeof(float) * (n_mat + n_light), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n_mat; ++i) d_materials[i] += g[i];
        for (size_t i = 0; i < n_light; ++i) d_lights[i] += g[n_mat + i];
    }
    return ctx->check_overflow();
}

} // extern "C"

namespace solhip {

int render_megakernel(solhip_scene& s, const sol_pt_config& cfg, uint32_t w, uint32_t h, uint64_t si, uint32_t sc,
                      float* d_accum) {
    Context* ctx = s.ctx;
    LaunchShape shape = ctx->shape(ctx->tuning.block_size, ctx->tuning.lds_stack);
    int st = ctx->ensure_scratch(shape.grid * shape.block);
    if (st) return st;
    HIP_TRY(launch_zero_u64(ctx->counters, 4, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev_begin, ctx->stream));
    HIP_TRY(launch_megakernel(s.view, cfg, w, h, si, sc, d_accum, ctx->counters, ctx->counters + 1, ctx->scratch,
                              shape, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev_end, ctx->stream));
    unsigned long long c[4];
    st = ctx->read_counters(c, 4);
    if (st) return st;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
    ctx->stats.render_ms += ms;
    ctx->stats.extension_rays += c[1];
    ctx->stats.shadow_rays += c[2];
    ctx->stats.samples += c[3];
    ctx->stats.launches += 1;
    return SOLHIP_OK;
}

} // namespace solhip
