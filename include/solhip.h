/*
 * C ABI of the GPU library libsolhip.so: hand-written HIP kernels for gfx950 behind plain
 * pointers and sizes. This is the drop-in boundary of the path-tracing hot path.
 *
 * Reference interfaces replaced (paths relative to the sol sources):
 *   solhip_render        <- sol::Renderer::render(Image&, sample_index, sample_count)
 *                           (include/sol/renderer.h:34) as implemented by PathTracer::render
 *                           (src/algorithms/path_tracer.cpp:9-23): adds, per pixel, the sum of
 *                           the samples [sample_index, sample_index + sample_count) into a
 *                           3-channel planar fp32 image (include/sol/image.h:54-60).
 *   solhip_intersect     <- sol::Geometry::intersect_closest / intersect_any
 *                           (include/sol/geometry.h:33-43; src/triangle_mesh.cpp:64-117),
 *                           batched.
 *   solhip_scene_upload  <- the TriangleMesh / Scene state the renderer reads
 *                           (include/sol/scene.h:34-57, include/sol/triangle_mesh.h:18-60).
 *   solhip_render_batch, solhip_backward, solhip_autotune: no counterpart in sol (many scenes
 *                           in one launch; adjoint path replay; launch-parameter search).
 *
 * Conventions:
 *   - return 0 on success, a negative SOLHIP_E_* code on failure; solhip_last_error() gives
 *     a thread-local message. Nothing throws across the ABI.
 *   - one HIP stream per context; calls on one context must be serialised by the caller (as a
 *     sol::RenderJob is controlled from one thread, include/sol/render_job.h:15-17);
 *     distinct contexts are independent.
 *   - every call returns after its GPU work has finished.
 *   - a "device" pointer argument is a HIP device address on the context's device; a "host"
 *     pointer is ordinary host memory. Each entry point says which.
 */
#ifndef SOLHIP_H
#define SOLHIP_H

#include <stddef.h>
#include <stdint.h>

#include "sol_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SOLHIP_OK          = 0,
    SOLHIP_E_INVALID   = -1,  /* invalid argument (null, zero size, bad enum)          */
    SOLHIP_E_NO_LIGHTS = -2,  /* rendering a scene without light (one uploaded for
                                 intersection queries only)                            */
    SOLHIP_E_HIP       = -3,  /* a HIP runtime call failed                             */
    SOLHIP_E_NO_DEVICE = -4,  /* no HIP device / device index out of range             */
    SOLHIP_E_TOO_DEEP  = -5,  /* BVH deeper than the traversal stack                   */
    SOLHIP_E_CONTEXT   = -6   /* objects from different contexts mixed                 */
};

enum {
    SOLHIP_RENDERER_MEGAKERNEL = 0, /* one thread per pixel runs whole paths (plain GPU)   */
    SOLHIP_RENDERER_WAVEFRONT  = 1  /* ray queues, material-sorted shading, graph loop    */
};

typedef struct solhip_ctx solhip_ctx;
typedef struct solhip_scene solhip_scene;

/* Flat scene, produced by the host library (sol_scene_gpu_desc). All pointers are host
   pointers; solhip_scene_upload copies them, the caller keeps ownership. */
typedef struct solhip_scene_desc {
    uint32_t num_nodes;
    const sol_gpu_node* nodes;      /* node 0 is the root                         */
    uint32_t num_tris;
    const sol_gpu_tri* tris;        /* leaf order                                 */
    const sol_gpu_shade* shade;     /* leaf order                                 */
    uint32_t num_materials;
    const sol_material* materials;
    uint32_t num_lights;
    const sol_light* lights;
    sol_camera camera;
    uint32_t bvh_depth;             /* longest root-to-leaf path (nodes)          */
    uint32_t top_nodes;             /* nodes [0, top_nodes) are the top levels    */
    /* The same tree collapsed to 4-wide nodes (the wavefront traversal kernel's layout). */
    uint32_t num_nodes4;
    const sol_gpu_node4* nodes4;    /* node 0 is the root                         */
    uint32_t bvh4_depth;            /* longest root-to-leaf path (4-wide nodes)   */
    uint32_t top_nodes4;            /* nodes4 [0, top_nodes4) are the top levels  */
} solhip_scene_desc;

/* Launch parameters the autotuner chooses. Zero fields mean "library default". */
typedef struct solhip_tuning {
    uint32_t block_size;    /* threads per workgroup of the traversal kernels: 64, 128, 256 */
    uint32_t lds_stack;     /* short-stack entries per lane kept in LDS: 8, 16, 32           */
    uint32_t path_slots;    /* wavefront: paths in flight, 4096..2^27 (default: the largest
                               power of two whose 348 B a slot take at most a quarter of the
                               device memory, 2^27 on a 288 GB MI355X)                      */
    uint32_t tile_w;        /* batch: tile shape of the global work queue                    */
    uint32_t tile_h;
    uint32_t graph_iters;   /* bounce-loop iterations per captured graph (a render whose
                               chunks all fit in the paths in flight captures max_path_len + 1
                               iterations, rounded up to even, instead: one replay finishes
                               each chunk)                                                  */
    uint32_t bvh_width;     /* 2 or 4: node layout of the traversal in the wavefront renderer,
                               the megakernel, the adjoint replay and the intersection kernel
                               (4 needs the scene's 4-wide layout; scenes without one use the
                               binary).                                                     */
    uint32_t chunk_items;   /* wavefront: work items (path samples) per chunk of work, the
                               capacity of the per-sample radiance buffer (16 B an item);
                               65536..2^28. Whole tiles make a chunk. Default: path_slots
                               (at least 65536, at most 192 Mi), so that every chunk fits in
                               the paths in flight and runs in place (each path keeps one
                               item, its record is the item's sample entry).                 */
    uint32_t queue_segments;/* wavefront queue split: segments per queue, 1 (default), 2, 4 or
                               8. Each segment has its own length counter, so the appends of
                               a pass spread their device atomics over that many addresses;
                               one atomic per queue and round of 512 entries is rarely
                               contended, and one segment keeps each queue contiguous.      */
    uint32_t shade_queues;  /* wavefront queue split: 3 (default) sorts the hits into diffuse,
                               glossy and specular queues shaded by one kernel each; 1 keeps one
                               queue shaded by a kernel for every BSDF. Images are the same.  */
} solhip_tuning;

/* Counters accumulated by the context since the last solhip_reset_stats. */
typedef struct solhip_stats {
    uint64_t extension_rays;   /* closest-hit queries                          */
    uint64_t shadow_rays;      /* any-hit queries                              */
    uint64_t samples;          /* paths started                                */
    uint64_t launches;         /* kernel launches (graph nodes count once each) */
    uint64_t graph_replays;
    double trace_ms;           /* summed durations measured with HIP events (profiling on):  */
    double sort_ms;            /*   traversal, hit sorting and shading kernels of the        */
    double shade_ms;           /*   wavefront loop; traversal of solhip_intersect            */
    double reserved_ms[2];
    uint64_t trace_launches;   /* launches covered by trace_ms                 */
    double render_ms;          /* wall time of render calls (HIP events)       */
} solhip_stats;

typedef struct solhip_device_info {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t wavefront_size;
    uint64_t total_memory;
    int32_t lds_per_block;
} solhip_device_info;

/* One image of a batched render. accum is a DEVICE pointer to 3 * width * height floats
   (planar R, G, B), to which the samples are added. */
typedef struct solhip_batch_job {
    solhip_scene* scene;
    uint32_t width;
    uint32_t height;
    uint64_t sample_index;
    uint32_t sample_count;
    uint32_t pad;
    float* accum;
} solhip_batch_job;

/* Raw mesh for the GPU BVH builder (host memory; solhip_scene_build copies it). */
typedef struct solhip_mesh_desc {
    uint32_t num_vertices;
    const float* positions;         /* 3 floats per vertex                              */
    const float* normals;           /* 3 floats per vertex                              */
    uint32_t num_tris;
    const uint32_t* indices;        /* 3 vertex indices per triangle                    */
    const int32_t* tri_material;    /* per triangle, -1 = no BSDF                       */
    const int32_t* tri_light;       /* per triangle, -1 = not emissive                  */
    uint32_t num_materials;
    const sol_material* materials;
    uint32_t num_lights;
    const sol_light* lights;
    sol_camera camera;
} solhip_mesh_desc;

/* Options of the GPU builder. Zero fields mean the defaults in brackets. */
typedef struct solhip_build_config {
    uint32_t max_leaf_size;         /* triangles per leaf, 1..8 [8]                     */
    float traversal_cost;           /* SAH cost of an inner-node visit [1]              */
    float intersection_cost;        /* SAH cost of a triangle test [1]                  */
    uint32_t search_radius;         /* clustering neighbourhood, 1..32 [16]             */
} solhip_build_config;

typedef struct solhip_build_stats {
    double build_ms;                /* wall time of the build on the device side        */
    uint32_t iterations;            /* clustering passes                                */
    uint32_t num_nodes;             /* binary inner nodes                               */
    uint32_t num_nodes4;            /* 4-wide nodes                                     */
    uint32_t num_leaves;
    uint32_t depth;                 /* longest root-to-leaf path, binary (nodes)         */
    uint32_t depth4;                /* the same, 4-wide                                 */
    float sah_cost;                 /* SAH cost of the binary tree / root area           */
} solhip_build_stats;

/* Sizes of an uploaded or built scene, for solhip_scene_download. */
typedef struct solhip_scene_info {
    uint32_t num_nodes, num_nodes4, num_tris, num_materials, num_lights;
    uint32_t bvh_depth, bvh4_depth, top_nodes, top_nodes4;
} solhip_scene_info;

const char* solhip_last_error(void);

int solhip_create(int device, solhip_ctx** out);
void solhip_destroy(solhip_ctx* ctx);
int solhip_device_query(solhip_ctx* ctx, solhip_device_info* out);

int solhip_set_tuning(solhip_ctx* ctx, const solhip_tuning* t);
int solhip_get_tuning(solhip_ctx* ctx, solhip_tuning* out);
/* profiling != 0: record HIP events around the traversal, sort and shade launches of the
   wavefront loop (launched eagerly then, not from the captured graph) into the stats. */
int solhip_set_profiling(solhip_ctx* ctx, int profiling);
int solhip_get_stats(solhip_ctx* ctx, solhip_stats* out);
int solhip_reset_stats(solhip_ctx* ctx);

/* Copies a flat scene to the context's device. A scene without lights (num_lights 0) serves
   intersection queries only (sol::Geometry through solhip_intersect); rendering it returns
   SOLHIP_E_NO_LIGHTS, since sol's path tracer divides by the light count (path_tracer.cpp:27,36). */
int solhip_scene_upload(solhip_ctx* ctx, const solhip_scene_desc* desc, solhip_scene** out);
void solhip_scene_destroy(solhip_scene* scene);

/* Builds the scene's BVH on the GPU: no counterpart in sol, whose TriangleMesh builds its BVH on
   the host (src/triangle_mesh.cpp:119-146; this library's host builder follows that one). Morton
   order, then agglomerative clustering of nearest neighbours by merged surface area within a
   window of the order (PLOC, Meister and Bittner 2018), SAH collapse of subtrees into leaves of up
   to max_leaf_size triangles, breadth-first layouts (binary and 4-wide). Renders of the result are
   bit-identical to renders of any other valid BVH of the same mesh. cfg and stats may be null. */
int solhip_scene_build(solhip_ctx* ctx, const solhip_mesh_desc* mesh, const solhip_build_config* cfg,
                       solhip_scene** out, solhip_build_stats* stats);

int solhip_scene_info_get(const solhip_scene* scene, solhip_scene_info* out);
/* Copies the device arrays of a scene to host memory (sizes from solhip_scene_info_get); null
   pointers are skipped. */
int solhip_scene_download(const solhip_scene* scene, sol_gpu_node* nodes, sol_gpu_node4* nodes4, sol_gpu_tri* tris,
                          sol_gpu_shade* shade);
/* Replace the material or light table (host arrays, same counts as uploaded). */
int solhip_scene_set_materials(solhip_scene* scene, const sol_material* m, uint32_t n);
int solhip_scene_set_lights(solhip_scene* scene, const sol_light* l, uint32_t n);

/* Batched intersection (Geometry::intersect_closest / intersect_any).
   rays: n sol_ray; any_hit == 0 -> out is n sol_hit (prim = -1 on a miss);
   any_hit != 0 -> out is n int32 (1 = occluded). device_ptrs != 0: rays/out are device
   pointers, else host pointers. Ties in t resolve to the smaller original triangle index. */
int solhip_intersect(solhip_scene* scene, const sol_ray* rays, uint64_t n, int any_hit,
                     void* out, int device_ptrs);

/* Renderer::render: accum += sum of samples [sample_index, sample_index + sample_count) per
   pixel; accum is 3 * width * height floats, planar RGB, pixel (0,0) top-left.
   device_ptr != 0: accum is a device pointer, else a host pointer. */
int solhip_render(solhip_scene* scene, const sol_pt_config* cfg, int renderer,
                  uint32_t width, uint32_t height, uint64_t sample_index,
                  uint32_t sample_count, float* accum, int device_ptr);

/* solhip_render with the megakernel into a DEVICE image (accum: 3 * width * height floats, planar
   RGB), allocating nothing: no counterpart in sol; for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Its
   scratch - counters, the stack-overflow flag and the traversal stack's spill area (64 entries per
   lane beyond the LDS stack, for up to 196,608 lanes) - is one static array of 48 MiB + 256 B per
   GPU, held until the process ends. Images equal solhip_render's bit for bit. Scenes whose BVH is
   deeper than that stack (binary depth > lds_stack + 65, or 4-wide depth with 3 (depth - 1) >
   lds_stack + 64) are refused with SOLHIP_E_TOO_DEEP. Calls that overlap in time on one GPU share
   the array and race: serialise them, as every call on one context already is. */
int solhip_render_noalloc(solhip_scene* scene, const sol_pt_config* cfg, uint32_t width,
                          uint32_t height, uint64_t sample_index, uint32_t sample_count,
                          float* accum);

/* Many scenes and image tiles in one wavefront run fed by a global work queue. All scenes
   must belong to ctx. Accumulation semantics per job as solhip_render. Jobs may share an image
   (or overlapping memory), e.g. two sample ranges of one frame: their sums are added in job order,
   exactly as successive solhip_render calls would add them. */
int solhip_render_batch(solhip_ctx* ctx, const solhip_batch_job* jobs, uint32_t num_jobs,
                        const sol_pt_config* cfg);

/* Adjoint path replay. d_image: 3 * width * height floats (planar RGB), dLoss/d(accum).
   Adds gradients into d_materials (num_materials x 4: d kd.r, d kd.g, d kd.b, d log ns; kd of a
   diffuse BSDF or of the diffuse lobe of SOL_BSDF_INTERP, ns of a Phong BSDF or lobe) and
   d_lights (num_lights x 3: d intensity r, g, b). Sampling decisions, MIS weights and Russian
   roulette are treated as constants (detached), pdfs use ns_sample. A forward pass of the
   wavefront renderer supplies each sample's radiance; the paths are then replayed once.
   device_ptrs != 0: all three arrays are device pointers, else host pointers. */
int solhip_backward(solhip_scene* scene, const sol_pt_config* cfg, uint32_t width,
                    uint32_t height, uint64_t sample_index, uint32_t sample_count,
                    const float* d_image, float* d_materials, float* d_lights, int device_ptrs);

/* Times candidate launch parameters on the given scene and image and keeps the fastest in the
   context (also written to *out if not null). budget_samples bounds the work per trial. */
int solhip_autotune(solhip_scene* scene, const sol_pt_config* cfg, uint32_t width,
                    uint32_t height, uint32_t budget_spp, solhip_tuning* out);

/* The same search on a batch: every trial is solhip_render_batch of the given jobs, rendered into
   scratch images (the jobs' own images are not written). Keeps the fastest parameters in the
   context (also written to *out if not null). No counterpart in sol. */
int solhip_autotune_batch(solhip_ctx* ctx, const solhip_batch_job* jobs, uint32_t num_jobs,
                          const sol_pt_config* cfg, solhip_tuning* out);

#ifdef __cplusplus
}
#endif

#endif
