/*
 * CPU oracle: a plain C++ restatement of sol's path-tracing hot path, used ONLY as the checker
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. Nothing in the product
 * (sol_amd/) links, loads or calls it.
 *
 * It consumes the same flat scene records as the GPU library (include/sol_types.h,
 * include/sol_host.h: sol_mesh_view and optionally a sol_bvh_view) and follows, function by
 * function, the sol sources (paths relative to the sol tree):
 *   oracle_render          PathTracer::render + trace_path   src/algorithms/path_tracer.cpp:9-120
 *   oracle_intersect       TriangleMesh::intersect_closest/any src/triangle_mesh.cpp:64-117
 *   oracle_pixel_seed      Renderer::pixel_seed              include/sol/renderer.h:44-46
 *   oracle_pcg             PcgGenerator + uniform float       include/sol/samplers.h:20-65
 * sol's own arithmetic library (proto), BVH library (bvh v2) and executors (par) are not
 * vendored; the oracle states their operations from the mathematics (see oracle.cpp).
 *
 * Parity status of each function, and what pins it, is listed in oracle.cpp's header.
 */
#ifndef SOL_ORACLE_H
#define SOL_ORACLE_H

#include <stdint.h>

#include "sol_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

/* Counters of an oracle run. */
typedef struct oracle_stats {
    uint64_t extension_rays;     /* closest-hit queries                       */
    uint64_t shadow_rays;        /* any-hit queries                           */
    uint64_t ext_inner_visits;   /* inner nodes processed by closest-hit      */
    uint64_t ext_tri_tests;      /* triangle tests by closest-hit             */
    uint64_t shadow_inner_visits;
    uint64_t shadow_tri_tests;
    uint64_t bsdf_hits;          /* surface hits with a BSDF                  */
    uint64_t samples;
} oracle_stats;

/* bvh == NULL: the oracle builds its own BVH (object-median split). Otherwise the given
   bvh v2-layout BVH is used as is (it must index the mesh's triangles). */
int oracle_scene_create(const sol_mesh_view* mesh, const sol_bvh_view* bvh, oracle_scene** out);
void oracle_scene_destroy(oracle_scene* s);
int oracle_scene_set_materials(oracle_scene* s, const sol_material* m, uint32_t n);
int oracle_scene_set_lights(oracle_scene* s, const sol_light* l, uint32_t n);

/* brute_force != 0: test every triangle (no BVH). out: sol_hit[n] or int32[n] as solhip_intersect. */
int oracle_intersect(const oracle_scene* s, const sol_ray* rays, uint64_t n, int any_hit, int brute_force,
                     void* out, oracle_stats* stats);

/* Renderer::render: accum (3 planes of w*h floats) += per-pixel sum of samples
   [sample_index, sample_index + sample_count). threads <= 0: OpenMP default.
   rows: render only pixel rows [row_begin, row_end) (row_end == 0: all rows). */
int oracle_render(const oracle_scene* s, const sol_pt_config* cfg, uint32_t w, uint32_t h, uint64_t sample_index,
                  uint32_t sample_count, float* accum, int threads, uint32_t row_begin, uint32_t row_end,
                  oracle_stats* stats);

/* Adjoint of oracle_render: adds d(sum(d_image * image))/d(parameter) into d_materials
   (num_materials x 4: kd r, g, b, log ns) and d_lights (num_lights x 3: intensity r, g, b), with
   sampling, MIS weights and Russian roulette held constant. */
int oracle_backward(const oracle_scene* s, const sol_pt_config* cfg, uint32_t w, uint32_t h, uint64_t sample_index,
                    uint32_t sample_count, const float* d_image, float* d_materials, float* d_lights, int threads);

uint32_t oracle_pixel_seed(uint64_t sample_index, uint64_t x, uint64_t y);
/* First n outputs of PcgGenerator(seed) and their floats. */
void oracle_pcg(uint64_t seed, uint32_t n, uint32_t* out_u32, float* out_float);
/* Sampling / basis primitives, exposed for closed-form checks. */
void oracle_sample_cosine_hemisphere(float u, float v, float* dir3, float* pdf);
void oracle_sample_cosine_power_hemisphere(float n, float u, float v, float* dir3, float* pdf);
void oracle_ortho_basis(const float* n3, float* t3, float* b3);
/* Intersection of one ray with one triangle (watertight test): returns 1 and t,u,v on a hit
   inside the edges (t not range-checked). */
int oracle_ray_triangle(const float* org3, const float* dir3, const float* v0, const float* v1, const float* v2,
                        float* t, float* u, float* v);
/* Glass/Phong/diffuse BSDF sampling with a scripted random sequence, for closed-form checks:
   normal n3 (unit), geometric normal = n3, out3 unit, front side; writes in3, pdf, cos, rgb.
   Returns 1 if the sample is valid. */
int oracle_bsdf_sample(const sol_material* m, const float* n3, const float* out3, int front, const float* randoms,
                       float* in3, float* pdf, float* cos, float* rgb);
float oracle_fresnel(float eta, float cos_i, float cos_t);
/* TriangleMesh::intersect_closest with its hit decoding (src/triangle_mesh.cpp:64-105): the
   surface record of each ray's closest hit (tex_coords left 0: the oracle has none), t = the ray's
   tmax after the query. */
int oracle_intersect_surface(const oracle_scene* s, const sol_ray* rays, uint64_t n, sol_surface_hit* out);
/* Renderer::sample_pixel (include/sol/renderer.h:50-54) of pixel (x, y) of a w x h image for the
   absolute sample index: the first two draws of the pixel's stream, as oracle_render makes them.
   uv[0] = u (left -1, right +1), uv[1] = v (top +1, bottom -1). */
void oracle_sample_pixel(uint64_t sample_index, uint32_t x, uint32_t y, uint32_t w, uint32_t h, float* uv);
/* PerspectiveCamera::generate_ray (src/cameras.cpp:23-25) of the scene's camera at (u, v), as
   oracle_render makes it: origin and unit direction. */
int oracle_camera_ray(const oracle_scene* s, float u, float v, float* org3, float* dir3);

#ifdef __cplusplus
}
#endif

#endif
