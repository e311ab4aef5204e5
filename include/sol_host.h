/*
 * C ABI of the host library libsol_host.so (C++20 inside).
 *
 * The host library keeps sol's scene model (Scene, TriangleMesh with its BVH builder, Bsdf,
 * Light, Camera) and turns it into flat arrays:
 *   - the raw mesh (sol_mesh_view), the form sol's TriangleMesh constructor takes
 *     (reference: include/sol/triangle_mesh.h:18-24);
 *   - the binary BVH in bvh v2 layout (sol_bvh_view), the result of
 *     TriangleMesh::build_bvh (reference: src/triangle_mesh.cpp:119-146);
 *   - the GPU scene descriptor (solhip_scene_desc) that solhip_scene_upload copies to HBM.
 * It also generates the procedural scenes and ray batches the tests and benchmarks use.
 *
 * Conventions: every function returns 0 on success and a negative SOL_E_* code on failure;
 * sol_host_last_error() returns a thread-local message for the last failure. No C++
 * exception crosses this boundary.
 */
#ifndef SOL_HOST_H
#define SOL_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "sol_types.h"
#include "solhip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SOL_OK            = 0,
    SOL_E_INVALID     = -1, /* invalid argument                          */
    SOL_E_NO_LIGHTS   = -2, /* scene without light (undefined in sol, path_tracer.cpp:27,36) */
    SOL_E_ALLOC       = -3, /* host allocation failed                    */
    SOL_E_INTERNAL    = -4,
    SOL_E_LOAD        = -5  /* scene / OBJ file cannot be opened or parsed */
};

typedef struct sol_scene sol_scene;

/* A closest hit as sol::Geometry::intersect_closest returns it (sol::Hit, reference
   include/sol/geometry.h:14-31), with the BSDF and emitter as indices into the scene's material
   and light tables. */
typedef struct sol_surface_hit {
    int32_t hit;             /* 1 if the ray hit a triangle within (tmin, tmax)              */
    int32_t front;           /* SurfaceInfo::is_front_side                                    */
    int32_t material;        /* Hit::bsdf as a material index, -1 = none                      */
    int32_t light;           /* Hit::light as a light index, -1 = none                        */
    float t;                 /* the ray's tmax after the query: the hit distance on a hit     */
    float point[3];          /* SurfaceInfo::point                                            */
    float face_normal[3];    /* SurfaceInfo::face_normal, towards the ray (not normalised)    */
    float tangent[3];        /* SurfaceInfo::local, column 0                                  */
    float bitangent[3];      /* column 1                                                      */
    float normal[3];         /* column 2: the shading normal, towards the ray                 */
    float surf_coords[2];    /* SurfaceInfo::surf_coords: the triangle test's (u, v)          */
    float tex_coords[2];     /* SurfaceInfo::tex_coords                                       */
} sol_surface_hit;

/* Options of the CPU BVH builder: sweep SAH (bvh v2 SweepSahBuilder semantics) followed by
   the reinsertion optimiser (bvh v2 SequentialReinsertionOptimizer semantics). */
typedef struct sol_bvh_config {
    uint32_t min_leaf_size;      /* ranges of this size or smaller always become leaves   */
    uint32_t max_leaf_size;      /* ranges larger than this are always split (<= 8)      */
    float traversal_cost;        /* SAH cost of one inner-node visit                      */
    float intersection_cost;     /* SAH cost of one triangle test                         */
    uint32_t reinsertion_iters;  /* passes of the reinsertion optimiser (0 = off)         */
    float reinsertion_batch;     /* fraction of nodes, largest area first, moved per pass */
    uint32_t top_levels;         /* BVH levels laid out first (BFS) for staging in LDS    */
} sol_bvh_config;

/* The raw mesh and scene tables, in original triangle order. Pointers stay owned by the
   sol_scene they come from. */
typedef struct sol_mesh_view {
    uint32_t num_vertices;
    const float* positions;     /* 3 floats per vertex */
    const float* normals;       /* 3 floats per vertex */
    uint32_t num_tris;
    const uint32_t* indices;    /* 3 vertex indices per triangle */
    const int32_t* tri_material;/* material per triangle, -1 = no BSDF */
    const int32_t* tri_light;   /* light per triangle, -1 = not emissive */
    uint32_t num_materials;
    const sol_material* materials;
    uint32_t num_lights;
    const sol_light* lights;
    sol_camera_params camera;
} sol_mesh_view;

typedef struct sol_bvh_view {
    uint32_t num_nodes;
    const sol_bvh_node* nodes;
    uint32_t num_prims;
    const uint32_t* prim_indices;
} sol_bvh_view;

typedef struct sol_build_stats {
    double build_seconds;        /* sweep SAH                        */
    double optimize_seconds;     /* reinsertion                      */
    double flatten_seconds;      /* GPU layout                       */
    float sah_cost_built;        /* SAH cost after the sweep         */
    float sah_cost_final;        /* SAH cost after reinsertion       */
    uint32_t depth;              /* longest root-to-leaf path, nodes */
    uint32_t num_nodes;
    uint32_t num_leaves;
} sol_build_stats;

/* Procedural scene parameters for the sphere field (400 icospheres on a ground quad, three
   quad lights and one point light). */
typedef struct sol_sphere_field_params {
    uint32_t num_spheres;   /* 400                                   */
    uint32_t subdivision;   /* icosphere level: 20 * 4^level triangles */
    uint64_t seed;          /* PCG seed of the generator             */
    uint32_t num_groups;    /* material groups; 0 = one per sphere   */
    float aspect;           /* camera aspect ratio (width / height)  */
    float extent;           /* sphere centres in [-extent, extent]^2 */
    float radius_min;       /* log-uniform radius range              */
    float radius_max;
} sol_sphere_field_params;

/* Values a scene file may leave out: sol::Scene::Defaults (reference include/sol/scene.h:23-29).
   sol's driver sets aspect_ratio to width / height (test/driver.cpp:197-200). */
typedef struct sol_scene_defaults {
    float fov;            /* horizontal, degrees: 60 */
    float aspect_ratio;   /* 1 */
    float eye[3];         /* 0 0 0 */
    float dir[3];         /* 0 0 1 */
    float up[3];          /* 0 1 0 */
} sol_scene_defaults;

const char* sol_host_last_error(void);

void sol_default_bvh_config(sol_bvh_config* out);
void sol_default_pt_config(sol_pt_config* out);
void sol_default_sphere_field_params(sol_sphere_field_params* out);

/* Procedural Cornell box: five walls, a tall and a short box, one ceiling quad light with
   emission (17, 12, 4); camera of cornell_box.toml:3-8 with the given aspect ratio. */
int sol_scene_cornell_box(float aspect, const sol_bvh_config* cfg, sol_scene** out);

/* Sphere field. */
int sol_scene_sphere_field(const sol_sphere_field_params* p, const sol_bvh_config* cfg,
                           sol_scene** out);

/* Random triangle soup: centres uniform in [0,1]^3, vertices offset by U[-0.01,0.01]^3.
   Every triangle is diffuse; one point light so that the scene is renderable. */
int sol_scene_triangle_soup(uint32_t num_tris, uint64_t seed, const sol_bvh_config* cfg,
                            sol_scene** out);

void sol_default_scene_defaults(sol_scene_defaults* out);

/* Scene file: sol::Scene::load (reference src/scene_loader.cpp:30-108). TOML with an optional
   [camera] table (type "perspective"; eye, dir, up, fov, aspect), [[objects]] entries of type
   "import" naming an OBJ file relative to the scene file, and the name of the root object.
   OBJ / MTL import follows src/formats/obj.cpp in its non-strict mode; a material with both Kd
   and Ks becomes SOL_BSDF_INTERP, each triangle of a face with non-zero Ke a triangle light.
   SOL_E_LOAD with the message in sol_host_last_error() if a file cannot be opened or parsed;
   defaults may be null (the values above). */
int sol_scene_load(const char* file_name, const sol_scene_defaults* defaults, const sol_bvh_config* cfg,
                   sol_scene** out);

/* One OBJ file as the whole scene (sol::obj::load, src/formats/obj.cpp:512-531) seen through
   the given camera. */
int sol_scene_load_obj(const char* file_name, const sol_camera_params* camera, const sol_bvh_config* cfg,
                       sol_scene** out);

/* Scene from caller arrays (copied). A SOL_BSDF_INTERP material adds its diffuse and Phong
   components to the material table after the caller's materials (indices 0 .. n-1 keep their
   meaning). */
int sol_scene_from_arrays(const sol_mesh_view* mesh, const sol_bvh_config* cfg,
                          sol_scene** out);

/* sol::Geometry of the scene's root (reference include/sol/geometry.h:33-43, answered by
   TriangleMesh, src/triangle_mesh.cpp:64-117): attaches the root mesh to the GPU context
   (TriangleMesh::attach_gpu), then answers queries through the C++ interface - one virtual
   call per ray (batched == 0: intersect_closest(Rayf&) / intersect_any(const Rayf&)) or one
   batched call for all rays (batched != 0). Host arrays; the context must outlive the scene's
   attachment (sol_scene_destroy or another attach ends it). */
int sol_scene_attach_gpu(sol_scene* scene, solhip_ctx* ctx);
int sol_scene_intersect_closest(const sol_scene* scene, const sol_ray* rays, uint64_t n, int batched,
                                sol_surface_hit* out);
int sol_scene_intersect_any(const sol_scene* scene, const sol_ray* rays, uint64_t n, int batched,
                            int32_t* out);

void sol_scene_destroy(sol_scene* scene);

int sol_scene_mesh(const sol_scene* scene, sol_mesh_view* out);
int sol_scene_bvh(const sol_scene* scene, sol_bvh_view* out);
int sol_scene_gpu_desc(const sol_scene* scene, solhip_scene_desc* out);
int sol_scene_build_stats(const sol_scene* scene, sol_build_stats* out);

/* Replaces the material / light tables (geometry and BVH unchanged). */
int sol_scene_set_materials(sol_scene* scene, const sol_material* m, uint32_t n);
int sol_scene_set_lights(sol_scene* scene, const sol_light* l, uint32_t n);

/* Image files: sol::Image::save (reference src/image.cpp:26-37). format: 0 = auto (EXR, sol's
   default), 4 = EXR (uncompressed FLOAT scanlines, channels B, G, R as src/formats/exr.cpp:48-100
   writes them), 5 = PFM; 1-3 (PNG, JPEG, TIFF) have no codec in this library and fail. pixels is
   planar fp32, channel by channel, rows top to bottom. */
enum {
    SOL_IMAGE_AUTO = 0, SOL_IMAGE_PNG = 1, SOL_IMAGE_JPEG = 2, SOL_IMAGE_TIFF = 3, SOL_IMAGE_EXR = 4,
    SOL_IMAGE_PFM = 5
};
int sol_image_save(const char* path, int format, const float* pixels, uint32_t width, uint32_t height,
                   uint32_t channels);

/* Computes the camera basis from its parameters (cameras.cpp:5-21). */
void sol_camera_from_params(const sol_camera_params* p, sol_camera* out);

/* Ray batches of the intersection benchmark.
   kind 0: origins U[-0.5,1.5]^3, unit directions uniform on the sphere, tmin 0, tmax +inf.
   kind 1: the same origins, towards targets U[0,1]^3 with between_points conventions:
           dir = target - org (not normalised), tmin = offset, tmax = 1 - offset. */
int sol_gen_rays(uint32_t kind, uint64_t num_rays, uint64_t seed, float offset, sol_ray* out);

#ifdef __cplusplus
}
#endif

#endif
