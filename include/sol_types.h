/*
 * Plain-C data records shared by the host library (libsol_host.so), the GPU library
 * (libsolhip.so) and their callers. Every record is a fixed-size POD so that it can cross the
 * C ABI, be memcpy'd to HBM unchanged and be mirrored by ctypes.
 *
 * The records restate, as flat data, the scene objects that sol keeps behind virtual classes:
 *   sol_material  <- DiffuseBsdf / PhongBsdf / GlassBsdf / MirrorBsdf with constant textures,
 *                    and InterpBsdf over a DiffuseBsdf and a PhongBsdf (the mix the OBJ loader
 *                    makes, src/formats/obj.cpp:388-409)
 *                    (reference: include/sol/bsdfs.h:97-182, include/sol/textures.h:105-144)
 *   sol_light     <- UniformTriangleLight / PointLight (include/sol/lights.h:91-139,
 *                    include/sol/shapes.h:49-60)
 *   sol_camera_params <- PerspectiveCamera constructor arguments (include/sol/cameras.h:39-46)
 *   sol_pt_config <- PathTracer::Config (include/sol/algorithms/path_tracer.h:23-29)
 */
#ifndef SOL_TYPES_H
#define SOL_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BSDF kinds. Values are stable: they index the material-sorted shade queues on the GPU. */
enum {
    SOL_BSDF_NONE    = -1, /* surface without BSDF: paths end there (path_tracer.cpp:62-63) */
    SOL_BSDF_DIFFUSE = 0,  /* Lambertian, Type::Diffuse  (bsdfs.cpp:31-52)            */
    SOL_BSDF_PHONG   = 1,  /* modified Phong, Type::Glossy (bsdfs.cpp:64-104)         */
    SOL_BSDF_GLASS   = 2,  /* dielectric, Type::Specular (bsdfs.cpp:131-165)          */
    SOL_BSDF_MIRROR  = 3,  /* perfect mirror, Type::Specular (bsdfs.cpp:108-119)      */
    SOL_BSDF_INTERP  = 4   /* InterpBsdf(DiffuseBsdf(kd), PhongBsdf(ks, ns), mix): Type::Diffuse
                              (bsdfs.cpp:179-230)                                        */
};

enum {
    SOL_LIGHT_TRIANGLE = 0, /* AreaLight<UniformTriangle> (lights.cpp:59-111) */
    SOL_LIGHT_POINT    = 1  /* PointLight (lights.cpp:11-44)                  */
};

/* One material: constant-texture parameters of one BSDF. 64 bytes. */
typedef struct sol_material {
    int32_t type;       /* SOL_BSDF_*                                                  */
    float kd[3];        /* diffuse albedo                                               */
    float ks[3];        /* Phong / glass / mirror specular colour                       */
    float kt[3];        /* glass transmission colour                                    */
    float ns;           /* Phong exponent used by eval()                                */
    float eta;          /* glass index ratio as the OBJ loader stores it (1/Ni)         */
    float ns_sample;    /* Phong exponent used by sample()/pdf(); equal to ns for plain
                           rendering, held fixed for frozen-sampling gradient checks    */
    int32_t group;      /* parameter group for gradients (-1 = not differentiated)      */
    float mix;          /* SOL_BSDF_INTERP: weight of the Phong lobe, the constant texture k
                           of InterpBsdf (obj.cpp: spec_k / (diff_k + spec_k))           */
    float pad;
} sol_material;

/* One light. 80 bytes. For a point light only v0 (position) and intensity are used. */
typedef struct sol_light {
    int32_t type;       /* SOL_LIGHT_*                                                   */
    float v0[3];
    float v1[3];
    float v2[3];
    float normal[3];    /* unit normal normalize(cross(v1-v0, v2-v0)) of a triangle light */
    float inv_area;     /* 1 / area of a triangle light                                  */
    float intensity[3]; /* emitted radiance (area light) or intensity (point light)      */
    int32_t group;      /* parameter group for emission gradients (-1 = none)            */
    int32_t pad[2];
} sol_light;

/* PerspectiveCamera constructor arguments (cameras.cpp:5-21). fov is horizontal, degrees. */
typedef struct sol_camera_params {
    float eye[3];
    float dir[3];
    float up[3];
    float fov;
    float aspect;
} sol_camera_params;

/* The camera basis computed from sol_camera_params exactly as cameras.cpp:5-21 does:
   dir normalised, right = normalize(cross(dir, up)) * w, up = cross(right_n, dir_n) * h. */
typedef struct sol_camera {
    float eye[3];
    float dir[3];
    float right[3];
    float up[3];
} sol_camera;

/* PathTracer::Config (path_tracer.h:23-29) plus the three compile-time switches of
   path_tracer.cpp:32-34 turned into run-time flags. */
enum {
    SOL_PT_DISABLE_MIS = 1,
    SOL_PT_DISABLE_NEE = 2,
    SOL_PT_DISABLE_RR  = 4
};

typedef struct sol_pt_config {
    uint32_t max_path_len;     /* default 64  */
    uint32_t min_rr_path_len;  /* default 3   */
    float min_survival_prob;   /* default 0.05 */
    float max_survival_prob;   /* default 0.75 */
    float ray_offset;          /* default 1e-5 */
    uint32_t flags;            /* SOL_PT_* */
} sol_pt_config;

/* A ray as the intersection boundary consumes it (32 bytes). */
typedef struct sol_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} sol_ray;

/* A closest-hit result (16 bytes). prim = original triangle index, -1 on a miss.
   u weights v1 and v (here "v" field) weights v2, as triangle_mesh.cpp:83-84 uses them. */
typedef struct sol_hit {
    int32_t prim;
    float t;
    float u;
    float v;
} sol_hit;

/* A node of a binary BVH in the layout of bvh v2 (children of an inner node are adjacent):
   prim_count == 0 -> inner node, children at index and index + 1;
   prim_count  > 0 -> leaf holding prim_indices[index .. index + prim_count). */
typedef struct sol_bvh_node {
    float bounds[6];   /* min x, max x, min y, max y, min z, max z */
    uint32_t index;
    uint32_t prim_count;
} sol_bvh_node;

/* GPU node: both children's boxes in the parent, 64 bytes (4 x 16-byte loads).
   child[i] >= 0: inner node index. child[i] < 0: leaf, ~child[i] = (first << 3) | (count - 1),
   first indexing the leaf-ordered triangle array. */
typedef struct sol_gpu_node {
    float b0[4];       /* child 0: lo.x hi.x lo.y hi.y */
    float b1[4];       /* child 1: lo.x hi.x lo.y hi.y */
    float bz[4];       /* child 0 lo.z hi.z, child 1 lo.z hi.z */
    int32_t child[2];
    int32_t pad[2];
} sol_gpu_node;

/* 4-wide GPU node (128 bytes): the boxes of up to four children, each coordinate as an array over
   the children, and the child references (child[i] >= 0: inner node index, < 0: leaf encoded as in
   sol_gpu_node, SOL_GPU_EMPTY: unused slot). An unused slot holds the inverted box lo = +inf,
   hi = -inf, which every ray misses, so the traversal tests its box like any other;
   solhip_scene_upload writes that box into every unused slot it is given. */
#define SOL_GPU_EMPTY ((int32_t)0x7FFFFFFF)
typedef struct sol_gpu_node4 {
    float lo_x[4];
    float hi_x[4];
    float lo_y[4];
    float hi_y[4];
    float lo_z[4];
    float hi_z[4];
    int32_t child[4];
    int32_t pad[4];
} sol_gpu_node4;

/* GPU triangle in leaf order: the three vertices (watertight test input) and the original
   triangle index. 48 bytes. */
typedef struct sol_gpu_tri {
    float v0[3];
    int32_t prim;
    float v1[3];
    int32_t pad0;      /* ignored on upload; the GPU library keeps the triangle's kind here (BSDF
                          of its material + 1 in bits 0-2, bit 3 if it is a light), recomputed on
                          upload and whenever the materials change, for the wavefront traversal */
    float v2[3];
    int32_t pad1;
} sol_gpu_tri;

/* Shading record in leaf order: the three vertex normals, the geometric normal
   cross(v1 - v0, v2 - v0) (computed in fp32 without fused multiply-add), material and light of
   the triangle, and the BSDF kind of that material (what the hit sorting reads, one dependent
   load fewer; the GPU library recomputes it on upload and whenever the materials change).
   64 bytes. */
typedef struct sol_gpu_shade {
    float n0[3];
    float n1[3];
    float n2[3];
    float face_n[3];
    int32_t material;  /* index into the material table, -1 = no BSDF */
    int32_t light;     /* index into the light table, -1 = not emissive */
    int32_t bsdf;      /* SOL_BSDF_* of the material, SOL_BSDF_NONE without one */
    int32_t pad;
} sol_gpu_shade;

#ifdef __cplusplus
}
#endif

#endif
