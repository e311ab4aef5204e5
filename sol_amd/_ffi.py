"""ctypes mirror of the C ABI (include/sol_types.h, include/sol_host.h, include/solhip.h).

The two product libraries are loaded from this package directory. There is no fallback: if
libsolhip.so or libsol_host.so is missing, importing this module raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

BSDF_NONE, BSDF_DIFFUSE, BSDF_PHONG, BSDF_GLASS, BSDF_MIRROR, BSDF_INTERP = -1, 0, 1, 2, 3, 4
LIGHT_TRIANGLE, LIGHT_POINT = 0, 1
PT_DISABLE_MIS, PT_DISABLE_NEE, PT_DISABLE_RR = 1, 2, 4
RENDERER_MEGAKERNEL, RENDERER_WAVEFRONT = 0, 1
SOL_E_INVALID, SOL_E_NO_LIGHTS, SOL_E_ALLOC, SOL_E_INTERNAL, SOL_E_LOAD = -1, -2, -3, -4, -5
IMAGE_AUTO, IMAGE_PNG, IMAGE_JPEG, IMAGE_TIFF, IMAGE_EXR, IMAGE_PFM = 0, 1, 2, 3, 4, 5


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("kd", C.c_float * 3), ("ks", C.c_float * 3), ("kt", C.c_float * 3),
                ("ns", C.c_float), ("eta", C.c_float), ("ns_sample", C.c_float), ("group", C.c_int32),
                ("mix", C.c_float), ("pad", C.c_float)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("v0", C.c_float * 3), ("v1", C.c_float * 3), ("v2", C.c_float * 3),
                ("normal", C.c_float * 3), ("inv_area", C.c_float), ("intensity", C.c_float * 3),
                ("group", C.c_int32), ("pad", C.c_int32 * 2)]


class CameraParams(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("dir", C.c_float * 3), ("up", C.c_float * 3), ("fov", C.c_float),
                ("aspect", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("dir", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3)]


class PtConfig(C.Structure):
    _fields_ = [("max_path_len", C.c_uint32), ("min_rr_path_len", C.c_uint32), ("min_survival_prob", C.c_float),
                ("max_survival_prob", C.c_float), ("ray_offset", C.c_float), ("flags", C.c_uint32)]


class Ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tmin", C.c_float), ("dir", C.c_float * 3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("prim", C.c_int32), ("t", C.c_float), ("u", C.c_float), ("v", C.c_float)]


class BvhNode(C.Structure):
    _fields_ = [("bounds", C.c_float * 6), ("index", C.c_uint32), ("prim_count", C.c_uint32)]


class GpuNode(C.Structure):
    _fields_ = [("b0", C.c_float * 4), ("b1", C.c_float * 4), ("bz", C.c_float * 4), ("child", C.c_int32 * 2),
                ("pad", C.c_int32 * 2)]


class GpuNode4(C.Structure):
    _fields_ = [("lo_x", C.c_float * 4), ("hi_x", C.c_float * 4), ("lo_y", C.c_float * 4), ("hi_y", C.c_float * 4),
                ("lo_z", C.c_float * 4), ("hi_z", C.c_float * 4), ("child", C.c_int32 * 4), ("pad", C.c_int32 * 4)]


GPU_EMPTY = 0x7FFFFFFF


class GpuTri(C.Structure):
    _fields_ = [("v0", C.c_float * 3), ("prim", C.c_int32), ("v1", C.c_float * 3), ("pad0", C.c_int32),
                ("v2", C.c_float * 3), ("pad1", C.c_int32)]


class GpuShade(C.Structure):
    _fields_ = [("n0", C.c_float * 3), ("n1", C.c_float * 3), ("n2", C.c_float * 3), ("face_n", C.c_float * 3),
                ("material", C.c_int32), ("light", C.c_int32), ("bsdf", C.c_int32), ("pad", C.c_int32)]


class BvhConfig(C.Structure):
    _fields_ = [("min_leaf_size", C.c_uint32), ("max_leaf_size", C.c_uint32), ("traversal_cost", C.c_float),
                ("intersection_cost", C.c_float), ("reinsertion_iters", C.c_uint32),
                ("reinsertion_batch", C.c_float), ("top_levels", C.c_uint32)]


class MeshView(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("num_tris", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("tri_material", C.POINTER(C.c_int32)), ("tri_light", C.POINTER(C.c_int32)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(Material)), ("num_lights", C.c_uint32),
                ("lights", C.POINTER(Light)), ("camera", CameraParams)]


class BvhView(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("nodes", C.POINTER(BvhNode)), ("num_prims", C.c_uint32),
                ("prim_indices", C.POINTER(C.c_uint32))]


class BuildStats(C.Structure):
    _fields_ = [("build_seconds", C.c_double), ("optimize_seconds", C.c_double), ("flatten_seconds", C.c_double),
                ("sah_cost_built", C.c_float), ("sah_cost_final", C.c_float), ("depth", C.c_uint32),
                ("num_nodes", C.c_uint32), ("num_leaves", C.c_uint32)]


class SphereFieldParams(C.Structure):
    _fields_ = [("num_spheres", C.c_uint32), ("subdivision", C.c_uint32), ("seed", C.c_uint64),
                ("num_groups", C.c_uint32), ("aspect", C.c_float), ("extent", C.c_float),
                ("radius_min", C.c_float), ("radius_max", C.c_float)]


class SceneDefaults(C.Structure):
    _fields_ = [("fov", C.c_float), ("aspect_ratio", C.c_float), ("eye", C.c_float * 3), ("dir", C.c_float * 3),
                ("up", C.c_float * 3)]


class SceneDesc(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("nodes", C.POINTER(GpuNode)), ("num_tris", C.c_uint32),
                ("tris", C.POINTER(GpuTri)), ("shade", C.POINTER(GpuShade)), ("num_materials", C.c_uint32),
                ("materials", C.POINTER(Material)), ("num_lights", C.c_uint32), ("lights", C.POINTER(Light)),
                ("camera", Camera), ("bvh_depth", C.c_uint32), ("top_nodes", C.c_uint32),
                ("num_nodes4", C.c_uint32), ("nodes4", C.POINTER(GpuNode4)), ("bvh4_depth", C.c_uint32),
                ("top_nodes4", C.c_uint32)]


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("num_tris", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("tri_material", C.POINTER(C.c_int32)), ("tri_light", C.POINTER(C.c_int32)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(Material)), ("num_lights", C.c_uint32),
                ("lights", C.POINTER(Light)), ("camera", Camera)]


class GpuBuildConfig(C.Structure):
    _fields_ = [("max_leaf_size", C.c_uint32), ("traversal_cost", C.c_float), ("intersection_cost", C.c_float),
                ("search_radius", C.c_uint32)]


class GpuBuildStats(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("iterations", C.c_uint32), ("num_nodes", C.c_uint32),
                ("num_nodes4", C.c_uint32), ("num_leaves", C.c_uint32), ("depth", C.c_uint32), ("depth4", C.c_uint32),
                ("sah_cost", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in ("num_nodes", "num_nodes4", "num_tris", "num_materials", "num_lights",
                                           "bvh_depth", "bvh4_depth", "top_nodes", "top_nodes4")]


class SurfaceHit(C.Structure):
    _fields_ = [("hit", C.c_int32), ("front", C.c_int32), ("material", C.c_int32), ("light", C.c_int32),
                ("t", C.c_float), ("point", C.c_float * 3), ("face_normal", C.c_float * 3), ("tangent", C.c_float * 3),
                ("bitangent", C.c_float * 3), ("normal", C.c_float * 3), ("surf_coords", C.c_float * 2),
                ("tex_coords", C.c_float * 2)]


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("lds_stack", C.c_uint32), ("path_slots", C.c_uint32),
                ("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("graph_iters", C.c_uint32), ("bvh_width", C.c_uint32),
                ("chunk_items", C.c_uint32), ("queue_segments", C.c_uint32), ("shade_queues", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("samples", C.c_uint64),
                ("launches", C.c_uint64), ("graph_replays", C.c_uint64), ("trace_ms", C.c_double),
                ("sort_ms", C.c_double), ("shade_ms", C.c_double), ("reserved_ms", C.c_double * 2),
                ("trace_launches", C.c_uint64), ("render_ms", C.c_double)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("wavefront_size", C.c_int32), ("total_memory", C.c_uint64), ("lds_per_block", C.c_int32)]


class BatchJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("sample_index", C.c_uint64),
                ("sample_count", C.c_uint32), ("pad", C.c_uint32), ("accum", C.c_void_p)]


assert C.sizeof(Material) == 64 and C.sizeof(Light) == 80 and C.sizeof(GpuNode) == 64
assert C.sizeof(GpuTri) == 48 and C.sizeof(GpuShade) == 64 and C.sizeof(Ray) == 32 and C.sizeof(Hit) == 16
assert C.sizeof(GpuNode4) == 128


def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() (make -C sol_amd)")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


hip = _load("libsolhip.so")
host = _load("libsol_host.so")

_P = C.c_void_p
_SIGS = {
    hip: {
        "solhip_last_error": (C.c_char_p, []),
        "solhip_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
        "solhip_destroy": (None, [_P]),
        "solhip_device_query": (C.c_int, [_P, C.POINTER(DeviceInfo)]),
        "solhip_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
        "solhip_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
        "solhip_set_profiling": (C.c_int, [_P, C.c_int]),
        "solhip_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
        "solhip_reset_stats": (C.c_int, [_P]),
        "solhip_scene_upload": (C.c_int, [_P, C.POINTER(SceneDesc), C.POINTER(_P)]),
        "solhip_scene_destroy": (None, [_P]),
        "solhip_scene_build": (C.c_int, [_P, C.POINTER(MeshDesc), C.POINTER(GpuBuildConfig), C.POINTER(_P),
                                         C.POINTER(GpuBuildStats)]),
        "solhip_scene_info_get": (C.c_int, [_P, C.POINTER(SceneInfo)]),
        "solhip_scene_download": (C.c_int, [_P, _P, _P, _P, _P]),
        "solhip_scene_set_materials": (C.c_int, [_P, C.POINTER(Material), C.c_uint32]),
        "solhip_scene_set_lights": (C.c_int, [_P, C.POINTER(Light), C.c_uint32]),
        "solhip_intersect": (C.c_int, [_P, _P, C.c_uint64, C.c_int, _P, C.c_int]),
        "solhip_render": (C.c_int, [_P, C.POINTER(PtConfig), C.c_int, C.c_uint32, C.c_uint32, C.c_uint64,
                                    C.c_uint32, _P, C.c_int]),
        "solhip_render_batch": (C.c_int, [_P, C.POINTER(BatchJob), C.c_uint32, C.POINTER(PtConfig)]),
        "solhip_backward": (C.c_int, [_P, C.POINTER(PtConfig), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                      _P, _P, _P, C.c_int]),
        "solhip_autotune": (C.c_int, [_P, C.POINTER(PtConfig), C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.POINTER(Tuning)]),
        "solhip_autotune_batch": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(PtConfig), C.POINTER(Tuning)]),
        "solhip_render_noalloc": (C.c_int, [_P, C.POINTER(PtConfig), C.c_uint32, C.c_uint32, C.c_uint64,
                                            C.c_uint32, _P]),
    },
    host: {
        "sol_host_last_error": (C.c_char_p, []),
        "sol_default_bvh_config": (None, [C.POINTER(BvhConfig)]),
        "sol_default_pt_config": (None, [C.POINTER(PtConfig)]),
        "sol_default_sphere_field_params": (None, [C.POINTER(SphereFieldParams)]),
        "sol_scene_cornell_box": (C.c_int, [C.c_float, C.POINTER(BvhConfig), C.POINTER(_P)]),
        "sol_scene_sphere_field": (C.c_int, [C.POINTER(SphereFieldParams), C.POINTER(BvhConfig), C.POINTER(_P)]),
        "sol_scene_triangle_soup": (C.c_int, [C.c_uint32, C.c_uint64, C.POINTER(BvhConfig), C.POINTER(_P)]),
        "sol_scene_from_arrays": (C.c_int, [C.POINTER(MeshView), C.POINTER(BvhConfig), C.POINTER(_P)]),
        "sol_default_scene_defaults": (None, [C.POINTER(SceneDefaults)]),
        "sol_image_save": (C.c_int, [C.c_char_p, C.c_int, _P, C.c_uint32, C.c_uint32, C.c_uint32]),
        "sol_scene_load": (C.c_int, [C.c_char_p, C.POINTER(SceneDefaults), C.POINTER(BvhConfig), C.POINTER(_P)]),
        "sol_scene_load_obj": (C.c_int, [C.c_char_p, C.POINTER(CameraParams), C.POINTER(BvhConfig), C.POINTER(_P)]),
        "sol_scene_destroy": (None, [_P]),
        "sol_scene_mesh": (C.c_int, [_P, C.POINTER(MeshView)]),
        "sol_scene_bvh": (C.c_int, [_P, C.POINTER(BvhView)]),
        "sol_scene_gpu_desc": (C.c_int, [_P, C.POINTER(SceneDesc)]),
        "sol_scene_build_stats": (C.c_int, [_P, C.POINTER(BuildStats)]),
        "sol_scene_set_materials": (C.c_int, [_P, C.POINTER(Material), C.c_uint32]),
        "sol_scene_set_lights": (C.c_int, [_P, C.POINTER(Light), C.c_uint32]),
        "sol_camera_from_params": (None, [C.POINTER(CameraParams), C.POINTER(Camera)]),
        "sol_gen_rays": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, C.c_float, _P]),
        "sol_scene_attach_gpu": (C.c_int, [_P, _P]),
        "sol_scene_intersect_closest": (C.c_int, [_P, _P, C.c_uint64, C.c_int, _P]),
        "sol_scene_intersect_any": (C.c_int, [_P, _P, C.c_uint64, C.c_int, _P]),
    },
}

for _lib, _sigs in _SIGS.items():
    for _name, (_res, _args) in _sigs.items():
        _fn = getattr(_lib, _name)
        _fn.restype = _res
        _fn.argtypes = _args


class SolError(RuntimeError):
    """A sol_* or solhip_* call returned a non-zero status."""

    def __init__(self, func, status, message):
        super().__init__(f"{func} failed with status {status}: {message}")
        self.status = status


def check_hip(status, func):
    if status != 0:
        raise SolError(func, status, hip.solhip_last_error().decode(errors="replace"))


def check_host(status, func):
    if status != 0:
        raise SolError(func, status, host.sol_host_last_error().decode(errors="replace"))


def exported_symbols():
    """Names every header in include/ declares, per library."""
    return {"libsolhip.so": sorted(_SIGS[hip]), "libsol_host.so": sorted(_SIGS[host])}
