"""Host-side scenes: thin Python handles over libsol_host.so (C++20 scene model, BVH builder,
flattening). A Scene owns the C++ sol::Scene and its flat arrays."""
import ctypes as C
import os

import numpy as np

from . import _ffi as F


def default_bvh_config(**overrides):
    cfg = F.BvhConfig()
    F.host.sol_default_bvh_config(C.byref(cfg))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def default_sphere_field_params(**overrides):
    p = F.SphereFieldParams()
    F.host.sol_default_sphere_field_params(C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def _array(ptr, count, dtype, shape=None):
    if count == 0:
        return np.zeros(shape or (0,), dtype=dtype)
    a = np.ctypeslib.as_array(ptr, shape=(count,)).copy()
    a = a.view(dtype) if a.dtype != dtype else a
    return a.reshape(shape) if shape else a


class Scene:
    """A procedural or array-built scene with its BVH (sweep SAH + reinsertion)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def cornell_box(cls, aspect=1.0, bvh=None):
        h = C.c_void_p()
        F.check_host(F.host.sol_scene_cornell_box(aspect, C.byref(bvh or default_bvh_config()), C.byref(h)),
                     "sol_scene_cornell_box")
        return cls(h.value)

    @classmethod
    def sphere_field(cls, params=None, bvh=None, **kw):
        p = params or default_sphere_field_params(**kw)
        h = C.c_void_p()
        F.check_host(F.host.sol_scene_sphere_field(C.byref(p), C.byref(bvh or default_bvh_config()), C.byref(h)),
                     "sol_scene_sphere_field")
        return cls(h.value)

    @classmethod
    def triangle_soup(cls, num_tris, seed=1, bvh=None):
        h = C.c_void_p()
        F.check_host(F.host.sol_scene_triangle_soup(num_tris, seed, C.byref(bvh or default_bvh_config()), C.byref(h)),
                     "sol_scene_triangle_soup")
        return cls(h.value)

    @classmethod
    def load(cls, path, width=None, height=None, aspect=None, bvh=None, **defaults):
        """A TOML scene file (sol::Scene::load). As sol's driver does, the default aspect ratio is
        width / height; other defaults (fov, eye, dir, up) may be overridden by keyword."""
        d = F.SceneDefaults()
        F.host.sol_default_scene_defaults(C.byref(d))
        if aspect is None and width and height:
            aspect = float(width) / float(height)
        if aspect is not None:
            d.aspect_ratio = aspect
        for k, v in defaults.items():
            if k in ("eye", "dir", "up"):
                getattr(d, k)[:] = [float(x) for x in v]
            else:
                setattr(d, k, v)
        h = C.c_void_p()
        F.check_host(F.host.sol_scene_load(os.fsencode(path), C.byref(d), C.byref(bvh or default_bvh_config()),
                                           C.byref(h)), "sol_scene_load")
        return cls(h.value)

    @classmethod
    def load_obj(cls, path, eye, dir, up, fov, aspect, bvh=None):
        """One OBJ file (with its MTL libraries) as the whole scene, seen through the given camera."""
        cam = F.CameraParams()
        cam.eye[:] = [float(x) for x in eye]
        cam.dir[:] = [float(x) for x in dir]
        cam.up[:] = [float(x) for x in up]
        cam.fov = fov
        cam.aspect = aspect
        h = C.c_void_p()
        F.check_host(F.host.sol_scene_load_obj(os.fsencode(path), C.byref(cam), C.byref(bvh or default_bvh_config()),
                                               C.byref(h)), "sol_scene_load_obj")
        return cls(h.value)

    @classmethod
    def from_arrays(cls, positions, normals, indices, tri_material, tri_light, materials, lights, camera, bvh=None):
        positions = np.ascontiguousarray(positions, np.float32)
        normals = np.ascontiguousarray(normals, np.float32)
        indices = np.ascontiguousarray(indices, np.uint32)
        tri_material = np.ascontiguousarray(tri_material, np.int32)
        tri_light = np.ascontiguousarray(tri_light, np.int32)
        mats = (F.Material * max(1, len(materials)))(*materials)
        lts = (F.Light * max(1, len(lights)))(*lights)
        mv = F.MeshView()
        mv.num_vertices = positions.size // 3
        mv.positions = positions.ctypes.data_as(C.POINTER(C.c_float))
        mv.normals = normals.ctypes.data_as(C.POINTER(C.c_float))
        mv.num_tris = indices.size // 3
        mv.indices = indices.ctypes.data_as(C.POINTER(C.c_uint32))
        mv.tri_material = tri_material.ctypes.data_as(C.POINTER(C.c_int32))
        mv.tri_light = tri_light.ctypes.data_as(C.POINTER(C.c_int32))
        mv.num_materials = len(materials)
        mv.materials = mats
        mv.num_lights = len(lights)
        mv.lights = lts
        mv.camera = camera
        h = C.c_void_p()
        F.check_host(F.host.sol_scene_from_arrays(C.byref(mv), C.byref(bvh or default_bvh_config()), C.byref(h)),
                     "sol_scene_from_arrays")
        return cls(h.value)

    def close(self):
        if self._h and self._h.value:
            F.host.sol_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- flat views ---------------------------------------------------------------------------
    def mesh_view(self):
        mv = F.MeshView()
        F.check_host(F.host.sol_scene_mesh(self._h, C.byref(mv)), "sol_scene_mesh")
        return mv

    def bvh_view(self):
        bv = F.BvhView()
        F.check_host(F.host.sol_scene_bvh(self._h, C.byref(bv)), "sol_scene_bvh")
        return bv

    def gpu_desc(self):
        d = F.SceneDesc()
        F.check_host(F.host.sol_scene_gpu_desc(self._h, C.byref(d)), "sol_scene_gpu_desc")
        return d

    def build_stats(self):
        s = F.BuildStats()
        F.check_host(F.host.sol_scene_build_stats(self._h, C.byref(s)), "sol_scene_build_stats")
        return s

    @property
    def num_tris(self):
        return self.mesh_view().num_tris

    def materials(self):
        mv = self.mesh_view()
        return [mv.materials[i] for i in range(mv.num_materials)]

    def lights(self):
        mv = self.mesh_view()
        return [mv.lights[i] for i in range(mv.num_lights)]

    def set_materials(self, mats):
        arr = (F.Material * len(mats))(*mats)
        F.check_host(F.host.sol_scene_set_materials(self._h, arr, len(mats)), "sol_scene_set_materials")

    def set_lights(self, lights):
        arr = (F.Light * len(lights))(*lights)
        F.check_host(F.host.sol_scene_set_lights(self._h, arr, len(lights)), "sol_scene_set_lights")

    # -- sol::Geometry of the root mesh ---------------------------------------------------------
    def attach_gpu(self, ctx):
        """TriangleMesh::attach_gpu: the mesh's BVH and triangles uploaded to the context for the
        Geometry queries below (the context must outlive the attachment)."""
        F.check_host(F.host.sol_scene_attach_gpu(self._h, ctx.handle), "sol_scene_attach_gpu")
        self._attached_ctx = ctx

    def geometry_intersect(self, rays, any_hit=False, batched=True):
        """The root's Geometry::intersect_closest / intersect_any on rays (n, 8) float32 (org, tmin,
        dir, tmax): one virtual call per ray (batched False) or one batched call. Closest: a
        SurfaceHit structured array; any: int32 flags."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        if any_hit:
            out = np.zeros(n, np.int32)
            F.check_host(F.host.sol_scene_intersect_any(self._h, rays.ctypes.data, n, int(batched), out.ctypes.data),
                         "sol_scene_intersect_any")
            return out
        out = (F.SurfaceHit * max(1, n))()
        F.check_host(F.host.sol_scene_intersect_closest(self._h, rays.ctypes.data, n, int(batched), C.addressof(out)),
                     "sol_scene_intersect_closest")
        return surface_hits_to_numpy(out, n)

    def arrays(self):
        """Copies of the raw mesh arrays (original triangle order)."""
        mv = self.mesh_view()
        return {
            "positions": _array(mv.positions, mv.num_vertices * 3, np.float32, (mv.num_vertices, 3)),
            "normals": _array(mv.normals, mv.num_vertices * 3, np.float32, (mv.num_vertices, 3)),
            "indices": _array(mv.indices, mv.num_tris * 3, np.uint32, (mv.num_tris, 3)),
            "tri_material": _array(mv.tri_material, mv.num_tris, np.int32),
            "tri_light": _array(mv.tri_light, mv.num_tris, np.int32),
        }

    def bvh_arrays(self):
        bv = self.bvh_view()
        nodes = np.ctypeslib.as_array(C.cast(bv.nodes, C.POINTER(C.c_uint32)), shape=(bv.num_nodes * 8,)).copy()
        nodes = nodes.reshape(bv.num_nodes, 8)
        bounds = nodes[:, :6].view(np.float32)
        prims = _array(bv.prim_indices, bv.num_prims, np.uint32)
        return bounds, nodes[:, 6].copy(), nodes[:, 7].copy(), prims


def surface_hits_to_numpy(arr, n):
    """Dict of numpy arrays from a ctypes array of SurfaceHit records."""
    raw = np.ctypeslib.as_array(C.cast(arr, C.POINTER(C.c_uint32)), shape=(max(1, n) * 24,)).copy()
    raw = raw.reshape(-1, 24)[:n]
    f = raw.view(np.float32)
    return {"hit": raw[:, 0].view(np.int32), "front": raw[:, 1].view(np.int32), "material": raw[:, 2].view(np.int32),
            "light": raw[:, 3].view(np.int32), "t": f[:, 4], "point": f[:, 5:8], "face_normal": f[:, 8:11],
            "tangent": f[:, 11:14], "bitangent": f[:, 14:17], "normal": f[:, 17:20], "surf_coords": f[:, 20:22],
            "tex_coords": f[:, 22:24]}


def gen_rays(kind, n, seed=7, offset=1e-5):
    """Ray batch of the intersection benchmark as a structured (n, 8) float32 array:
    org xyz, tmin, dir xyz, tmax."""
    out = np.empty((n, 8), np.float32)
    F.check_host(F.host.sol_gen_rays(kind, n, seed, offset, out.ctypes.data), "sol_gen_rays")
    return out


def camera_basis(params):
    cam = F.Camera()
    F.host.sol_camera_from_params(C.byref(params), C.byref(cam))
    return cam


_IMAGE_FORMATS = {"auto": F.IMAGE_AUTO, "png": F.IMAGE_PNG, "jpeg": F.IMAGE_JPEG, "tiff": F.IMAGE_TIFF,
                  "exr": F.IMAGE_EXR, "pfm": F.IMAGE_PFM}


def save_image(path, image, format="auto"):
    """Writes a planar (channels, height, width) float32 image as sol::Image::save does: EXR by
    default, or PFM; PNG / JPEG / TIFF have no codec in this library and raise."""
    img = np.ascontiguousarray(image, np.float32)
    if img.ndim != 3:
        raise ValueError("image must be (channels, height, width)")
    F.check_host(F.host.sol_image_save(os.fsencode(path), _IMAGE_FORMATS[format], img.ctypes.data, img.shape[2],
                                       img.shape[1], img.shape[0]), "sol_image_save")
