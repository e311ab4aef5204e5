"""GPU side: contexts, uploaded scenes, and the GpuPathTracer renderer, over libsolhip.so.

GpuPathTracer mirrors sol::Renderer / sol::PathTracer (include/sol/renderer.h:21-64,
include/sol/algorithms/path_tracer.h:23-54 of sol): render(image, sample_index, sample_count)
adds the samples [sample_index, sample_index + sample_count) of every pixel into a planar RGB
float32 image of shape (3, height, width), pixel (0, 0) top-left.
"""
import atexit
import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _ffi as F


@dataclass
class PathTracerConfig:
    """sol::PathTracer::Config (path_tracer.h:23-29) plus run-time forms of the
    disable_mis / disable_nee / disable_rr switches (path_tracer.cpp:32-34)."""
    max_path_len: int = 64
    min_rr_path_len: int = 3
    min_survival_prob: float = 0.05
    max_survival_prob: float = 0.75
    ray_offset: float = 1e-5
    disable_mis: bool = False
    disable_nee: bool = False
    disable_rr: bool = False

    def to_c(self):
        c = F.PtConfig()
        c.max_path_len = self.max_path_len
        c.min_rr_path_len = self.min_rr_path_len
        c.min_survival_prob = self.min_survival_prob
        c.max_survival_prob = self.max_survival_prob
        c.ray_offset = self.ray_offset
        c.flags = ((F.PT_DISABLE_MIS if self.disable_mis else 0) | (F.PT_DISABLE_NEE if self.disable_nee else 0)
                   | (F.PT_DISABLE_RR if self.disable_rr else 0))
        return c


_live_scenes = weakref.WeakSet()
_live_contexts = weakref.WeakSet()


@atexit.register
def _close_all():
    # Scenes before contexts, while the HIP runtime is still up.
    for s in list(_live_scenes):
        s.close()
    for c in list(_live_contexts):
        c.close()


def new_image(width, height):
    """An empty planar RGB image, sol::Image(width, height, 3)."""
    return np.zeros((3, height, width), np.float32)


class Context:
    """One HIP device and stream (solhip_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        F.check_hip(F.hip.solhip_create(device, C.byref(h)), "solhip_create")
        self._h = h
        _live_contexts.add(self)

    def close(self):
        if self._h and self._h.value:
            F.hip.solhip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_info(self):
        info = F.DeviceInfo()
        F.check_hip(F.hip.solhip_device_query(self._h, C.byref(info)), "solhip_device_query")
        return info

    def tuning(self):
        t = F.Tuning()
        F.check_hip(F.hip.solhip_get_tuning(self._h, C.byref(t)), "solhip_get_tuning")
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        F.check_hip(F.hip.solhip_set_tuning(self._h, C.byref(t)), "solhip_set_tuning")

    def set_profiling(self, on):
        F.check_hip(F.hip.solhip_set_profiling(self._h, 1 if on else 0), "solhip_set_profiling")

    def stats(self):
        s = F.Stats()
        F.check_hip(F.hip.solhip_get_stats(self._h, C.byref(s)), "solhip_get_stats")
        return s

    def reset_stats(self):
        F.check_hip(F.hip.solhip_reset_stats(self._h), "solhip_reset_stats")

    def upload(self, scene):
        return GpuScene(self, scene)

    def build(self, scene, max_leaf_size=0, search_radius=0, traversal_cost=0.0, intersection_cost=0.0):
        """The scene's mesh with a BVH built on the GPU (solhip_scene_build); the stats of the
        build are in the result's build_stats."""
        return GpuScene(self, scene, gpu_build=F.GpuBuildConfig(max_leaf_size, traversal_cost, intersection_cost,
                                                             search_radius))

    @staticmethod
    def _jobs(jobs):
        arr = (F.BatchJob * len(jobs))()
        for i, (gs, w, h, si, sc, ptr) in enumerate(jobs):
            arr[i].scene = gs.handle.value
            arr[i].width, arr[i].height = w, h
            arr[i].sample_index, arr[i].sample_count = si, sc
            arr[i].accum = ptr
        return arr

    def render_batch(self, jobs, config=None):
        """jobs: list of (GpuScene, width, height, sample_index, sample_count, device_ptr)."""
        cfg = (config or PathTracerConfig()).to_c()
        F.check_hip(F.hip.solhip_render_batch(self._h, self._jobs(jobs), len(jobs), C.byref(cfg)),
                    "solhip_render_batch")

    def autotune_batch(self, jobs, config=None):
        """solhip_autotune_batch over the jobs of render_batch; returns the chosen Tuning (kept in
        the context)."""
        cfg = (config or PathTracerConfig()).to_c()
        t = F.Tuning()
        F.check_hip(F.hip.solhip_autotune_batch(self._h, self._jobs(jobs), len(jobs), C.byref(cfg), C.byref(t)),
                    "solhip_autotune_batch")
        return t


class GpuScene:
    """A scene uploaded to HBM (solhip_scene)."""

    def __init__(self, ctx, scene, gpu_build=None):
        self.ctx = ctx
        self.scene = scene
        self.build_stats = None
        h = C.c_void_p()
        if gpu_build is None:
            desc = scene.gpu_desc()
            F.check_hip(F.hip.solhip_scene_upload(ctx.handle, C.byref(desc), C.byref(h)), "solhip_scene_upload")
        else:
            mv = scene.mesh_view()
            md = F.MeshDesc()
            for k in ("num_vertices", "positions", "normals", "num_tris", "indices", "tri_material", "tri_light",
                      "num_materials", "materials", "num_lights", "lights"):
                setattr(md, k, getattr(mv, k))
            F.host.sol_camera_from_params(C.byref(mv.camera), C.byref(md.camera))
            st = F.GpuBuildStats()
            F.check_hip(F.hip.solhip_scene_build(ctx.handle, C.byref(md), C.byref(gpu_build), C.byref(h), C.byref(st)),
                        "solhip_scene_build")
            self.build_stats = st
        self._h = h
        _live_scenes.add(self)

    def info(self):
        i = F.SceneInfo()
        F.check_hip(F.hip.solhip_scene_info_get(self._h, C.byref(i)), "solhip_scene_info_get")
        return i

    def download(self):
        """The scene's device arrays as numpy structured arrays: nodes, nodes4, tris, shade."""
        i = self.info()
        nodes = (F.GpuNode * max(1, i.num_nodes))()
        nodes4 = (F.GpuNode4 * max(1, i.num_nodes4))()
        tris = (F.GpuTri * i.num_tris)()
        shade = (F.GpuShade * i.num_tris)()
        F.check_hip(F.hip.solhip_scene_download(self._h, C.addressof(nodes), C.addressof(nodes4) if i.num_nodes4 else None,
                                                C.addressof(tris), C.addressof(shade)), "solhip_scene_download")
        return {"nodes": nodes[:i.num_nodes], "nodes4": nodes4[:i.num_nodes4], "tris": tris[:], "shade": shade[:],
                "info": i}

    def close(self):
        if self._h and self._h.value:
            F.hip.solhip_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_materials(self, mats):
        arr = (F.Material * len(mats))(*mats)
        F.check_hip(F.hip.solhip_scene_set_materials(self._h, arr, len(mats)), "solhip_scene_set_materials")

    def set_lights(self, lights):
        arr = (F.Light * len(lights))(*lights)
        F.check_hip(F.hip.solhip_scene_set_lights(self._h, arr, len(lights)), "solhip_scene_set_lights")

    def intersect(self, rays, any_hit=False):
        """rays: (n, 8) float32 (org, tmin, dir, tmax). Returns (prim, t, u, v) arrays, or an int32
        occlusion array for any_hit."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        if any_hit:
            out = np.empty(n, np.int32)
        else:
            out = np.empty((n, 4), np.float32)
        F.check_hip(F.hip.solhip_intersect(self._h, rays.ctypes.data, n, 1 if any_hit else 0, out.ctypes.data, 0),
                    "solhip_intersect")
        if any_hit:
            return out
        return out[:, 0].view(np.int32).copy(), out[:, 1].copy(), out[:, 2].copy(), out[:, 3].copy()

    def intersect_device(self, rays_ptr, n, out_ptr, any_hit=False):
        F.check_hip(F.hip.solhip_intersect(self._h, rays_ptr, n, 1 if any_hit else 0, out_ptr, 1),
                    "solhip_intersect")

    def render(self, image, sample_index, sample_count, config=None, renderer=F.RENDERER_WAVEFRONT):
        if image.dtype != np.float32 or not image.flags.c_contiguous or image.ndim != 3 or image.shape[0] != 3:
            raise ValueError("image must be a C-contiguous float32 array of shape (3, height, width)")
        cfg = (config or PathTracerConfig()).to_c()
        F.check_hip(F.hip.solhip_render(self._h, C.byref(cfg), renderer, image.shape[2], image.shape[1],
                                        sample_index, sample_count, image.ctypes.data, 0), "solhip_render")

    def render_noalloc(self, ptr, width, height, sample_index, sample_count, config=None):
        """solhip_render_noalloc: the megakernel into the device image at ptr, allocating nothing."""
        cfg = (config or PathTracerConfig()).to_c()
        F.check_hip(F.hip.solhip_render_noalloc(self._h, C.byref(cfg), width, height, sample_index, sample_count,
                                                ptr), "solhip_render_noalloc")

    def render_device(self, ptr, width, height, sample_index, sample_count, config=None,
                      renderer=F.RENDERER_WAVEFRONT):
        cfg = (config or PathTracerConfig()).to_c()
        F.check_hip(F.hip.solhip_render(self._h, C.byref(cfg), renderer, width, height, sample_index, sample_count,
                                        ptr, 1), "solhip_render")

    def backward(self, d_image, sample_index, sample_count, config=None):
        """Adjoint path replay: returns (d_materials (M, 4), d_lights (L, 3)) for dLoss/d(image)."""
        d_image = np.ascontiguousarray(d_image, np.float32)
        mv = self.scene.mesh_view()
        dm = np.zeros((mv.num_materials, 4), np.float32)
        dl = np.zeros((mv.num_lights, 3), np.float32)
        cfg = (config or PathTracerConfig()).to_c()
        F.check_hip(F.hip.solhip_backward(self._h, C.byref(cfg), d_image.shape[2], d_image.shape[1], sample_index,
                                          sample_count, d_image.ctypes.data, dm.ctypes.data, dl.ctypes.data, 0),
                    "solhip_backward")
        return dm, dl

    def backward_device(self, d_image_ptr, width, height, sample_index, sample_count, d_materials_ptr, d_lights_ptr,
                        config=None):
        """solhip_backward on device buffers: d_image (3, height, width), d_materials (M, 4) and
        d_lights (L, 3) float32; the gradients are added to the last two."""
        cfg = (config or PathTracerConfig()).to_c()
        F.check_hip(F.hip.solhip_backward(self._h, C.byref(cfg), width, height, sample_index, sample_count,
                                          d_image_ptr, d_materials_ptr, d_lights_ptr, 1), "solhip_backward")

    def autotune(self, width, height, budget_spp=1, config=None):
        cfg = (config or PathTracerConfig()).to_c()
        t = F.Tuning()
        F.check_hip(F.hip.solhip_autotune(self._h, C.byref(cfg), width, height, budget_spp, C.byref(t)),
                    "solhip_autotune")
        return t


class GpuPathTracer:
    """Drop-in for sol::PathTracer: same Config, same render() semantics, on the GPU."""

    name = "GpuPathTracer"

    def __init__(self, scene, config=None, kind="wavefront", ctx=None, device=0):
        self.config = config or PathTracerConfig()
        self.kind = {"wavefront": F.RENDERER_WAVEFRONT, "megakernel": F.RENDERER_MEGAKERNEL}[kind]
        self.ctx = ctx or Context(device)
        self.gpu_scene = self.ctx.upload(scene)

    def render(self, image, sample_index, sample_count=1):
        self.gpu_scene.render(image, sample_index, sample_count, self.config, self.kind)
