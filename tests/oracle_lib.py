"""Loads the CPU oracle (oracle/liboracle.so) for the tests. Test infrastructure only."""
import ctypes as C
import os

import numpy as np

from sol_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
if not os.path.exists(_PATH):
    raise ImportError(f"{_PATH} is not built; run __graft_entry__.build() (make -C oracle)")
lib = C.CDLL(_PATH)


class OracleStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("extension_rays", "shadow_rays", "ext_inner_visits", "ext_tri_tests",
                                           "shadow_inner_visits", "shadow_tri_tests", "bsdf_hits", "samples")]


_P = C.c_void_p
lib.oracle_scene_create.argtypes = [C.POINTER(F.MeshView), C.POINTER(F.BvhView), C.POINTER(_P)]
lib.oracle_scene_destroy.argtypes = [_P]
lib.oracle_scene_destroy.restype = None
lib.oracle_scene_set_materials.argtypes = [_P, C.POINTER(F.Material), C.c_uint32]
lib.oracle_scene_set_lights.argtypes = [_P, C.POINTER(F.Light), C.c_uint32]
lib.oracle_intersect.argtypes = [_P, _P, C.c_uint64, C.c_int, C.c_int, _P, C.POINTER(OracleStats)]
lib.oracle_render.argtypes = [_P, C.POINTER(F.PtConfig), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, _P, C.c_int,
                              C.c_uint32, C.c_uint32, C.POINTER(OracleStats)]
lib.oracle_backward.argtypes = [_P, C.POINTER(F.PtConfig), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, _P, _P, _P,
                                C.c_int]
lib.oracle_pixel_seed.restype = C.c_uint32
lib.oracle_pixel_seed.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
lib.oracle_pcg.restype = None
lib.oracle_pcg.argtypes = [C.c_uint64, C.c_uint32, _P, _P]
for name in ("oracle_sample_cosine_hemisphere",):
    getattr(lib, name).argtypes = [C.c_float, C.c_float, _P, _P]
    getattr(lib, name).restype = None
lib.oracle_sample_cosine_power_hemisphere.argtypes = [C.c_float, C.c_float, C.c_float, _P, _P]
lib.oracle_sample_cosine_power_hemisphere.restype = None
lib.oracle_ortho_basis.argtypes = [_P, _P, _P]
lib.oracle_ortho_basis.restype = None
lib.oracle_ray_triangle.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P]
lib.oracle_bsdf_sample.argtypes = [C.POINTER(F.Material), _P, _P, C.c_int, _P, _P, _P, _P, _P]
lib.oracle_fresnel.restype = C.c_float
lib.oracle_fresnel.argtypes = [C.c_float, C.c_float, C.c_float]
lib.oracle_sample_pixel.restype = None
lib.oracle_sample_pixel.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]
lib.oracle_camera_ray.argtypes = [_P, C.c_float, C.c_float, _P, _P]
lib.oracle_intersect_surface.argtypes = [_P, _P, C.c_uint64, _P]


def f32(*v):
    return np.array(v, np.float32)


def material_struct(rec):
    """sol_material of an oracle loader material record (oracle/scene_files.py)."""
    m = F.Material()
    m.type = rec["type"]
    m.group = -1
    for key in ("kd", "ks", "kt"):
        if key in rec:
            getattr(m, key)[:] = [float(x) for x in rec[key]]
    if "ns" in rec:
        m.ns = m.ns_sample = float(rec["ns"])
    if "eta" in rec:
        m.eta = float(rec["eta"])
    if "mix" in rec:
        m.mix = float(rec["mix"])
    return m


def light_struct(rec):
    """sol_light of an oracle loader light record."""
    lt = F.Light()
    lt.type = F.LIGHT_TRIANGLE
    for name, v in zip(("v0", "v1", "v2"), rec["v"]):
        getattr(lt, name)[:] = [float(x) for x in v]
    lt.normal[:] = [float(x) for x in rec["normal"]]
    lt.inv_area = float(rec["inv_area"])
    lt.intensity[:] = [float(x) for x in rec["intensity"]]
    lt.group = -1
    return lt


class OracleScene:
    """The oracle's copy of a sol_amd.Scene (own BVH unless use_product_bvh), or of the tables of
    the oracle's scene-file loader (from_tables)."""

    def __init__(self, scene, use_product_bvh=False):
        mv = scene.mesh_view()
        bv = scene.bvh_view() if use_product_bvh else None
        self._create(mv, bv)

    def _create(self, mv, bv=None):
        h = C.c_void_p()
        if lib.oracle_scene_create(C.byref(mv), C.byref(bv) if bv is not None else None, C.byref(h)) != 0:
            raise RuntimeError("oracle_scene_create failed")
        self._h = h

    @classmethod
    def with_bvh(cls, scene, bounds, index, count, prims):
        """The oracle's copy of scene traversing the given bvh v2-layout tree: bounds (N, 6) float32
        (min x, max x, min y, max y, min z, max z), index and count (N,) (count 0: inner node with
        children at index, index + 1; else a leaf of prims[index, index + count)), prims the
        original triangle index of every leaf entry."""
        self = cls.__new__(cls)
        rows = np.zeros((len(index), 8), np.uint32)
        rows[:, :6] = np.ascontiguousarray(bounds, np.float32).view(np.uint32)
        rows[:, 6] = index
        rows[:, 7] = count
        prims = np.ascontiguousarray(prims, np.uint32)
        bv = F.BvhView()
        bv.num_nodes = rows.shape[0]
        bv.nodes = C.cast(rows.ctypes.data, C.POINTER(F.BvhNode))
        bv.num_prims = prims.shape[0]
        bv.prim_indices = prims.ctypes.data_as(C.POINTER(C.c_uint32))
        self._create(scene.mesh_view(), bv)       # the oracle copies the arrays
        return self

    @classmethod
    def with_gpu_nodes(cls, scene, nodes, prims):
        """The oracle's copy of scene traversing a tree given in the GPU's binary layout (sol_gpu_node:
        per node both child boxes and child words, a leaf as ~(first << 3 | count - 1) over the
        leaf-order triangles, whose original indices are prims), rewritten breadth-first into the
        bvh v2 layout: root box = union of the root's child boxes, an inner node's children at
        index and index + 1."""
        assert len(nodes) >= 1 and nodes[0].child[0] != nodes[0].child[1]

        def child_boxes(g):
            return [(g.b0[0], g.b0[1], g.b0[2], g.b0[3], g.bz[0], g.bz[1]),
                    (g.b1[0], g.b1[1], g.b1[2], g.b1[3], g.bz[2], g.bz[3])]

        c0, c1 = child_boxes(nodes[0])
        bounds = [(min(c0[0], c1[0]), max(c0[1], c1[1]), min(c0[2], c1[2]), max(c0[3], c1[3]),
                   min(c0[4], c1[4]), max(c0[5], c1[5]))]
        index, count = [0], [0]
        work = [(0, 0)]                       # (GPU node, row of the rewritten tree)
        head = 0
        while head < len(work):
            gi, row = work[head]
            head += 1
            g = nodes[gi]
            first = len(bounds)
            index[row] = first
            for c, box in enumerate(child_boxes(g)):
                bounds.append(box)
                ref = int(g.child[c])
                if ref >= 0:
                    index.append(0)
                    count.append(0)
                    work.append((ref, first + c))
                else:
                    code = (~ref) & 0xFFFFFFFF
                    index.append(code >> 3)
                    count.append((code & 7) + 1)
        return cls.with_bvh(scene, np.float32(bounds), np.uint32(index), np.uint32(count), np.uint32(prims))

    @classmethod
    def from_tables(cls, t):
        """t: the dict oracle.scene_files.load_scene returns (camera included)."""
        self = cls.__new__(cls)
        pos = np.ascontiguousarray(t["positions"], np.float32)
        nrm = np.ascontiguousarray(t["normals"], np.float32)
        idx = np.ascontiguousarray(t["indices"], np.uint32)
        tm = np.ascontiguousarray(t["tri_material"], np.int32)
        tl = np.ascontiguousarray(t["tri_light"], np.int32)
        mats = (F.Material * max(1, len(t["materials"])))(*[material_struct(r) for r in t["materials"]])
        lts = (F.Light * max(1, len(t["lights"])))(*[light_struct(r) for r in t["lights"]])
        mv = F.MeshView()
        mv.num_vertices = pos.shape[0]
        mv.positions = pos.ctypes.data_as(C.POINTER(C.c_float))
        mv.normals = nrm.ctypes.data_as(C.POINTER(C.c_float))
        mv.num_tris = idx.shape[0]
        mv.indices = idx.ctypes.data_as(C.POINTER(C.c_uint32))
        mv.tri_material = tm.ctypes.data_as(C.POINTER(C.c_int32))
        mv.tri_light = tl.ctypes.data_as(C.POINTER(C.c_int32))
        mv.num_materials = len(t["materials"])
        mv.materials = mats
        mv.num_lights = len(t["lights"])
        mv.lights = lts
        cam = t["camera"]
        for key in ("eye", "dir", "up"):
            getattr(mv.camera, key)[:] = [float(x) for x in cam[key]]
        mv.camera.fov = float(cam["fov"])
        mv.camera.aspect = float(cam["aspect"])
        self._create(mv)
        return self

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.oracle_scene_destroy(self._h)

    def set_materials(self, mats):
        arr = (F.Material * len(mats))(*mats)
        assert lib.oracle_scene_set_materials(self._h, arr, len(mats)) == 0

    def set_lights(self, lights):
        arr = (F.Light * len(lights))(*lights)
        assert lib.oracle_scene_set_lights(self._h, arr, len(lights)) == 0

    def camera_ray(self, u, v):
        """(origin, unit direction) of the scene camera's ray through (u, v) in [-1, 1]^2."""
        org, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
        assert lib.oracle_camera_ray(self._h, float(u), float(v), org.ctypes.data, d.ctypes.data) == 0
        return org, d

    def intersect_surface(self, rays):
        """Decoded closest hits (the dict of sol_amd.scene.surface_hits_to_numpy)."""
        from sol_amd.scene import surface_hits_to_numpy
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        out = (F.SurfaceHit * max(1, n))()
        assert lib.oracle_intersect_surface(self._h, rays.ctypes.data, n, C.addressof(out)) == 0
        return surface_hits_to_numpy(out, n)

    def intersect(self, rays, any_hit=False, brute_force=False, stats=None):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        out = np.empty(n, np.int32) if any_hit else np.empty((n, 4), np.float32)
        st = stats if stats is not None else OracleStats()
        assert lib.oracle_intersect(self._h, rays.ctypes.data, n, int(any_hit), int(brute_force), out.ctypes.data,
                                    C.byref(st)) == 0
        if any_hit:
            return out
        return out[:, 0].view(np.int32).copy(), out[:, 1].copy(), out[:, 2].copy(), out[:, 3].copy()

    def render(self, image, sample_index, sample_count, config, threads=0, rows=(0, 0), stats=None):
        assert image.dtype == np.float32 and image.flags.c_contiguous and image.shape[0] == 3
        cfg = config.to_c()
        st = stats if stats is not None else OracleStats()
        assert lib.oracle_render(self._h, C.byref(cfg), image.shape[2], image.shape[1], sample_index, sample_count,
                                 image.ctypes.data, threads, rows[0], rows[1], C.byref(st)) == 0
        return st


def backward(oscene, d_image, sample_index, sample_count, config, num_materials, num_lights, threads=0):
    """oracle_backward: (d_materials (M, 4), d_lights (L, 3)) of sum(d_image * image)."""
    d_image = np.ascontiguousarray(d_image, np.float32)
    dm = np.zeros((num_materials, 4), np.float32)
    dl = np.zeros((num_lights, 3), np.float32)
    cfg = config.to_c()
    assert lib.oracle_backward(oscene._h, C.byref(cfg), d_image.shape[2], d_image.shape[1], sample_index, sample_count,
                               d_image.ctypes.data, dm.ctypes.data, dl.ctypes.data, threads) == 0
    return dm, dl


def pcg(seed, n):
    u = np.empty(n, np.uint32)
    f = np.empty(n, np.float32)
    lib.oracle_pcg(seed, n, u.ctypes.data, f.ctypes.data)
    return u, f


def sample_pixel(sample_index, x, y, w, h):
    """(u, v) of sample_pixel for pixel (x, y) of a w x h image."""
    uv = np.zeros(2, np.float32)
    lib.oracle_sample_pixel(sample_index, x, y, w, h, uv.ctypes.data)
    return float(uv[0]), float(uv[1])


def ray_triangle(org, d, v0, v1, v2):
    t, u, v = (np.zeros(1, np.float32) for _ in range(3))
    args = [f32(*a) for a in (org, d, v0, v1, v2)]   # kept alive across the call
    hit = lib.oracle_ray_triangle(*(a.ctypes.data for a in args), t.ctypes.data, u.ctypes.data, v.ctypes.data)
    return bool(hit), float(t[0]), float(u[0]), float(v[0])


def bsdf_sample(mat, n, out, randoms, front=True):
    n, out = f32(*n), f32(*out)
    rnd = np.asarray(randoms, np.float32)
    i3, rgb = np.zeros(3, np.float32), np.zeros(3, np.float32)
    pdf, cos = np.zeros(1, np.float32), np.zeros(1, np.float32)
    ok = lib.oracle_bsdf_sample(C.byref(mat), n.ctypes.data, out.ctypes.data, int(front), rnd.ctypes.data,
                                i3.ctypes.data, pdf.ctypes.data, cos.ctypes.data, rgb.ctypes.data)
    return bool(ok), i3, float(pdf[0]), float(cos[0]), rgb
