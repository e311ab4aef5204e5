"""sol::Geometry (include/sol/geometry.h:33-43) answered by TriangleMesh on the GPU, against the oracle's
TriangleMesh::intersect_closest / intersect_any with sol's hit decoding (src/triangle_mesh.cpp:64-117).

Each query is a solhip_intersect launch decoded on the host: the hit flag, the side, the BSDF and
emitter, the ray's new tmax, the hit point, both normals and the shading frame, and the surface
coordinates equal the oracle's bit for bit, through sol's one-ray virtual interface and through
the batched overload. A scene uploaded without lights serves such queries but is not rendered.
"""
import ctypes as C
import os

import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

FIELDS = ("hit", "front", "material", "light", "t", "point", "face_normal", "tangent", "bitangent", "normal",
          "surf_coords")
GLOSSY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell", "cornell_box_glossy.toml")


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


def rays_inside(scene, n, seed):
    """Rays from points spread over the scene's bounding box towards uniform directions."""
    pos = scene.arrays()["positions"]
    lo, hi = pos.min(0), pos.max(0)
    rng = np.random.default_rng(seed)
    org = lo + (hi - lo) * rng.uniform(0.05, 0.95, (n, 3))
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = 1e-5
    rays[:, 4:7] = d
    rays[:, 7] = np.float32(3.4028235e38)
    rays[::7, 7] = np.float32(0.5)          # some short rays: misses within (tmin, tmax)
    return rays


@pytest.mark.parametrize("name", ["cornell", "spheres", "glossy"])
def test_geometry_queries_match_the_oracle(ctx, name):
    scene = {"cornell": lambda: sol_amd.Scene.cornell_box(),
             "spheres": lambda: sol_amd.Scene.sphere_field(num_spheres=30, subdivision=2, seed=5, extent=3.0,
                                                           radius_max=1.0),
             "glossy": lambda: sol_amd.Scene.load(GLOSSY, width=32, height=32)}[name]()
    scene.attach_gpu(ctx)
    oracle = O.OracleScene(scene)
    rays = rays_inside(scene, 4096, seed=len(name))
    want = oracle.intersect_surface(rays)
    want_any = oracle.intersect(rays, any_hit=True)
    assert want["hit"].mean() > 0.2 and (want["light"] >= 0).any()
    for batched in (True, False):
        got = scene.geometry_intersect(rays if batched else rays[:1024], batched=batched)
        n = len(got["hit"])
        print(f"{name}: batched {batched}, {n} rays, {got['hit'].mean():.3f} hit", flush=True)
        for f in FIELDS:
            np.testing.assert_array_equal(got[f], want[f][:n], err_msg=f)
        occ = scene.geometry_intersect(rays if batched else rays[:1024], any_hit=True, batched=batched)
        np.testing.assert_array_equal(occ, want_any[:n])


def test_scene_without_lights_intersects_but_does_not_render(ctx):
    # the soup's one light is a point light: no triangle's record refers to a light
    scene = sol_amd.Scene.triangle_soup(2000, seed=4)
    desc = scene.gpu_desc()
    desc.num_lights = 0
    desc.lights = None
    h = C.c_void_p()
    F.check_hip(F.hip.solhip_scene_upload(ctx.handle, C.byref(desc), C.byref(h)), "solhip_scene_upload")
    try:
        rays = sol_amd.gen_rays(0, 256, seed=3)
        out = np.zeros((256, 4), np.float32)
        F.check_hip(F.hip.solhip_intersect(h, rays.ctypes.data, 256, 0, out.ctypes.data, 0), "solhip_intersect")
        want = O.OracleScene(scene).intersect(rays)
        np.testing.assert_array_equal(out[:, 0].view(np.int32), want[0])
        img = np.zeros((3, 4, 4), np.float32)
        cfg = sol_amd.PathTracerConfig().to_c()
        st = F.hip.solhip_render(h, C.byref(cfg), F.RENDERER_MEGAKERNEL, 4, 4, 0, 1, img.ctypes.data, 0)
        assert st == -2 and b"without light" in F.hip.solhip_last_error()
    finally:
        F.hip.solhip_scene_destroy(h)
