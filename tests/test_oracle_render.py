"""Pins the oracle's path tracer against what the mathematics fixes (CPU only):
closed-form direct lighting from a point light, the furnace identity of a closed emitting box,
and the agreement of the three estimators sol can switch between (MIS, NEE only, BSDF only)."""
import math
import os

import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from tests import oracle_lib as O


def _material(kind, **kw):
    m = F.Material()
    m.type = kind
    m.group = -1
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = v
        else:
            setattr(m, k, v)
    return m


def _camera(eye, d, fov=60.0, aspect=1.0, up=(0, 1, 0)):
    c = F.CameraParams()
    c.eye[:], c.dir[:], c.up[:] = eye, d, up
    c.fov, c.aspect = fov, aspect
    return c


def _quad(p0, p1, p2, p3):
    pos = np.float32([p0, p1, p2, p3])
    n = np.cross(pos[1] - pos[0], pos[2] - pos[0])
    n /= np.linalg.norm(n)
    return pos, np.float32([n] * 4), np.uint32([[0, 1, 2], [0, 2, 3]])


def test_point_light_direct_lighting_closed_form():
    # A diffuse floor of albedo rho lit by a point light of intensity I at height H: the radiance
    # seen at the floor point below the light is rho / pi * I / H^2 (one bounce, no other
    # surface). The camera looks straight down at that point through the central pixel.
    rho, I, H = 0.6, 10.0, 2.0
    pos, nrm, idx = _quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50))
    light = F.Light()
    light.type = F.LIGHT_POINT
    light.v0[:] = (0, H, 0)
    light.intensity[:] = (I, I, I)
    light.group = -1
    cam = _camera((0, 1.0, 0), (0, -1, 0), fov=0.05, up=(0, 0, -1))
    s = sol_amd.Scene.from_arrays(pos, nrm, idx, [0, 0], [-1, -1], [_material(F.BSDF_DIFFUSE, kd=(rho, rho, rho))],
                                  [light], cam)
    img = sol_amd.new_image(3, 3)
    O.OracleScene(s).render(img, 0, 16, sol_amd.PathTracerConfig(max_path_len=1))
    want = rho / math.pi * I / H ** 2
    assert abs(img[0, 1, 1] / 16 - want) < 1e-4 * want


def _inward_box(albedo, le):
    # Closed cube [-1,1]^3 with every face facing inwards, emissive (le) and diffuse (albedo).
    c = np.float32([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)])
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    pos, nrm, idx = [], [], []
    for f in faces:
        p = c[list(f)]
        n = -np.cross(p[1] - p[0], p[2] - p[0])
        n /= np.linalg.norm(n)
        base = len(pos)
        # wind so that the geometric normal points inwards
        for q in (p[0], p[3], p[2], p[1]):
            pos.append(q)
            nrm.append(-n if np.dot(-n, -q) > 0 else n)
        idx += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
    pos, nrm, idx = np.float32(pos), np.float32(nrm), np.uint32(idx)
    geo = np.cross(pos[idx[:, 1]] - pos[idx[:, 0]], pos[idx[:, 2]] - pos[idx[:, 0]])
    assert np.all(np.einsum("ij,ij->i", geo, -pos[idx].mean(1)) > 0), "faces must face the centre"
    lights = []
    for t in range(len(idx)):
        v = pos[idx[t]]
        l = F.Light()
        l.type = F.LIGHT_TRIANGLE
        l.v0[:], l.v1[:], l.v2[:] = v[0], v[1], v[2]
        g = np.cross(v[1] - v[0], v[2] - v[0])
        l.normal[:] = g / np.linalg.norm(g)
        l.inv_area = 2.0 / np.linalg.norm(g)
        l.intensity[:] = (le, le, le)
        l.group = -1
        lights.append(l)
    return sol_amd.Scene.from_arrays(pos, nrm, idx, [0] * len(idx), list(range(len(idx))),
                                     [_material(F.BSDF_DIFFUSE, kd=(albedo,) * 3)], lights,
                                     _camera((0.1, 0.05, 0.0), (0.3, -0.2, -1.0)))


@pytest.mark.parametrize("flags", [{}, {"disable_mis": True}, {"disable_nee": True}])
def test_furnace_closed_box(flags):
    # Inside a closed box whose walls emit Le and reflect a fraction a, the radiance is
    # Le / (1 - a) everywhere, whatever the estimator.
    a, le = 0.5, 1.0
    scene = _inward_box(a, le)
    img = sol_amd.new_image(16, 16)
    spp = 64
    O.OracleScene(scene).render(img, 0, spp, sol_amd.PathTracerConfig(max_path_len=40, **flags))
    mean = img.mean() / spp
    assert abs(mean - le / (1 - a)) < 0.01 * le / (1 - a)


def test_estimators_agree_on_cornell_box():
    # MIS, NEE-only and BSDF-only (path_tracer.cpp:32-34) estimate the same image.
    scene = sol_amd.Scene.cornell_box()
    o = O.OracleScene(scene)
    means = {}
    for name, flags in {"mis": {}, "nee": {"disable_mis": True}, "bsdf": {"disable_nee": True}}.items():
        img = sol_amd.new_image(24, 24)
        o.render(img, 1000, 256, sol_amd.PathTracerConfig(max_path_len=8, **flags))
        means[name] = img.reshape(3, -1).mean(1) / 256
    for k in ("nee", "bsdf"):
        np.testing.assert_allclose(means[k], means["mis"], rtol=0.04)


def test_oracle_render_is_deterministic_and_thread_independent():
    scene = sol_amd.Scene.cornell_box()
    o = O.OracleScene(scene)
    a, b = sol_amd.new_image(20, 16), sol_amd.new_image(20, 16)
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    o.render(a, 3, 4, cfg, threads=1)
    o.render(b, 3, 4, cfg, threads=8)
    np.testing.assert_array_equal(a, b)
    # row ranges compose
    c = sol_amd.new_image(20, 16)
    o.render(c, 3, 4, cfg, rows=(0, 7))
    o.render(c, 3, 4, cfg, rows=(7, 16))
    np.testing.assert_array_equal(a, c)


def test_oracle_product_bvh_and_own_bvh_render_identically():
    scene = sol_amd.Scene.sphere_field(num_spheres=15, subdivision=2, seed=4, extent=3.0)
    a, b = sol_amd.new_image(32, 18), sol_amd.new_image(32, 18)
    cfg = sol_amd.PathTracerConfig(max_path_len=16)
    O.OracleScene(scene, use_product_bvh=True).render(a, 0, 4, cfg)
    O.OracleScene(scene, use_product_bvh=False).render(b, 0, 4, cfg)
    np.testing.assert_array_equal(a, b)


CONVERGED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell_64_4096spp.npy")


def test_converged_cornell_fixture_reproduces_and_converges():
    # The committed 4096-spp image is the oracle's own (rows 0..3 re-rendered bit for bit), and a
    # 256-spp render from other sample indices agrees with it to Monte Carlo accuracy.
    ref = np.load(CONVERGED)
    scene = sol_amd.Scene.cornell_box()
    osc = O.OracleScene(scene)
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    part = sol_amd.new_image(64, 64)
    osc.render(part, 0, 4096, cfg, rows=(0, 4))
    np.testing.assert_array_equal((part / np.float32(4096)).astype(np.float32)[:, :4], ref[:, :4])
    img = sol_amd.new_image(64, 64)
    osc.render(img, 100_000, 256, cfg)
    img /= 256
    assert abs(img.mean() - ref.mean()) <= 0.01 * ref.mean()
    rmse_256 = np.sqrt(((img - ref) ** 2).mean())
    img16 = sol_amd.new_image(64, 64)
    osc.render(img16, 100_000, 16, cfg)
    rmse_16 = np.sqrt(((img16 / 16 - ref) ** 2).mean())
    # Monte Carlo error falls as 1 / sqrt(spp): 16x the samples, about a quarter of the error
    assert 0.15 < rmse_256 / rmse_16 < 0.4
