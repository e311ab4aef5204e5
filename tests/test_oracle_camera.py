"""Pins the oracle's camera (PerspectiveCamera, src/cameras.cpp:5-25), its pixel sampling
(Renderer::sample_pixel, include/sol/renderer.h:50-54) and the mirror BSDF (bsdfs.cpp:112-119)
against what geometry fixes, computed here in float64 from the camera's definition rather than from
the oracle's expressions (CPU only):

* the horizontal field of view is the angle between the rays through the left and right edges, and
  the vertical extent is the horizontal one divided by the aspect ratio;
* the ray through the image centre is the viewing direction, the top edge lies on the side of the
  up vector (pixel (0, 0) is the top-left corner), the right edge on the side of dir x up;
* every sample of pixel (x, y) lies in that pixel's footprint, its jitter is the first two draws of
  the pixel's stream, x first;
* a render of an emitter placed off-centre lights exactly the pixels its projection covers.

A slip such as a y-flip, fov instead of fov / 2, h = w * aspect or a swapped cross product fails at
least one check.
"""
import math

import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from tests import oracle_lib as O


def _camera(eye, d, up, fov, aspect):
    c = F.CameraParams()
    c.eye[:], c.dir[:], c.up[:] = eye, d, up
    c.fov, c.aspect = fov, aspect
    return c


def _scene_with_camera(cam, tris=None, light_tris=(0,), le=(1.0, 1.0, 1.0)):
    """A scene of the given triangles (default: one far-away triangle), each in light_tris an
    emitter without BSDF, viewed by cam."""
    if tris is None:
        tris = [((100, 100, 100), (101, 100, 100), (100, 101, 100))]
    pos = np.float32([v for t in tris for v in t])
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    nrm = []
    lights = []
    for k, t in enumerate(tris):
        p = np.float64(t)
        g = np.cross(p[1] - p[0], p[2] - p[0])
        n = g / np.linalg.norm(g)
        nrm += [n] * 3
        if k in light_tris:
            lt = F.Light()
            lt.type = F.LIGHT_TRIANGLE
            lt.v0[:], lt.v1[:], lt.v2[:] = p[0], p[1], p[2]
            lt.normal[:] = n
            lt.inv_area = 2.0 / np.linalg.norm(g)
            lt.intensity[:] = le
            lt.group = -1
            lights.append(lt)
    m = F.Material()
    m.type = F.BSDF_DIFFUSE
    m.kd[:] = (0.5, 0.5, 0.5)
    m.group = -1
    tri_light = []
    for k in range(len(tris)):
        tri_light.append(light_tris.index(k) if k in light_tris else -1)
    s = sol_amd.Scene.from_arrays(pos, np.float32(nrm), idx, [-1] * len(tris), tri_light, [m], lights, cam)
    return s, O.OracleScene(s)


def _basis64(d, up, fov, aspect):
    """Unit viewing direction, unit right (dir x up), unit up' (right x dir), and the half extents
    w = tan(fov / 2), h = w / aspect of the image plane at distance 1."""
    d = np.float64(d) / np.linalg.norm(d)
    r = np.cross(d, np.float64(up))
    r /= np.linalg.norm(r)
    u = np.cross(r, d)
    w = math.tan(math.radians(fov) / 2)
    return d, r, u, w, w / aspect


CAMERAS = [
    # eye, dir, up, horizontal fov (degrees), aspect
    ((0.0, 0.9, 2.5), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 60.0, 1.0),      # sol's cornell_box.toml camera
    ((1.0, -2.0, 0.5), (0.3, 0.2, -1.0), (0.0, 1.0, 0.0), 75.0, 16 / 9),   # off-axis, wide image
    ((0.0, 0.0, 0.0), (0.1, -0.2, -1.0), (0.3, 1.0, 0.2), 40.0, 0.6),     # up not orthogonal, tall image
]


@pytest.mark.parametrize("eye,d,up,fov,aspect", CAMERAS)
def test_camera_centre_is_the_viewing_direction(eye, d, up, fov, aspect):
    _, o = _scene_with_camera(_camera(eye, d, up, fov, aspect))
    org, dirv = o.camera_ray(0.0, 0.0)
    np.testing.assert_array_equal(org, np.float32(eye))
    want = np.float64(d) / np.linalg.norm(d)
    np.testing.assert_allclose(dirv, want, atol=2e-7)


@pytest.mark.parametrize("eye,d,up,fov,aspect", CAMERAS)
def test_camera_corners_and_edges(eye, d, up, fov, aspect):
    _, o = _scene_with_camera(_camera(eye, d, up, fov, aspect))
    D, R, U, w, h = _basis64(d, up, fov, aspect)
    for su in (-1.0, 1.0):
        for sv in (-1.0, 1.0):
            want = D + su * w * R + sv * h * U
            want /= np.linalg.norm(want)
            _, got = o.camera_ray(su, sv)
            np.testing.assert_allclose(got, want, atol=1e-6)
            # the top row (v = +1) leans towards the world's up vector, the right column to dir x up
            assert np.sign(np.dot(got, U)) == sv and np.sign(np.dot(got, R)) == su
    # horizontal field of view: the angle between the left and right edge rays at v = 0
    left, right = o.camera_ray(-1.0, 0.0)[1], o.camera_ray(1.0, 0.0)[1]
    assert abs(math.degrees(math.acos(np.clip(np.dot(left, right), -1, 1))) - fov) < 1e-3
    # vertical extent: tan of half the angle between the top and bottom edge rays is w / aspect
    top, bottom = o.camera_ray(0.0, 1.0)[1], o.camera_ray(0.0, -1.0)[1]
    half = math.acos(np.clip(np.dot(top, bottom), -1, 1)) / 2
    assert abs(math.tan(half) - w / aspect) < 1e-5 * (1 + w / aspect)


def test_camera_right_handed_for_the_usual_pose():
    # looking down -z with +y up: right is +x, the top-left corner ray points to (-x, +y, -z)
    _, o = _scene_with_camera(_camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 90.0, 2.0))
    _, tl = o.camera_ray(-1.0, 1.0)
    want = np.float64([-1.0, 0.5, -1.0])
    np.testing.assert_allclose(tl, want / np.linalg.norm(want), atol=1e-6)


@pytest.mark.parametrize("w,h", [(7, 5), (16, 16), (3, 11)])
def test_sample_pixel_stays_in_the_pixel_footprint(w, h):
    for si in (0, 1, 63, 1000):
        for y in range(h):
            for x in range(w):
                u, v = O.sample_pixel(si, x, y, w, h)
                # footprint of pixel (x, y): columns from the left, rows from the top
                assert x * 2.0 / w - 1 - 1e-6 <= u <= (x + 1) * 2.0 / w - 1 + 1e-6
                assert 1 - (y + 1) * 2.0 / h - 1e-6 <= v <= 1 - y * 2.0 / h + 1e-6
    u, v = O.sample_pixel(0, 0, 0, w, h)
    assert u < -1 + 2.0 / w + 1e-6 and v > 1 - 2.0 / h - 1e-6       # pixel (0, 0) is the top-left


def test_sample_pixel_jitter_is_the_first_two_draws_x_first():
    w, h = 13, 9
    for si, x, y in [(0, 0, 0), (5, 12, 8), (77, 3, 4), (2 ** 33 + 5, 7, 1)]:
        _, f = O.pcg(O.lib.oracle_pixel_seed(si, x, y), 2)
        u, v = O.sample_pixel(si, x, y, w, h)
        jx = (u + 1.0) * w / 2.0 - x
        jy = (1.0 - v) * h / 2.0 - y
        assert abs(jx - float(f[0])) < 1e-5 and abs(jy - float(f[1])) < 1e-5


def test_render_lights_the_pixels_an_off_centre_emitter_covers():
    # Camera at the origin looking down -z, fov 90, aspect 2 (w = 1, h = 1/2); an emitting
    # rectangle at z = -2 over x in [-1.6, -0.8], y in [0.3, 0.8] projects to u = x / 2,
    # v = y (the top-left of the image). With a path length of one, a pixel sees Le for every
    # sample whose ray hits the emitter, and black otherwise.
    W, H, spp, le = 32, 16, 4, 2.0
    quad = [(-1.6, 0.3, -2.0), (-0.8, 0.3, -2.0), (-0.8, 0.8, -2.0), (-1.6, 0.8, -2.0)]
    tris = [(quad[0], quad[1], quad[2]), (quad[0], quad[2], quad[3])]   # facing +z, the camera
    s, o = _scene_with_camera(_camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 90.0, 2.0), tris, light_tris=(0, 1),
                              le=(le, le, le))
    img = sol_amd.new_image(W, H)
    o.render(img, 0, spp, sol_amd.PathTracerConfig(max_path_len=1))
    # footprint of each pixel in (u, v), and the emitter's projection
    u_lo, u_hi = -1.6 / 2, -0.8 / 2
    v_lo, v_hi = 0.3, 0.8
    lit = 0
    for y in range(H):
        for x in range(W):
            pu0, pu1 = x * 2.0 / W - 1, (x + 1) * 2.0 / W - 1
            pv0, pv1 = 1 - (y + 1) * 2.0 / H, 1 - y * 2.0 / H
            inside = pu0 >= u_lo and pu1 <= u_hi and pv0 >= v_lo and pv1 <= v_hi
            outside = pu1 <= u_lo or pu0 >= u_hi or pv1 <= v_lo or pv0 >= v_hi
            if inside:
                lit += 1
                np.testing.assert_array_equal(img[:, y, x], np.float32([spp * le] * 3))
            elif outside:
                np.testing.assert_array_equal(img[:, y, x], np.float32([0, 0, 0]))
    assert lit >= 8


@pytest.mark.parametrize("out", [(0.0, 0.0, 1.0), (0.6, 0.0, 0.8), (-0.28, 0.48, 0.83)])
def test_mirror_reflection_law(out):
    # MirrorBsdf::sample: in = reflect(-out, n) = 2 (out . n) n - out, pdf = cos = 1, colour ks,
    # no draw. The tangential part of out is reversed and the normal part kept.
    m = F.Material()
    m.type = F.BSDF_MIRROR
    m.ks[:] = (0.9, 0.5, 0.25)
    m.group = -1
    n = np.float64([0.0, 0.0, 1.0])
    o = np.float64(out) / np.linalg.norm(out)
    ok, i3, pdf, cos, rgb = O.bsdf_sample(m, n, o.astype(np.float32), [])
    assert ok and pdf == 1.0 and cos == 1.0
    np.testing.assert_array_equal(rgb, np.float32([0.9, 0.5, 0.25]))
    np.testing.assert_allclose(i3, [-o[0], -o[1], o[2]], atol=1e-6)
    # tilted normal: the angle to n is kept and in, n, out are coplanar
    nt = np.float64([0.3, -0.2, 0.9])
    nt /= np.linalg.norm(nt)
    ok, i3, _, _, _ = O.bsdf_sample(m, nt.astype(np.float32), o.astype(np.float32), [])
    if np.dot(o, nt) > 0:
        assert ok
        assert abs(np.dot(i3, nt) - np.dot(o, nt)) < 1e-6
        assert abs(np.dot(np.cross(o, nt), i3)) < 1e-6
        np.testing.assert_allclose(i3 + o, 2 * np.dot(o, nt) * nt, atol=1e-6)


def test_mirror_rejects_directions_below_the_surface():
    # out below the face: the reflected direction is below it too, and validate_sample
    # (bsdfs.h:89-94) rejects it.
    m = F.Material()
    m.type = F.BSDF_MIRROR
    m.ks[:] = (1, 1, 1)
    m.group = -1
    ok, i3, _, _, _ = O.bsdf_sample(m, (0, 0, 1), (0.6, 0.0, -0.8), [])
    assert not ok and i3[2] < 0
