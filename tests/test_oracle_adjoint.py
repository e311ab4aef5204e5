"""Pins the oracle's adjoint (oracle_backward) on the CPU.

* Euler's identity for emission: every term of a path carries exactly one emission factor, so the
  image is homogeneous of degree 1 in the light intensities and sum_l Le_l * dI/dLe_l = I. This
  holds for any configuration (Russian roulette included: its decisions depend on the throughput,
  not on Le).
* Euler's identity for albedo at path lengths 1 and 2, where the image is a polynomial of degree
  1 or 2 in the albedos: sum_m kd_m * dI/dkd_m = d/dt I(t kd) at t = 1, read off renders at
  t = 0, 1/2, 1.
* Central finite differences with common random numbers (h = 1e-3, Russian roulette off so that
  no decision moves with the parameters) on an L2 loss against a target rendered with perturbed
  parameters, for albedo, Phong log-exponent and emission: within 1 % relative for every parameter
  whose gradient is at least 1 % of the largest of its kind.
"""
import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from sol_amd.diff import apply_parameters, base_parameters
from tests import oracle_lib as O

H = 1e-3


def render(osc, mats, lights, w, h, si, sc, cfg):
    osc.set_materials(mats)
    osc.set_lights(lights)
    img = sol_amd.new_image(w, h)
    osc.render(img, si, sc, cfg)
    return img


def grads(osc, mats, lights, d_image, si, sc, cfg):
    osc.set_materials(mats)
    osc.set_lights(lights)
    return O.backward(osc, d_image, si, sc, cfg, len(mats), len(lights))


@pytest.fixture(scope="module")
def cornell():
    s = sol_amd.Scene.cornell_box()
    return s, O.OracleScene(s)


def test_emission_euler_identity(cornell):
    scene, osc = cornell
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    mats, lights = scene.materials(), scene.lights()
    img = render(osc, mats, lights, 24, 24, 0, 8, cfg)
    dm, dl = grads(osc, mats, lights, np.ones_like(img), 0, 8, cfg)
    _, _, le = base_parameters(mats, lights)
    lhs = float((le.astype(np.float64) * dl).sum())
    rhs = float(img.astype(np.float64).sum())
    assert abs(lhs - rhs) <= 1e-4 * rhs


@pytest.mark.parametrize("max_len", [1, 2])
def test_albedo_euler_identity(cornell, max_len):
    # With max_path_len 1 the image is affine in the albedo of the first vertex; with 2 (NEE at the
    # second vertex carries both albedos) it is quadratic. I(t kd) = I0 + t I1 + t^2 I2 from
    # t = 0, 1/2, 1, and sum kd dI/dkd = d/dt I(t kd) at t = 1 = I1 + 2 I2.
    scene, osc = cornell
    cfg = sol_amd.PathTracerConfig(max_path_len=max_len)
    mats, lights = scene.materials(), scene.lights()
    albedo, log_ns, le = base_parameters(mats, lights)
    tot = {}
    for t in (0.0, 0.5, 1.0):
        img = render(osc, *apply_parameters(mats, lights, albedo * np.float32(t), log_ns, le), 24, 24, 0, 8, cfg)
        tot[t] = img.astype(np.float64).sum()
    i2 = 2 * (tot[1.0] - 2 * tot[0.5] + tot[0.0])
    i1 = tot[1.0] - tot[0.0] - i2
    if max_len == 1:
        assert abs(i2) <= 1e-5 * abs(i1)
    dm, _ = grads(osc, mats, lights, np.ones((3, 24, 24), np.float32), 0, 8, cfg)
    lhs = float((albedo.astype(np.float64) * dm[:, :3]).sum())
    rhs = i1 + 2 * i2
    assert rhs > 0
    assert abs(lhs - rhs) <= 1e-4 * rhs


def test_gradient_is_linear_in_the_image_adjoint(cornell):
    scene, osc = cornell
    cfg = sol_amd.PathTracerConfig(max_path_len=6)
    mats, lights = scene.materials(), scene.lights()
    rng = np.random.default_rng(3)
    d1 = rng.standard_normal((3, 16, 16)).astype(np.float32)
    d2 = rng.standard_normal((3, 16, 16)).astype(np.float32)
    a = grads(osc, mats, lights, d1, 0, 4, cfg)
    b = grads(osc, mats, lights, d2, 0, 4, cfg)
    c = grads(osc, mats, lights, d1 + d2, 0, 4, cfg)
    for x, y, z in zip(a, b, c):
        np.testing.assert_allclose(z, x + y, rtol=1e-4, atol=1e-4 * np.abs(z).max())
    z = grads(osc, mats, lights, np.zeros_like(d1), 0, 4, cfg)
    assert not z[0].any() and not z[1].any()


def fd_check(osc, mats, lights, w, h, si, sc, cfg, params, seed, max_fd=None):
    """Adjoint vs central differences of L = sum((I - target)^2). Returns the compared pairs."""
    albedo, log_ns, le = params
    rng = np.random.default_rng(seed)
    t_albedo = np.clip(albedo + rng.normal(0, 0.1, albedo.shape), 0.02, 0.98).astype(np.float32)
    t_log_ns = (log_ns + rng.normal(0, 0.1, log_ns.shape)).astype(np.float32)
    t_le = np.maximum(le + rng.normal(0, 0.1, le.shape), 0).astype(np.float32)
    target = render(osc, *apply_parameters(mats, lights, t_albedo, t_log_ns, t_le), w, h, si, sc, cfg)

    def loss(a, n, e):
        img = render(osc, *apply_parameters(mats, lights, a, n, e), w, h, si, sc, cfg)
        return float(((img.astype(np.float64) - target) ** 2).sum()), img

    _, img = loss(albedo, log_ns, le)
    dm, dl = grads(osc, *apply_parameters(mats, lights, albedo, log_ns, le), 2 * (img - target), si, sc, cfg)
    # parameters: (kind, index, channel, adjoint value)
    cand = []
    for i, m in enumerate(mats):
        if m.type == F.BSDF_DIFFUSE:
            cand += [("albedo", i, c, dm[i, c]) for c in range(3)]
        elif m.type == F.BSDF_PHONG:
            cand.append(("log_ns", i, 0, dm[i, 3]))
    cand += [("emission", i, c, dl[i, c]) for i in range(len(lights)) for c in range(3)]
    if max_fd:
        # the largest max_fd adjoint entries of each kind
        kinds = {}
        for p in cand:
            kinds.setdefault(p[0], []).append(p)
        cand = [p for ps in kinds.values() for p in sorted(ps, key=lambda p: -abs(p[3]))[:max_fd]]
    pairs = []
    for kind, i, c, adj in cand:
        arrs = [albedo.copy(), log_ns.copy(), le.copy()]
        k = {"albedo": 0, "log_ns": 1, "emission": 2}[kind]
        idx = (i, c) if k != 1 else (i,)
        plus = [x.copy() for x in arrs]
        minus = [x.copy() for x in arrs]
        plus[k][idx] += H
        minus[k][idx] -= H
        fd = (loss(*plus)[0] - loss(*minus)[0]) / (2 * H)
        pairs.append((kind, i, c, float(adj), fd))
    return pairs


def assert_fd_agrees(pairs, rel=0.01):
    by_kind = {}
    for p in pairs:
        by_kind.setdefault(p[0], []).append(p)
    checked = 0
    for kind, ps in by_kind.items():
        scale = max(abs(p[4]) for p in ps)
        for k, i, c, adj, fd in ps:
            if abs(fd) < 0.01 * scale:
                continue
            checked += 1
            print(f"{k:9s} {i:3d} {c} adjoint {adj:+.6e} fd {fd:+.6e} rel {abs(adj - fd) / abs(fd):.2e}")
            assert abs(adj - fd) <= rel * abs(fd), (k, i, c, adj, fd)
    assert checked >= 3


def test_adjoint_matches_central_differences_cornell(cornell):
    scene, osc = cornell
    cfg = sol_amd.PathTracerConfig(max_path_len=6, disable_rr=True)
    mats, lights = scene.materials(), scene.lights()
    pairs = fd_check(osc, mats, lights, 24, 24, 0, 8, cfg, base_parameters(mats, lights), seed=1)
    assert any(p[0] == "albedo" for p in pairs) and any(p[0] == "emission" for p in pairs)
    assert_fd_agrees(pairs)


def test_adjoint_matches_central_differences_spheres():
    scene = sol_amd.Scene.sphere_field(num_spheres=24, subdivision=2, seed=5, extent=3.0, radius_min=0.4,
                                       radius_max=1.0, num_groups=12)
    osc = O.OracleScene(scene)
    mats, lights = scene.materials(), scene.lights()
    assert any(m.type == F.BSDF_PHONG for m in mats) and any(m.type == F.BSDF_DIFFUSE for m in mats)
    cfg = sol_amd.PathTracerConfig(max_path_len=6, disable_rr=True)
    pairs = fd_check(osc, mats, lights, 32, 18, 0, 8, cfg, base_parameters(mats, lights), seed=2, max_fd=6)
    assert any(p[0] == "log_ns" for p in pairs)
    assert_fd_agrees(pairs)
