"""GPU adjoint (solhip_backward) against the oracle's adjoint and against central finite differences
of GPU renders; the torch.autograd.Function wrapper against the direct call.

The GPU replays each path and subtracts what it has attributed (path replay); the oracle stores the
path and sums suffixes. The two agree to float rounding (1e-3 relative, gradients summed with
atomics in a different order). Finite differences: h = 1e-3, common random numbers, Russian
roulette off, L2 loss against a target rendered with parameters perturbed by N(0, 0.1^2); within
1 % relative for every parameter whose gradient is at least 1 % of the largest of its kind.
"""
import numpy as np
import pytest
import torch

import sol_amd
from sol_amd import _ffi as F
from sol_amd.diff import DiffRenderer, apply_parameters, base_parameters
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

H = 1e-3


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


def small_spheres():
    return sol_amd.Scene.sphere_field(num_spheres=24, subdivision=2, seed=5, extent=3.0, radius_min=0.4,
                                      radius_max=1.0, num_groups=12)


def compare(a, b, rtol=1e-3):
    scale = max(np.abs(b).max(), 1e-12)
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * scale)


@pytest.mark.parametrize("scene_kind", ["cornell", "spheres"])
@pytest.mark.parametrize("rr", [True, False])
def test_adjoint_matches_oracle_adjoint(ctx, scene_kind, rr):
    scene = sol_amd.Scene.cornell_box() if scene_kind == "cornell" else small_spheres()
    cfg = sol_amd.PathTracerConfig(max_path_len=8, disable_rr=not rr)
    w, h, spp = 40, 32, 4
    rng = np.random.default_rng(11)
    d_image = rng.standard_normal((3, h, w)).astype(np.float32)
    g = ctx.upload(scene)
    dm, dl = g.backward(d_image, 5, spp, cfg)
    mats, lights = scene.materials(), scene.lights()
    om, ol = O.backward(O.OracleScene(scene), d_image, 5, spp, cfg, len(mats), len(lights))
    print(f"max |d_materials| {np.abs(om).max():.4e}, max |d_lights| {np.abs(ol).max():.4e}")
    compare(dm, om)
    compare(dl, ol)


def test_adjoint_matches_oracle_at_256_material_groups(ctx):
    # The differentiable configuration's scene: the 400-sphere field (2,048,008 triangles) with 256
    # material groups, depth 16, Russian roulette off; a loss gradient over a 256 x 256 image at 4 spp,
    # the oracle traversing the same BVH.
    scene = sol_amd.Scene.sphere_field(num_groups=256, aspect=1.0)
    mats, lights = scene.materials(), scene.lights()
    assert len(mats) >= 256
    cfg = sol_amd.PathTracerConfig(max_path_len=16, disable_rr=True)
    w = h = 256
    d_image = np.random.default_rng(17).standard_normal((3, h, w)).astype(np.float32)
    print(f"adjoint: {scene.num_tris} triangles, {len(mats)} materials, {w}x{h} 4 spp", flush=True)
    dm, dl = ctx.upload(scene).backward(d_image, 0, 4, cfg)
    om, ol = O.backward(O.OracleScene(scene, use_product_bvh=True), d_image, 0, 4, cfg, len(mats), len(lights))
    # the diffuse groups (about a third) receive albedo gradients, the Phong ones exponent gradients
    assert (np.abs(om[:, :3]).max(axis=1) > 0).sum() >= 64 and (np.abs(om[:, 3]) > 0).sum() >= 64
    compare(dm, om)
    compare(dl, ol)


def test_adjoint_host_and_device_pointers_agree_and_accumulate(ctx):
    scene = sol_amd.Scene.cornell_box()
    cfg = sol_amd.PathTracerConfig(max_path_len=6)
    w, h = 32, 32
    d_image = np.random.default_rng(2).standard_normal((3, h, w)).astype(np.float32)
    g = ctx.upload(scene)
    dm, dl = g.backward(d_image, 0, 4, cfg)
    d_img = torch.from_numpy(d_image).cuda()
    tm = torch.ones((len(scene.materials()), 4), dtype=torch.float32, device="cuda")
    tl = torch.ones((len(scene.lights()), 3), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    g.backward_device(d_img.data_ptr(), w, h, 0, 4, tm.data_ptr(), tl.data_ptr(), cfg)
    compare(tm.cpu().numpy() - 1.0, dm, rtol=1e-4)
    compare(tl.cpu().numpy() - 1.0, dl, rtol=1e-4)


def fd_check(dr, seed, kinds_wanted, per_kind=6, min_checked=6, types=None):
    """Adjoint gradients of an L2 loss against a target rendered with parameters perturbed by
    N(0, 0.1^2), against central differences (h = H, common random numbers) of GPU renders through
    DiffRenderer: the largest `per_kind` gradients of each kind, each parameter of a material or
    light the renderer reads (of materials of the given BSDF types, if types is given)."""
    albedo, log_ns, le = dr.parameters()
    rng = np.random.default_rng(seed)
    t_albedo = np.clip(albedo + rng.normal(0, 0.1, albedo.shape), 0.02, 0.98).astype(np.float32)
    t_log_ns = (log_ns + rng.normal(0, 0.1, log_ns.shape)).astype(np.float32)
    t_le = np.maximum(le + rng.normal(0, 0.1, le.shape), 0).astype(np.float32)
    target = dr.render_with(t_albedo, t_log_ns, t_le).double()

    def loss(a, n, e):
        return float(((dr.render_with(a, n, e).double() - target) ** 2).sum())

    img = dr.render_with(albedo, log_ns, le)
    da, dn, de = dr.gradients(2 * (img.double() - target).float(), albedo, log_ns, le)
    da, dn, de = da.cpu().numpy(), dn.cpu().numpy(), de.cpu().numpy()
    cand = []
    for i, m in enumerate(dr.materials):
        picked = types is None or m.type in types
        if m.type in (F.BSDF_DIFFUSE, F.BSDF_INTERP) and picked:
            cand += [("albedo", (i, c), da[i, c]) for c in range(3)]
        if m.type in (F.BSDF_PHONG, F.BSDF_INTERP) and picked:
            cand.append(("log_ns", (i,), dn[i]))
        if m.type not in (F.BSDF_DIFFUSE, F.BSDF_INTERP):
            assert not da[i].any()            # parameters the renderer does not read get no gradient
        if m.type not in (F.BSDF_PHONG, F.BSDF_INTERP):
            assert dn[i] == 0
    cand += [("emission", (i, c), de[i, c]) for i in range(len(dr.lights)) for c in range(3)]
    kinds = {}
    for p in cand:
        kinds.setdefault(p[0], []).append(p)
    assert set(kinds) == set(kinds_wanted)
    checked = 0
    for kind, ps in kinds.items():
        ps = sorted(ps, key=lambda p: -abs(p[2]))[:per_kind]
        k = {"albedo": 0, "log_ns": 1, "emission": 2}[kind]
        fds = []
        for _, idx, adj in ps:
            plus = [albedo.copy(), log_ns.copy(), le.copy()]
            minus = [albedo.copy(), log_ns.copy(), le.copy()]
            plus[k][idx] += H
            minus[k][idx] -= H
            fds.append((idx, float(adj), (loss(*plus) - loss(*minus)) / (2 * H)))
        scale = max(abs(f[2]) for f in fds)
        for idx, adj, fd in fds:
            print(f"{kind:9s} {idx} adjoint {adj:+.6e} fd {fd:+.6e}")
            if abs(fd) >= 0.01 * scale:
                checked += 1
                assert abs(adj - fd) <= 0.01 * abs(fd), (kind, idx, adj, fd)
    assert checked >= min_checked


def test_adjoint_matches_central_differences(ctx):
    cfg = sol_amd.PathTracerConfig(max_path_len=8, disable_rr=True)
    dr = DiffRenderer(ctx.upload(small_spheres()), 128, 72, 0, 8, cfg)
    fd_check(dr, 4, {"albedo", "log_ns", "emission"})


def test_adjoint_matches_central_differences_on_mixed_bsdfs(ctx):
    # sol's glossy Cornell box: diffuse walls, Phong materials and two diffuse-Phong mixes (the
    # floor's and the tall box's), whose albedo and Phong exponent DiffRenderer moves and
    # solhip_backward differentiates.
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell", "cornell_box_glossy.toml")
    scene = sol_amd.Scene.load(path, width=64, height=48)
    assert any(m.type == F.BSDF_INTERP for m in scene.materials())
    cfg = sol_amd.PathTracerConfig(max_path_len=8, disable_rr=True)
    dr = DiffRenderer(ctx.upload(scene), 64, 48, 0, 16, cfg)
    fd_check(dr, 9, {"albedo", "log_ns", "emission"}, per_kind=6, min_checked=6, types={F.BSDF_INTERP})


def test_autograd_function_matches_direct_call(ctx):
    scene = small_spheres()
    cfg = sol_amd.PathTracerConfig(max_path_len=6, disable_rr=True)
    dr = DiffRenderer(ctx.upload(scene), 64, 36, 0, 4, cfg)
    albedo, log_ns, le = dr.parameters()
    target = dr.render_with(albedo * np.float32(0.9), log_ns + np.float32(0.1), le * np.float32(1.1))
    a = torch.tensor(albedo, requires_grad=True)
    n = torch.tensor(log_ns, requires_grad=True)
    e = torch.tensor(le, requires_grad=True)
    img = dr(a, n, e)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    da, dn, de = dr.gradients(2 * (img.detach() - target), albedo, log_ns, le)
    compare(a.grad.numpy(), da.cpu().numpy(), rtol=1e-4)
    compare(n.grad.numpy(), dn.cpu().numpy(), rtol=1e-4)
    compare(e.grad.numpy(), de.cpu().numpy(), rtol=1e-4)
    # the forward pass is the ordinary renderer with the parameter tables applied
    ref = sol_amd.new_image(64, 36)
    g2 = ctx.upload(scene)
    mats, lights = apply_parameters(dr.materials, dr.lights, albedo, log_ns, le)
    g2.set_materials(mats)
    g2.set_lights(lights)
    g2.render(ref, 0, 4, cfg)
    np.testing.assert_array_equal(img.detach().cpu().numpy(), ref)


def test_backward_rejects_bad_arguments(ctx):
    g = ctx.upload(sol_amd.Scene.cornell_box())
    with pytest.raises(F.SolError):
        g.backward(np.zeros((3, 0, 4), np.float32), 0, 1)
