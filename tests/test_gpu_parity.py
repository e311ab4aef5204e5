"""GPU parity: the HIP kernels, called through the C ABI, against the CPU oracle on the same inputs.

Bar: intersection results bit-exact (prim id, t, u, v; any-hit flag). Images bit-exact: the GPU
and the oracle use the same per-(pixel, sample) PCG streams, the same fp32 operation order without
contraction, correctly rounded division and square root, and cos, sin and pow evaluated in fp64 and
rounded once to fp32, so every pixel of every image is identical.
"""
import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


def compare_images(gpu, ref):
    """Bit-exact image parity; the fraction of identical pixels is printed to diagnose a failure."""
    assert gpu.shape == ref.shape
    assert np.all(np.isfinite(gpu)) and ref.any()
    same = np.all(gpu == ref, axis=0)
    print(f"bit-identical pixels {same.mean():.6f}")
    np.testing.assert_array_equal(gpu, ref)


@pytest.fixture(scope="module")
def soup():
    s = sol_amd.Scene.triangle_soup(100_000, seed=21)
    return s, O.OracleScene(s, use_product_bvh=True), O.OracleScene(s, use_product_bvh=False)


@pytest.mark.parametrize("block,lds", [(64, 16), (256, 8), (128, 32)])
def test_intersect_closest_bit_exact(ctx, soup, block, lds):
    scene, oracle, oracle_own = soup
    ctx.set_tuning(block_size=block, lds_stack=lds)
    g = ctx.upload(scene)
    rays = sol_amd.gen_rays(0, 200_000, seed=22)
    got = g.intersect(rays)
    want = oracle.intersect(rays)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    own = oracle_own.intersect(rays)              # a different tree gives the same answer
    np.testing.assert_array_equal(got[0], own[0])
    assert (got[0] >= 0).mean() > 0.1
    ctx.set_tuning(block_size=64, lds_stack=16)


def test_intersect_any_exact(ctx, soup):
    scene, oracle, _ = soup
    g = ctx.upload(scene)
    rays = sol_amd.gen_rays(1, 200_000, seed=23)
    got = g.intersect(rays, any_hit=True)
    np.testing.assert_array_equal(got, oracle.intersect(rays, any_hit=True))
    assert 0.05 < got.mean() < 0.95


def test_intersect_edge_cases(ctx):
    s = sol_amd.Scene.triangle_soup(1, seed=3)          # single leaf at the root
    g = ctx.upload(s)
    o = O.OracleScene(s, use_product_bvh=True)
    rays = sol_amd.gen_rays(0, 5000, seed=2)
    a = s.arrays()
    c = a["positions"][a["indices"][0]].mean(0)
    aim = rays.copy()
    aim[:, 4:7] = c - aim[:, :3]                        # rays through the triangle's centre
    for r in (rays, aim):
        for x, y in zip(g.intersect(r), o.intersect(r)):
            np.testing.assert_array_equal(x, y)
    assert (g.intersect(aim)[0] == 0).all()
    empty = np.zeros((0, 8), np.float32)
    assert g.intersect(empty)[0].size == 0


def _cfg(**kw):
    return sol_amd.PathTracerConfig(**kw)


@pytest.fixture(scope="module")
def cornell():
    s = sol_amd.Scene.cornell_box()
    return s, O.OracleScene(s, use_product_bvh=False)


@pytest.mark.parametrize("kind", ["megakernel", "wavefront"])
@pytest.mark.parametrize("flags", [{}, {"disable_mis": True}, {"disable_nee": True}, {"disable_rr": True}])
def test_render_cornell_matches_oracle(ctx, cornell, kind, flags):
    scene, oracle = cornell
    cfg = _cfg(max_path_len=8, **flags)
    img = sol_amd.new_image(64, 64)
    ref = sol_amd.new_image(64, 64)
    pt = sol_amd.GpuPathTracer(scene, cfg, kind=kind, ctx=ctx)
    pt.render(img, 0, 16)
    oracle.render(ref, 0, 16, cfg)
    compare_images(img, ref)


@pytest.mark.parametrize("kind", ["megakernel", "wavefront"])
def test_render_sphere_field_matches_oracle(ctx, kind):
    scene = sol_amd.Scene.sphere_field(num_spheres=40, subdivision=2, seed=5, extent=4.0, radius_max=1.0)
    oracle = O.OracleScene(scene, use_product_bvh=False)
    cfg = _cfg(max_path_len=16)
    img, ref = sol_amd.new_image(96, 54), sol_amd.new_image(96, 54)
    sol_amd.GpuPathTracer(scene, cfg, kind=kind, ctx=ctx).render(img, 3, 8)
    oracle.render(ref, 3, 8, cfg)
    compare_images(img, ref)


@pytest.mark.parametrize("kind", ["megakernel", "wavefront"])
def test_render_accumulates_and_resumes(ctx, cornell, kind):
    # Renderer::render adds into the image; render(0, 8) == render(0, 3) + render(3, 5) up to
    # the per-call summation order.
    scene, _ = cornell
    pt = sol_amd.GpuPathTracer(scene, _cfg(max_path_len=8), kind=kind, ctx=ctx)
    whole = sol_amd.new_image(32, 32)
    pt.render(whole, 0, 8)
    parts = np.full((3, 32, 32), 0.5, np.float32)
    pt.render(parts, 0, 3)
    pt.render(parts, 3, 5)
    np.testing.assert_allclose(parts - 0.5, whole, rtol=2e-5, atol=2e-5)


def test_render_rejects_bad_arguments(ctx, cornell):
    scene, _ = cornell
    g = ctx.upload(scene)
    with pytest.raises(F.SolError):
        g.render(sol_amd.new_image(0, 4), 0, 1)
    with pytest.raises(F.SolError):
        g.render(sol_amd.new_image(4, 4), 0, 1, _cfg(min_survival_prob=0.9, max_survival_prob=0.1))
    img = sol_amd.new_image(4, 4)
    g.render(img, 0, 0)                      # zero samples: no-op
    assert not img.any()


@pytest.mark.parametrize("slots", [5000, 8192, 70000])
def test_wavefront_equals_megakernel_bitwise(ctx, slots):
    # Same streams, same operations, same per-pixel summation order: identical images. 23,040 work
    # items: 5000 and 8192 slots refill many times (5000: the start fill ends in a partial round of
    # 392 slots), 70000 slots leave the last segments' start entries as holes.
    scene = sol_amd.Scene.sphere_field(num_spheres=30, subdivision=2, seed=8, extent=4.0, radius_max=1.0)
    cfg = _cfg(max_path_len=12)
    g = ctx.upload(scene)
    a, b = sol_amd.new_image(80, 48), sol_amd.new_image(80, 48)
    g.render(a, 0, 6, cfg, renderer=F.RENDERER_MEGAKERNEL)
    ctx.set_tuning(path_slots=slots, tile_w=16, tile_h=8, graph_iters=4)   # many chunks of work, tiny queues
    g.render(b, 0, 6, cfg, renderer=F.RENDERER_WAVEFRONT)
    ctx.set_tuning(path_slots=0, tile_w=0, tile_h=0, graph_iters=0)       # back to the defaults
    np.testing.assert_array_equal(a, b)


def test_material_kind_changes_reach_the_wavefront_queues(ctx):
    # The hit sorting reads each triangle's BSDF kind from its shading record: after
    # set_materials turns diffuse materials into mirrors and Phong into the diffuse-Phong mix, the
    # wavefront must still render what the megakernel (which reads the material table) renders.
    scene = sol_amd.Scene.sphere_field(num_spheres=30, subdivision=2, seed=8, extent=4.0, radius_max=1.0)
    g = ctx.upload(scene)
    mats = scene.materials()
    for m in mats:
        if m.type == F.BSDF_DIFFUSE:
            m.type = F.BSDF_MIRROR
            m.ks[:] = list(m.kd)
        elif m.type == F.BSDF_PHONG:
            m.type = F.BSDF_INTERP
            m.kd[:] = [0.3, 0.3, 0.3]
            m.mix = 0.25
    g.set_materials(mats)
    cfg = _cfg(max_path_len=10)
    a, b = sol_amd.new_image(64, 40), sol_amd.new_image(64, 40)
    g.render(a, 0, 4, cfg, renderer=F.RENDERER_MEGAKERNEL)
    g.render(b, 0, 4, cfg, renderer=F.RENDERER_WAVEFRONT)
    assert a.any()
    np.testing.assert_array_equal(a, b)


def test_wavefront_launches_only_the_queues_the_materials_use(ctx):
    # A scene whose materials are all of one kind shades with that kind's kernel (and the diffuse
    # one, which also takes the lights); changing the kinds, and back, changes the kernels the
    # captured bounce loop launches.
    scene = sol_amd.Scene.sphere_field(num_spheres=30, subdivision=2, seed=9, extent=4.0, radius_max=1.0)
    g = ctx.upload(scene)
    base = scene.materials()
    cfg = _cfg(max_path_len=10)
    a, b = sol_amd.new_image(48, 32), sol_amd.new_image(48, 32)
    for kind in (F.BSDF_DIFFUSE, F.BSDF_PHONG, F.BSDF_MIRROR, F.BSDF_GLASS, F.BSDF_INTERP, None):
        mats = scene.materials() if kind is not None else base
        if kind is not None:
            for m in mats:
                if m.type != F.BSDF_NONE:
                    m.type = kind
                    m.ks[:] = [0.4, 0.4, 0.4]
                    m.kt[:] = [1.0, 1.0, 1.0]
                    m.ns = m.ns_sample = 30.0
                    m.eta = 1.0 / 1.5
                    m.mix = 0.5
        g.set_materials(mats)
        a[:] = 0
        b[:] = 0
        g.render(a, 0, 4, cfg, renderer=F.RENDERER_MEGAKERNEL)
        g.render(b, 0, 4, cfg, renderer=F.RENDERER_WAVEFRONT)
        assert a.any()
        np.testing.assert_array_equal(a, b)


def test_triangle_records_carry_their_kind(ctx):
    # The traversal kernel packs each hit triangle's kind (BSDF + 1 in bits 0-2, light bit 3) from
    # the fourth word of its record's second row; the hit sorting reads nothing else. That word must
    # follow the shading record's material and light on upload, on a GPU build and after
    # set_materials.
    scene = sol_amd.Scene.sphere_field(num_spheres=20, subdivision=1, seed=3, extent=3.0, radius_max=1.0)

    def check(g):
        d = g.download()
        kinds = np.array([t.pad0 for t in d["tris"]])
        want = np.array([(sh.bsdf + 1) | (8 if sh.light >= 0 else 0) for sh in d["shade"]])
        np.testing.assert_array_equal(kinds, want)
        return set(kinds.tolist())

    g = ctx.upload(scene)
    seen = check(g)
    assert len(seen) >= 3                     # diffuse, Phong or glass spheres, lights, the ground
    check(ctx.build(scene))
    mats = scene.materials()
    for m in mats:
        if m.type == F.BSDF_DIFFUSE:
            m.type = F.BSDF_MIRROR
    g.set_materials(mats)
    assert check(g) != seen


def test_wavefront_profiling_mode_matches_graph_mode(ctx, cornell):
    scene, _ = cornell
    g = ctx.upload(scene)
    a, b = sol_amd.new_image(48, 40), sol_amd.new_image(48, 40)
    cfg = _cfg(max_path_len=8)
    g.render(a, 5, 4, cfg)
    ctx.set_profiling(True)
    ctx.reset_stats()
    g.render(b, 5, 4, cfg)
    st = ctx.stats()
    ctx.set_profiling(False)
    np.testing.assert_array_equal(a, b)
    assert st.trace_ms > 0 and st.trace_launches > 0
    assert st.samples == 48 * 40 * 4 and st.extension_rays >= st.samples


def test_batch_of_scenes_equals_single_renders(ctx):
    import torch
    scenes = [sol_amd.Scene.cornell_box(aspect=1.5),
              sol_amd.Scene.sphere_field(num_spheres=12, subdivision=1, seed=3, extent=3.0, aspect=1.0)]
    gs = [ctx.upload(s) for s in scenes]
    sizes = [(37, 25, 0, 5), (64, 64, 7, 3), (20, 33, 2, 4)]
    which = [0, 1, 1]
    cfg = _cfg(max_path_len=10)
    want = []
    for (w, h, si, sc), k in zip(sizes, which):
        img = sol_amd.new_image(w, h)
        gs[k].render(img, si, sc, cfg, renderer=F.RENDERER_MEGAKERNEL)
        want.append(img)
    bufs = [torch.zeros((3, h, w), dtype=torch.float32, device="cuda") for (w, h, _, _) in sizes]
    torch.cuda.synchronize()
    jobs = [(gs[k], w, h, si, sc, b.data_ptr()) for (w, h, si, sc), k, b in zip(sizes, which, bufs)]
    ctx.render_batch(jobs, cfg)
    for b, wimg in zip(bufs, want):
        np.testing.assert_array_equal(b.cpu().numpy(), wimg)


@pytest.mark.parametrize("kind", ["megakernel", "wavefront"])
def test_equal_rmse_against_converged_image(ctx, cornell, kind):
    # Acceptance criterion of the renderer: at equal spp, the GPU image's RMSE against the
    # converged (4096-spp oracle) image is within 2 % of the oracle's own, and its mean radiance
    # within 0.1 % of the oracle's.
    import os
    ref = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell_64_4096spp.npy"))
    scene, oracle = cornell
    cfg = _cfg(max_path_len=8)
    gpu, cpu = sol_amd.new_image(64, 64), sol_amd.new_image(64, 64)
    sol_amd.GpuPathTracer(scene, cfg, kind=kind, ctx=ctx).render(gpu, 100_000, 16)
    oracle.render(cpu, 100_000, 16, cfg)
    rmse_gpu = float(np.sqrt(((gpu / 16 - ref) ** 2).mean()))
    rmse_cpu = float(np.sqrt(((cpu / 16 - ref) ** 2).mean()))
    print(f"RMSE vs converged: gpu {rmse_gpu:.6f}, oracle {rmse_cpu:.6f}")
    assert abs(rmse_gpu - rmse_cpu) <= 0.02 * rmse_cpu
    assert abs(gpu.mean() - cpu.mean()) <= 1e-3 * cpu.mean()


@pytest.mark.parametrize("w,h,spp", [(5, 3, 1), (5, 3, 7), (17, 9, 3), (32, 32, 3), (70, 40, 2)])
def test_wavefront_small_renders_take_every_work_item(ctx, cornell, w, h, spp):
    # Renders with fewer work items than a round per queue segment: every item must still be
    # handed out (segments whose own range is never reached are drained by the others).
    scene, _ = cornell
    cfg = _cfg(max_path_len=8)
    a, b = sol_amd.new_image(w, h), sol_amd.new_image(w, h)
    g = ctx.upload(scene)
    g.render(a, 11, spp, cfg, renderer=F.RENDERER_WAVEFRONT)
    g.render(b, 11, spp, cfg, renderer=F.RENDERER_MEGAKERNEL)
    np.testing.assert_array_equal(a, b)


def test_binary_and_four_wide_traversal_agree(ctx):
    # The closest hit is the minimum of (t, original triangle index) over every triangle the exact
    # test accepts, so the node layout cannot change any result.
    scene = sol_amd.Scene.sphere_field(num_spheres=40, subdivision=2, seed=5, extent=4.0, radius_max=1.0)
    g = ctx.upload(scene)
    soup = ctx.upload(sol_amd.Scene.triangle_soup(50_000, seed=8))
    rays = sol_amd.gen_rays(0, 1 << 16)
    shadow = sol_amd.gen_rays(1, 1 << 16)
    cfg = _cfg(max_path_len=12)
    out = {}
    try:
        for width in (2, 4):
            ctx.set_tuning(bvh_width=width)
            img = sol_amd.new_image(96, 54)
            g.render(img, 0, 4, cfg, renderer=F.RENDERER_WAVEFRONT)
            out[width] = (img, soup.intersect(rays), soup.intersect(shadow, any_hit=True))
    finally:
        ctx.set_tuning(bvh_width=0)
    np.testing.assert_array_equal(out[2][0], out[4][0])
    for a, b in zip(out[2][1], out[4][1]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(out[2][2], out[4][2])
