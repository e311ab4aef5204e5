"""GPU BVH construction (solhip_scene_build).

The reference builds its BVH on the host; a BVH only decides which nodes a ray visits, never which
hit it returns (the closest hit is unique by t and original triangle index), so a GPU-built scene
must render and intersect bit for bit like the host-built upload of the same mesh. Before any ray
touches a built tree, its structure is checked on the host: every reference in range, the leaves a
partition of the triangles, child boxes containing their triangles, the leaf-order triangle and
shading records equal to the mesh's, the 4-wide layout referring to the same leaves.
"""
import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

STACK = 8 + 88          # LDS + spill entries of the traversal stack


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


def scenes():
    return {
        "cornell": lambda: sol_amd.Scene.cornell_box(),
        "spheres": lambda: sol_amd.Scene.sphere_field(num_spheres=60, subdivision=2, seed=9, extent=4.0,
                                                      radius_max=1.0),
        "soup": lambda: sol_amd.Scene.triangle_soup(20000, seed=4),
        "water": lambda: sol_amd.Scene.load("tests/golden/cornell/cornell_box_water.toml", 64, 48),
    }


def decode(ref):
    code = (~int(ref)) & 0xFFFFFFFF
    return code >> 3, (code & 7) + 1


def check_structure(scene, built):
    a = scene.arrays()
    pos = a["positions"].reshape(-1, 3)
    idx = a["indices"].reshape(-1, 3)
    d = built.download()
    info = d["info"]
    n = idx.shape[0]
    assert info.num_tris == n and info.bvh_depth <= STACK and 3 * (info.bvh4_depth - 1) <= STACK
    tris = d["tris"]
    prim = np.array([t.prim for t in tris])
    assert sorted(prim.tolist()) == list(range(n))
    v = np.array([[list(t.v0), list(t.v1), list(t.v2)] for t in tris], np.float32)
    np.testing.assert_array_equal(v, pos[idx[prim]])
    # shading records equal the host flattening's record of the same triangle
    up = scene.gpu_desc()
    hp = np.array([up.tris[i].prim for i in range(n)])
    hshade = {int(p): bytes(up.shade[i]) for i, p in enumerate(hp)}
    assert all(bytes(s) == hshade[int(p)] for s, p in zip(d["shade"], prim))
    # binary nodes: references in range, leaves partition [0, n), boxes contain their triangles
    seen = np.zeros(n, np.int32)
    leaves = set()
    nodes = d["nodes"]
    for k, g in enumerate(nodes):
        boxes = [(g.b0[0], g.b0[1], g.b0[2], g.b0[3], g.bz[0], g.bz[1]),
                 (g.b1[0], g.b1[1], g.b1[2], g.b1[3], g.bz[2], g.bz[3])]
        for c in range(2 if len(nodes) > 1 or g.child[0] != g.child[1] else 1):
            ref = g.child[c]
            lx, hx, ly, hy, lz, hz = boxes[c]
            if ref >= 0:
                assert k < ref < len(nodes)       # breadth-first: children after parents
                ch = nodes[ref]
                for b in ((ch.b0[0], ch.b0[1], ch.b0[2], ch.b0[3], ch.bz[0], ch.bz[1]),
                          (ch.b1[0], ch.b1[1], ch.b1[2], ch.b1[3], ch.bz[2], ch.bz[3])):
                    assert lx <= b[0] and b[1] <= hx and ly <= b[2] and b[3] <= hy and lz <= b[4] and b[5] <= hz
            else:
                first, count = decode(ref)
                assert first + count <= n
                leaves.add(ref)
                seen[first:first + count] += 1
                tv = v[first:first + count].reshape(-1, 3)
                assert (tv[:, 0] >= lx).all() and (tv[:, 0] <= hx).all()
                assert (tv[:, 1] >= ly).all() and (tv[:, 1] <= hy).all()
                assert (tv[:, 2] >= lz).all() and (tv[:, 2] <= hz).all()
    assert (seen == 1).all()
    # 4-wide nodes: the same leaves, every node reachable once
    refs4 = set()
    reached = np.zeros(len(d["nodes4"]), np.int32)
    reached[0] = 1
    for k, g in enumerate(d["nodes4"]):
        used = [c for c in g.child if c != F.GPU_EMPTY]
        assert used
        for s in range(4):   # unused slots: the inverted box, which every ray misses
            if g.child[s] == F.GPU_EMPTY:
                assert g.lo_x[s] == g.lo_y[s] == g.lo_z[s] == np.inf
                assert g.hi_x[s] == g.hi_y[s] == g.hi_z[s] == -np.inf
        for c in used:
            if c >= 0:
                assert k < c < len(d["nodes4"])
                reached[c] += 1
            else:
                refs4.add(c)
    assert (reached == 1).all() and refs4 == leaves
    assert info.top_nodes <= 63 and info.top_nodes4 <= 85


@pytest.mark.parametrize("name", ["cornell", "spheres", "soup", "water"])
def test_gpu_built_tree_is_valid(ctx, name):
    scene = scenes()[name]()
    print(f"build {name} ({scene.num_tris} triangles)", flush=True)
    built = ctx.build(scene)
    st = built.build_stats
    print(f"  {st.build_ms:.1f} ms, {st.iterations} passes, {st.num_nodes} nodes, depth {st.depth}, "
          f"4-wide {st.num_nodes4} nodes depth {st.depth4}, SAH {st.sah_cost:.2f}", flush=True)
    check_structure(scene, built)


@pytest.mark.parametrize("n", [1, 2, 3, 9])
def test_tiny_and_degenerate_meshes(ctx, n):
    rng = np.random.default_rng(n)
    pos = rng.random((3 * n, 3)).astype(np.float32)
    if n == 9:
        pos[:] = 0.5                                # every triangle degenerate, every centroid equal
    cam = sol_amd.Scene.cornell_box().mesh_view().camera
    light = F.Light()
    light.type = F.LIGHT_POINT
    light.v0[:] = [0.5, 2.0, 0.5]
    light.intensity[:] = [1, 1, 1]
    light.group = -1
    mat = F.Material()
    mat.type = F.BSDF_DIFFUSE
    mat.kd[:] = [0.5, 0.5, 0.5]
    mat.group = -1
    scene = sol_amd.Scene.from_arrays(pos, np.tile([0, 0, 1], (3 * n, 1)), np.arange(3 * n).reshape(n, 3),
                                      np.zeros(n, np.int32), -np.ones(n, np.int32), [mat], [light], cam)
    built = ctx.build(scene)
    check_structure(scene, built)


def oracle_of_built(scene, built):
    """The oracle over the GPU-built tree itself (not the host's, not its own)."""
    d = built.download()
    return O.OracleScene.with_gpu_nodes(scene, d["nodes"], [t.prim for t in d["tris"]])


@pytest.mark.parametrize("name", ["cornell", "spheres", "water"])
def test_gpu_built_scene_renders_like_the_oracle_on_its_tree(ctx, name):
    # The oracle traverses the GPU-built tree (not the host's, not its own): renders of both GPU
    # renderers over the GPU-built tree equal the oracle's bit for bit.
    scene = scenes()[name]()
    built = ctx.build(scene)
    oracle = oracle_of_built(scene, built)
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    ref = sol_amd.new_image(48, 32)
    oracle.render(ref, 2, 4, cfg)
    assert ref.any()
    for kind in (F.RENDERER_MEGAKERNEL, F.RENDERER_WAVEFRONT):
        for width in (2, 4):
            ctx.set_tuning(bvh_width=width)
            print(f"render {name} renderer {kind} bvh width {width} vs oracle", flush=True)
            img = sol_amd.new_image(48, 32)
            built.render(img, 2, 4, cfg, renderer=kind)
            np.testing.assert_array_equal(img, ref)
    ctx.set_tuning(bvh_width=0)


def test_gpu_built_scene_intersects_like_the_oracle_on_its_tree(ctx):
    scene = sol_amd.Scene.triangle_soup(30000, seed=12)
    built = ctx.build(scene)
    oracle = oracle_of_built(scene, built)
    rays = sol_amd.gen_rays(0, 1 << 15, seed=5)
    shadow = sol_amd.gen_rays(1, 1 << 15, seed=6)
    want = oracle.intersect(rays)
    want_any = oracle.intersect(shadow, any_hit=True)
    assert (want[0] >= 0).mean() > 0.05 and 0.05 < want_any.mean() < 0.95
    for width in (2, 4):
        ctx.set_tuning(bvh_width=width)
        for x, y in zip(built.intersect(rays), want):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(built.intersect(shadow, any_hit=True), want_any)
    ctx.set_tuning(bvh_width=0)


@pytest.mark.parametrize("name", ["cornell", "spheres", "water"])
def test_gpu_built_scene_renders_like_the_host_built_one(ctx, name):
    scene = scenes()[name]()
    host = ctx.upload(scene)
    built = ctx.build(scene)
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    for kind in (F.RENDERER_MEGAKERNEL, F.RENDERER_WAVEFRONT):
        for width in (2, 4):
            ctx.set_tuning(bvh_width=width)
            print(f"render {name} renderer {kind} bvh width {width}", flush=True)
            a, b = sol_amd.new_image(48, 32), sol_amd.new_image(48, 32)
            host.render(a, 0, 4, cfg, renderer=kind)
            built.render(b, 0, 4, cfg, renderer=kind)
            assert a.any()
            np.testing.assert_array_equal(a, b)
    ctx.set_tuning(bvh_width=0)


def test_gpu_built_scene_intersects_like_the_host_built_one(ctx):
    scene = sol_amd.Scene.triangle_soup(50000, seed=6)
    host = ctx.upload(scene)
    built = ctx.build(scene)
    rays = sol_amd.gen_rays(0, 1 << 16, seed=3)
    for width in (2, 4):
        ctx.set_tuning(bvh_width=width)
        for any_hit in (False, True):
            ha = host.intersect(rays, any_hit=any_hit)
            hb = built.intersect(rays, any_hit=any_hit)
            for x, y in zip(ha if isinstance(ha, tuple) else (ha,), hb if isinstance(hb, tuple) else (hb,)):
                np.testing.assert_array_equal(x, y)
    ctx.set_tuning(bvh_width=0)


def test_gpu_build_quality_and_errors(ctx):
    scene = sol_amd.Scene.sphere_field(num_spheres=200, subdivision=3, seed=2)
    built = ctx.build(scene)
    host_sah = scene.build_stats().sah_cost_final
    st = built.build_stats
    print(f"SAH: GPU {st.sah_cost:.2f} host {host_sah:.2f}; GPU build {st.build_ms:.1f} ms", flush=True)
    assert st.sah_cost < 1.5 * host_sah
    with pytest.raises(F.SolError):
        ctx.build(scene, max_leaf_size=9)
    with pytest.raises(F.SolError):
        ctx.build(scene, search_radius=33)


def test_batch_of_gpu_built_scenes_equals_uploaded_ones(ctx):
    import torch
    from sol_amd.workloads import BATCH_DEPTH, batch_specs
    specs = batch_specs(count=6, seed=21, max_spheres=120, sides=(64, 96), spps=(1, 3))
    cfg = sol_amd.PathTracerConfig(max_path_len=BATCH_DEPTH)
    scenes = [sp.scene() for sp in specs]
    out = []
    for make in (ctx.upload, ctx.build):
        gs = [make(sc) for sc in scenes]
        bufs = [torch.zeros((3, sp.height, sp.width), dtype=torch.float32, device="cuda") for sp in specs]
        torch.cuda.synchronize()
        ctx.render_batch([(g, sp.width, sp.height, 0, sp.spp, b.data_ptr()) for sp, g, b in zip(specs, gs, bufs)], cfg)
        out.append([b.cpu().numpy() for b in bufs])
    for a, b in zip(*out):
        assert a.any()
        np.testing.assert_array_equal(a, b)


def test_built_arrays_match_recorded_digests(ctx):
    # The builder is deterministic: its arrays for these scenes equal, bit for bit, those recorded
    # in tests/golden/gpu_build_digests.json (written by tests/golden/make_gpu_build_digests.py
    # when the Morton sort and the prefix sums came from a library; the hand-written radix sort and
    # scans must reproduce them).
    import json
    import os

    from tests.golden.make_gpu_build_digests import SCENES, digest, make_scene
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gpu_build_digests.json")) as f:
        want = json.load(f)
    assert sorted(want) == sorted(SCENES)
    for name, spec in SCENES.items():
        built = ctx.build(make_scene(sol_amd, spec))
        got = digest(built)
        built.close()
        print(f"{name}: {got['num_nodes']} nodes, {got['num_nodes4']} 4-wide nodes", flush=True)
        assert got == want[name], name
