"""Host library (libsol_host.so): BVH builder invariants, GPU layout, procedural scenes, and the
oracle's traversal against brute force on the product BVH and on its own BVH (CPU only)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from tests import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _tri_boxes(scene):
    a = scene.arrays()
    v = a["positions"][a["indices"]]             # (T, 3, 3)
    return v.min(1), v.max(1)


def _check_bvh(scene):
    bounds, index, count, prims = scene.bvh_arrays()
    n = scene.num_tris
    lo_t, hi_t = _tri_boxes(scene)
    seen = np.zeros(n, np.int64)
    stack = [0]
    while stack:
        i = stack.pop()
        lo, hi = bounds[i, 0::2], bounds[i, 1::2]
        if count[i]:
            ids = prims[index[i]:index[i] + count[i]]
            seen[ids] += 1
            assert np.all(lo_t[ids] >= lo) and np.all(hi_t[ids] <= hi)
        else:
            for c in (index[i], index[i] + 1):
                clo, chi = bounds[c, 0::2], bounds[c, 1::2]
                assert np.all(clo >= lo) and np.all(chi <= hi)
                stack.append(c)
    assert np.all(seen == 1), "every triangle in exactly one leaf"
    assert count.max() <= 8


@pytest.mark.parametrize("n", [1, 2, 7, 9, 1000, 20000])
def test_bvh_invariants_soup(n):
    s = sol_amd.Scene.triangle_soup(n, seed=n)
    _check_bvh(s)
    st = s.build_stats()
    assert st.sah_cost_final <= st.sah_cost_built * (1 + 1e-6)


def test_reinsertion_never_increases_cost_and_median_is_worse():
    s = sol_amd.Scene.triangle_soup(30000, seed=5, bvh=sol_amd.default_bvh_config(reinsertion_iters=4))
    st = s.build_stats()
    assert st.sah_cost_final <= st.sah_cost_built
    _check_bvh(s)


def test_gpu_layout_covers_every_leaf():
    s = sol_amd.Scene.triangle_soup(5000, seed=3)
    d = s.gpu_desc()
    nodes = np.ctypeslib.as_array(C.cast(d.nodes, C.POINTER(C.c_int32)), shape=(d.num_nodes * 16,)).reshape(-1, 16)
    tris = np.ctypeslib.as_array(C.cast(d.tris, C.POINTER(C.c_int32)), shape=(d.num_tris * 12,)).reshape(-1, 12)
    covered = np.zeros(d.num_tris, np.int64)
    referenced = np.zeros(d.num_nodes, np.int64)
    for c in nodes[:, 12:14].ravel():
        if c >= 0:
            referenced[c] += 1
        else:
            code = ~int(c) & 0xFFFFFFFF
            first, cnt = code >> 3, (code & 7) + 1
            covered[first:first + cnt] += 1
    assert np.all(covered == 1)
    assert referenced[0] == 0 and np.all(referenced[1:] == 1)
    assert sorted(tris[:, 3].tolist()) == list(range(d.num_tris))   # leaf order is a permutation
    assert d.top_nodes >= 1 and d.bvh_depth >= 2


@pytest.mark.parametrize("make", [lambda: sol_amd.Scene.triangle_soup(5000, seed=3),
                                  lambda: sol_amd.Scene.sphere_field(num_spheres=30, subdivision=2, seed=4),
                                  lambda: sol_amd.Scene.cornell_box()])
def test_gpu_layout4_covers_every_leaf_and_bounds_its_subtrees(make):
    s = make()
    d = s.gpu_desc()
    n4 = d.num_nodes4
    raw = np.ctypeslib.as_array(C.cast(d.nodes4, C.POINTER(C.c_float)), shape=(n4 * 32,)).reshape(-1, 32)
    boxes = raw[:, :24].reshape(n4, 6, 4)             # lo_x, hi_x, lo_y, hi_y, lo_z, hi_z x 4 slots
    child = raw[:, 24:28].view(np.int32)
    tris = np.ctypeslib.as_array(C.cast(d.tris, C.POINTER(C.c_float)), shape=(d.num_tris * 12,)).reshape(-1, 12)
    verts = tris[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3, 3)
    covered = np.zeros(d.num_tris, np.int64)
    referenced = np.zeros(n4, np.int64)
    # box of every subtree, bottom-up (children come after parents in the layout)
    sub_lo = np.full((n4, 3), np.inf, np.float32)
    sub_hi = np.full((n4, 3), -np.inf, np.float32)
    filled = []
    for i in range(n4 - 1, -1, -1):
        used = 0
        for k in range(4):
            c = int(child[i, k])
            if c == F.GPU_EMPTY:   # an unused slot holds the inverted box, which every ray misses
                assert np.all(boxes[i, [0, 2, 4], k] == np.inf) and np.all(boxes[i, [1, 3, 5], k] == -np.inf)
                continue
            used += 1
            lo = boxes[i, [0, 2, 4], k]
            hi = boxes[i, [1, 3, 5], k]
            if c >= 0:
                assert c > i
                referenced[c] += 1
                assert np.all(lo <= sub_lo[c]) and np.all(hi >= sub_hi[c])   # child box bounds its subtree
            else:
                code = ~c & 0xFFFFFFFF
                first, cnt = code >> 3, (code & 7) + 1
                covered[first:first + cnt] += 1
                v = verts[first:first + cnt].reshape(-1, 3)
                assert np.all(lo <= v.min(0)) and np.all(hi >= v.max(0))
            sub_lo[i] = np.minimum(sub_lo[i], lo)
            sub_hi[i] = np.maximum(sub_hi[i], hi)
        assert used >= (1 if n4 == 1 else 2)
        filled.append(used)
    assert np.all(covered == 1)
    assert referenced[0] == 0 and np.all(referenced[1:] == 1)
    assert 1 <= d.top_nodes4 <= 85 and 2 <= d.bvh4_depth <= d.bvh_depth
    if d.num_tris > 1000:
        assert np.mean(filled) > 3.0                    # collapsing fills most slots


def test_cornell_box_scene():
    s = sol_amd.Scene.cornell_box()
    assert s.num_tris == 36
    lights = s.lights()
    assert len(lights) == 2 and all(l.type == F.LIGHT_TRIANGLE for l in lights)
    assert all(abs(l.normal[1] + 1) < 1e-6 for l in lights)       # ceiling light faces down
    np.testing.assert_allclose([l.intensity[:] for l in lights], [[17, 12, 4]] * 2)
    area = 0.5 * 0.4
    np.testing.assert_allclose(sum(1 / l.inv_area for l in lights), area, rtol=1e-5)
    mats = s.materials()
    kd = sorted(tuple(np.round(m.kd[:], 3)) for m in mats)
    assert (0.65, 0.05, 0.05) in kd and (0.12, 0.45, 0.15) in kd and (0.73, 0.73, 0.73) in kd
    _check_bvh(s)


@pytest.mark.parametrize("level", [0, 1, 2])
def test_sphere_field_icospheres_face_outwards(level):
    s = sol_amd.Scene.sphere_field(num_spheres=5, subdivision=level, seed=9)
    a = s.arrays()
    assert s.num_tris == 2 + 5 * 20 * 4 ** level + 6
    v = a["positions"][a["indices"]]
    sph = a["tri_material"] > 0
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    # every sphere triangle's normal agrees with its interpolated vertex normals (outwards)
    vn = a["normals"][a["indices"]].sum(1)
    assert np.all(np.einsum("ij,ij->i", n[sph], vn[sph]) > 0)
    lights = s.lights()
    assert len(lights) == 7 and sum(l.type == F.LIGHT_POINT for l in lights) == 1


def test_sphere_field_triangle_count_at_full_size_level():
    p = sol_amd.default_sphere_field_params()
    assert p.num_spheres == 400 and p.subdivision == 4
    assert 400 * 20 * 4 ** 4 == 2_048_000   # ~2M triangles


def test_rays_of_intersection_benchmark():
    r = sol_amd.gen_rays(0, 10000, seed=4)
    assert np.all(r[:, :3] >= -0.5) and np.all(r[:, :3] <= 1.5)
    np.testing.assert_allclose(np.linalg.norm(r[:, 4:7], axis=1), 1, atol=1e-5)
    assert np.all(r[:, 3] == 0) and np.all(np.isinf(r[:, 7]))
    s = sol_amd.gen_rays(1, 10000, seed=4, offset=1e-5)
    np.testing.assert_array_equal(s[:, :3], r[:, :3])               # same origins
    tgt = s[:, :3] + s[:, 4:7]
    assert np.all(tgt >= -1e-6) and np.all(tgt <= 1 + 1e-6)
    assert np.all(s[:, 3] == np.float32(1e-5)) and np.all(s[:, 7] == np.float32(1 - 1e-5))


@pytest.mark.parametrize("use_product_bvh", [False, True])
def test_oracle_traversal_equals_brute_force(use_product_bvh):
    scene = sol_amd.Scene.triangle_soup(30000, seed=11)
    o = O.OracleScene(scene, use_product_bvh=use_product_bvh)
    rays = sol_amd.gen_rays(0, 20000, seed=12)
    bvh = o.intersect(rays)
    brute = o.intersect(rays, brute_force=True)
    for a, b in zip(bvh, brute):
        np.testing.assert_array_equal(a, b)
    assert (bvh[0] >= 0).mean() > 0.05
    shadow = sol_amd.gen_rays(1, 20000, seed=12)
    np.testing.assert_array_equal(o.intersect(shadow, any_hit=True), o.intersect(shadow, any_hit=True, brute_force=True))


def test_oracle_tie_break_prefers_smaller_index():
    # Two coincident triangles: the closest hit must be the one with the smaller index,
    # whatever the tree.
    pos = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0]])
    nrm = np.float32([[0, 0, 1]] * 6)
    idx = np.uint32([[3, 4, 5], [0, 1, 2]])
    m = F.Material(); m.type = F.BSDF_DIFFUSE; m.kd[:] = (0.5, 0.5, 0.5); m.group = -1
    l = F.Light(); l.type = F.LIGHT_POINT; l.v0[:] = (0, 0, 5); l.intensity[:] = (1, 1, 1); l.group = -1
    cam = F.CameraParams(); cam.eye[:] = (0.2, 0.2, 2); cam.dir[:] = (0, 0, -1); cam.up[:] = (0, 1, 0)
    cam.fov, cam.aspect = 60, 1
    s = sol_amd.Scene.from_arrays(pos, nrm, idx, [0, 0], [-1, -1], [m], [l], cam)
    rays = np.float32([[0.2, 0.2, 1, 0, 0, 0, -1, np.inf]])
    for product in (False, True):
        prim, t, _, _ = O.OracleScene(s, use_product_bvh=product).intersect(rays)
        assert prim[0] == 0 and t[0] == 1.0


def _declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\*?((?:sol|solhip)_[a-z_0-9]+)\(", text, re.M)))


def test_libraries_export_every_declared_symbol():
    for header, lib in (("solhip.h", "libsolhip.so"), ("sol_host.h", "libsol_host.so")):
        names = _declared_functions(header)
        assert len(names) >= 8
        out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "sol_amd", lib)], capture_output=True,
                             text=True, check=True).stdout
        exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
        missing = [n for n in names if n not in exported]
        assert not missing, f"{lib} lacks {missing}"
        assert set(names) <= set(F.exported_symbols()[lib])


def test_invalid_arguments_are_rejected():
    with pytest.raises(F.SolError):
        sol_amd.Scene.triangle_soup(0)
    with pytest.raises(F.SolError):
        sol_amd.Scene.cornell_box(aspect=-1.0)
    with pytest.raises(F.SolError):
        sol_amd.Scene.sphere_field(num_spheres=0)


def test_oracle_over_the_gpu_node_layout_equals_the_product_bvh():
    # The rewrite of a binary GPU-layout tree into the oracle's bvh v2 layout (used to run the
    # oracle over GPU-built trees) checked on the host's own flattening of the same tree: the
    # oracle then renders and intersects exactly as over the product BVH.
    s = sol_amd.Scene.sphere_field(num_spheres=20, subdivision=2, seed=4, extent=3.0, radius_max=1.0)
    d = s.gpu_desc()
    nodes = [d.nodes[i] for i in range(d.num_nodes)]
    prims = [d.tris[i].prim for i in range(d.num_tris)]
    conv = O.OracleScene.with_gpu_nodes(s, nodes, prims)
    prod = O.OracleScene(s, use_product_bvh=True)
    rays = sol_amd.gen_rays(0, 4096, seed=9)
    for a, b in zip(conv.intersect(rays), prod.intersect(rays)):
        np.testing.assert_array_equal(a, b)
    st_conv, st_prod = O.OracleStats(), O.OracleStats()
    cfg = sol_amd.PathTracerConfig(max_path_len=6)
    a, b = sol_amd.new_image(24, 16), sol_amd.new_image(24, 16)
    conv.render(a, 0, 2, cfg, stats=st_conv)
    prod.render(b, 0, 2, cfg, stats=st_prod)
    assert a.any()
    np.testing.assert_array_equal(a, b)
    # the same tree: every ray visits the same nodes
    assert st_conv.ext_inner_visits == st_prod.ext_inner_visits and st_conv.ext_tri_tests == st_prod.ext_tri_tests
