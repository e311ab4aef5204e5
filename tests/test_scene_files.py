"""Scene files: OBJ / MTL import and TOML scene loading (CPU).

The oracle loader (oracle/scene_files.py) is pinned first, on hand-written OBJ / MTL / TOML cases
with expected values worked out by hand from sol's semantics (src/formats/obj.cpp,
src/scene_loader.cpp), and on counts read off the raw fixture files. The product loader
(libsol_host.so, sol_scene_load / sol_scene_load_obj) must then produce the oracle's flat tables bit
for bit on the same cases and on sol's own test scenes (test/data/cornell_box*.{toml,obj,mtl},
copied into tests/golden/cornell with their copyright notice).
"""
import os

import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from oracle import scene_files as SF
from tests import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "cornell")
NAMES = ["cornell_box", "cornell_box_glossy", "cornell_box_specular", "cornell_box_water"]
CAM = dict(eye=(0, 0.9, 2.5), dir=(0, 0, -1), up=(0, 1, 0), fov=60.0, aspect=1.0)

# A light face every hand-written case carries (the product needs one light to flatten a scene).
LIGHT_MTL = "newmtl lamp\nKd 0 0 0\nKe 5 5 5\n"
LIGHT_OBJ = "v 0 5 0\nv 1 5 0\nv 0 5 1\nusemtl lamp\nf -3 -2 -1\n"


def write_case(tmp_path, obj, mtl="", name="case"):
    (tmp_path / f"{name}.mtl").write_text(mtl + LIGHT_MTL)
    (tmp_path / f"{name}.obj").write_text(f"mtllib {name}.mtl\n" + obj + LIGHT_OBJ)
    return str(tmp_path / f"{name}.obj")


def product_tables(scene):
    a = scene.arrays()
    return {"positions": a["positions"].reshape(-1, 3), "normals": a["normals"].reshape(-1, 3),
            "indices": a["indices"].reshape(-1, 3), "tri_material": a["tri_material"], "tri_light": a["tri_light"]}


def assert_same_tables(scene, want):
    got = product_tables(scene)
    for k in ("positions", "normals", "indices", "tri_material", "tri_light"):
        g, w = np.ascontiguousarray(got[k]), np.ascontiguousarray(want[k]).reshape(got[k].shape)
        assert g.dtype.itemsize == w.dtype.itemsize and g.tobytes() == w.tobytes(), k
    assert [bytes(m) for m in scene.materials()] == [bytes(O.material_struct(r)) for r in want["materials"]]
    assert [bytes(lt) for lt in scene.lights()] == [bytes(O.light_struct(r)) for r in want["lights"]]


# ---- the oracle loader against hand-worked cases ------------------------------------------------

def test_quad_with_negative_indices_fans_and_gets_its_face_normal(tmp_path):
    path = write_case(tmp_path, "v 0 0 0\nv 2 0 0\nv 2 1 0\nv 0 1 0\nf -4 -3 -2 -1\n", "newmtl white\nKd 0.5 0.5 0.5\n")
    m = SF.load_obj(path)
    # one quad -> two triangles (0 1 2) (0 2 3), then the light triangle
    assert m["indices"].tolist() == [[0, 1, 2], [0, 2, 3], [4, 5, 6]]
    np.testing.assert_array_equal(m["positions"][:4], np.float32([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]]))
    # no vn: each vertex gets the normalised sum of unit face normals, here +z
    np.testing.assert_array_equal(m["normals"][:4], np.float32([[0, 0, 1]] * 4))
    # no usemtl before the quad: the dummy material, no BSDF
    assert m["tri_material"].tolist()[:2] == [-1, -1]
    assert m["tri_light"].tolist() == [-1, -1, 0]


def test_vertex_texture_normal_triples_are_shared_only_when_all_three_match(tmp_path):
    obj = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvt 1 1\nvn 0 0 2\nvn 0 0 -1\n"
           "usemtl white\nf 1/1/1 2/1/1 3/1/1\nf 2/1/1 4/2/1 3/1/1\nf 1//2 2//2 3//2\n")
    m = SF.load_obj(write_case(tmp_path, obj, "newmtl white\nKd 0.5 0.5 0.5\n"))
    # (v, n, t) triples in order of first use: (1,1,1) (2,1,1) (3,1,1) (4,1,2) (1,2,0) (2,2,0) (3,2,0)
    assert m["indices"].tolist()[:3] == [[0, 1, 2], [1, 3, 2], [4, 5, 6]]
    assert len(m["positions"]) == 7 + 3
    # normals from the file are kept as written (not normalised)
    np.testing.assert_array_equal(m["normals"][:4], np.float32([[0, 0, 2]] * 4))
    np.testing.assert_array_equal(m["normals"][4:7], np.float32([[0, 0, -1]] * 3))


def test_invalid_faces_are_skipped(tmp_path):
    obj = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl white\n"
           "f 1 2\n"            # two indices
           "f 0 1 2\n"          # vertex index 0
           "f 1 2 9\n"          # beyond the vertices declared so far
           "f -4 -3 -1\n"       # relative index before the first vertex
           "f 1 2 3\n")         # the only valid face
    m = SF.load_obj(write_case(tmp_path, obj, "newmtl white\nKd 0.5 0.5 0.5\n"))
    assert m["indices"].tolist() == [[0, 1, 2], [3, 4, 5]]


def test_material_mapping_by_illum_and_colours(tmp_path):
    mtl = ("newmtl mirror\nillum 5\nKs 0.9 0.8 0.7\nKd 1 1 1\n"
           "newmtl glass\nillum 7\nKs 1 1 1\nTf 0.5 0.6 0.7\nNi 1.5\n"
           "newmtl diffuse\nKd 0 1 0\n"
           "newmtl phong\nKs 1 1 1\nNs 50\n"
           "newmtl mix\nKd 0 1 0\nKs 1 1 1\nNs 50\n"
           "newmtl black\nKd 0 0 0\n")
    obj = "v 0 0 0\nv 1 0 0\nv 0 1 0\n" + "".join(f"usemtl {n}\nf 1 2 3\n" for n in
                                                   ("mirror", "glass", "diffuse", "phong", "mix", "black", "nosuch"))
    m = SF.load_obj(write_case(tmp_path, obj, mtl))
    mats = m["materials"]
    kinds = [mats[i]["type"] if i >= 0 else None for i in m["tri_material"][:7]]
    assert kinds == [F.BSDF_MIRROR, F.BSDF_GLASS, F.BSDF_DIFFUSE, F.BSDF_PHONG, F.BSDF_INTERP, None, None]
    glass = mats[m["tri_material"][1]]
    assert glass["eta"] == np.float32(1.0) / np.float32(1.5)
    mix = mats[m["tri_material"][4]]
    # lum(Ks) / (lum(Kd) + lum(Ks)) = 1 / (0.7152 + 1)
    assert abs(float(mix["mix"]) - 1.0 / 1.7152) < 1e-6
    # the mix reuses the diffuse and Phong BSDFs made earlier: 5 BSDFs, creation order kept
    assert [r["type"] for r in mats] == [F.BSDF_MIRROR, F.BSDF_GLASS, F.BSDF_DIFFUSE, F.BSDF_PHONG, F.BSDF_INTERP]


def test_mtl_quirks(tmp_path):
    mtl = ("Kd 0.1 0.1 0.1\n"                    # before any newmtl: edits the dummy material
           "newmtl a\nKd 0.3 0.3 0.3\nKs -1 0 2\nNs 7\n"
           "newmtl a\nKs 0 0 0\n")              # a repeated newmtl edits the same entry
    obj = "v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl nosuch\nf 1 2 3\nusemtl a\nf 1 2 3\n"
    m = SF.load_obj(write_case(tmp_path, obj, mtl))
    mats = m["materials"]
    dummy = mats[m["tri_material"][0]]
    assert dummy["type"] == F.BSDF_DIFFUSE and float(dummy["kd"][0]) == np.float32(0.1)
    a = mats[m["tri_material"][1]]
    assert a["type"] == F.BSDF_DIFFUSE and float(a["kd"][0]) == np.float32(0.3)


def test_emissive_faces_give_one_light_per_triangle(tmp_path):
    mtl = "newmtl glow\nKd 0.5 0.5 0.5\nKe 1 2 3\n"
    obj = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nusemtl glow\nf 1 2 3 4\n"
    m = SF.load_obj(write_case(tmp_path, obj, mtl))
    assert m["tri_light"].tolist() == [0, 1, 2]
    for lt in m["lights"][:2]:
        np.testing.assert_array_equal(np.float32(lt["normal"]), np.float32([0, 0, 1]))
        assert lt["inv_area"] == np.float32(2.0)           # each half of the unit square
        np.testing.assert_array_equal(np.float32(lt["intensity"]), np.float32([1, 2, 3]))
    # the light's BSDF stays: emissive faces are diffuse too
    assert m["materials"][m["tri_material"][0]]["type"] == F.BSDF_DIFFUSE


def test_overlong_line_ends_the_file(tmp_path):
    obj = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl white\nf 1 2 3\n# " + "x" * 1100 + "\nf 1 2 3\n")
    path = tmp_path / "long.obj"
    path.write_text("mtllib long.mtl\n" + obj)
    (tmp_path / "long.mtl").write_text("newmtl white\nKd 1 1 1\nKe 1 1 1\n")
    m = SF.load_obj(str(path))
    assert len(m["indices"]) == 1
    s = sol_amd.Scene.load_obj(str(path), **CAM)
    assert s.num_tris == 1


def test_fixture_counts_read_off_the_files():
    """Triangle and light counts, emission and light areas from the raw OBJ / MTL text."""
    for name in NAMES:
        emit, cur = {}, None
        for ln in open(os.path.join(FIXTURES, name + ".mtl")).read().splitlines():
            tok = ln.split()
            if tok[:1] == ["newmtl"]:
                cur = tok[1]
            elif tok[:1] == ["Ke"] and any(float(x) != 0 for x in tok[1:4]):
                emit[cur] = [float(x) for x in tok[1:4]]
        lines = open(os.path.join(FIXTURES, name + ".obj")).read().splitlines()
        want_tris = want_lights = 0
        cur, used = None, set()
        for ln in lines:
            if ln.startswith("usemtl"):
                cur = ln.split()[1]
            elif ln.startswith("f "):
                want_tris += len(ln.split()) - 3
                if cur in emit:
                    want_lights += len(ln.split()) - 3
                    used.add(cur)
        m = SF.load_scene(os.path.join(FIXTURES, name + ".toml"))
        assert len(m["indices"]) == want_tris, name
        assert (m["tri_light"] >= 0).sum() == len(m["lights"]) == want_lights == 2, name
        (ke,) = [emit[n] for n in used]
        for lt in m["lights"]:
            np.testing.assert_array_equal(np.float32(lt["intensity"]), np.float32(ke))
            # every fixture's light is a ceiling quad facing down into the box
            np.testing.assert_allclose(np.float32(lt["normal"]), [0, -1, 0], atol=1e-6)
            v = np.array(lt["v"], np.float64)
            area = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
            assert abs(float(lt["inv_area"]) * area - 1) < 1e-5


def test_fixture_cameras_and_glossy_mix_weights():
    m = SF.load_scene(os.path.join(FIXTURES, "cornell_box_glossy.toml"), aspect=4 / 3)
    cam = m["camera"]
    assert [float(x) for x in cam["eye"]] == [0.0, np.float32(0.9), 2.5]
    assert [float(x) for x in cam["dir"]] == [0, 0, -1] and float(cam["fov"]) == 60.0
    assert cam["aspect"] == np.float32(4 / 3)      # the default: the file has no aspect
    mixes = sorted(float(r["mix"]) for r in m["materials"] if r["type"] == F.BSDF_INTERP)
    lum = lambda c: 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    floor = lum((1, 1, 1)) / (lum((0.725, 0.71, 0.68)) + lum((1, 1, 1)))    # floor: Kd white, Ks 1
    assert mixes == pytest.approx([0.5, floor], abs=1e-6)                     # tallBox: Ks = Kd


# ---- the product loader against the oracle ------------------------------------------------------

@pytest.mark.parametrize("case", ["quad", "triples", "invalid", "materials", "quirks", "emissive"])
def test_product_obj_loader_matches_oracle(tmp_path, case):
    cases = {
        "quad": ("v 0 0 0\nv 2 0 0\nv 2 1.5 0.25\nv 0 1 0\nusemtl white\nf -4 -3 -2 -1\n", "newmtl white\nKd 0.5 0.5 0.5\n"),
        "triples": ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvt 1 1\nvn 0 0 2\nvn 0 0 -1\nusemtl white\n"
                    "f 1/1/1 2/1/1 3/1/1\nf 2/1/1 4/2/1 3/1/1\nf 1//2 2//2 3//2\n", "newmtl white\nKd 0.5 0.5 0.5\n"),
        "invalid": ("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl white\nf 1 2\nf 0 1 2\nf 1 2 9\nf -4 -3 -1\nf 1 2 3\n",
                    "newmtl white\nKd 0.5 0.5 0.5\n"),
        "materials": ("v 0 0 0\nv 1 0 0\nv 0 1 0\n" + "".join(f"usemtl {n}\nf 1 2 3\n" for n in
                      ("mirror", "glass", "diffuse", "phong", "mix", "black", "nosuch")),
                      "newmtl mirror\nillum 5\nKs 0.9 0.8 0.7\nnewmtl glass\nillum 7\nKs 1 1 1\nTf 0.5 0.6 0.7\n"
                      "Ni 1.5\nnewmtl diffuse\nKd 0 1 0\nnewmtl phong\nKs 1 1 1\nNs 50\nnewmtl mix\nKd 0 1 0\n"
                      "Ks 1 1 1\nNs 50\nnewmtl black\nKd 0 0 0\n"),
        "quirks": ("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl nosuch\nf 1 2 3\nusemtl a\nf 1 2 3\n",
                   "Kd 0.1 0.1 0.1\nnewmtl a\nKd 0.3 0.3 0.3\nKs -1 0 2\nNs 7\nnewmtl a\nKs 0 0 0\n"),
        "emissive": ("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nusemtl glow\nf 1 2 3 4\n",
                     "newmtl glow\nKd 0.5 0.5 0.5\nKe 1 2 3\n"),
    }
    obj, mtl = cases[case]
    path = write_case(tmp_path, obj, mtl)
    assert_same_tables(sol_amd.Scene.load_obj(path, **CAM), SF.load_obj(path))


@pytest.mark.parametrize("name", NAMES)
def test_product_scene_loader_matches_oracle_on_sol_test_scenes(name):
    path = os.path.join(FIXTURES, name + ".toml")
    scene = sol_amd.Scene.load(path, width=320, height=240)
    want = SF.load_scene(path, aspect=320 / 240)
    assert_same_tables(scene, want)
    p = scene.mesh_view().camera
    c = want["camera"]
    assert [list(p.eye), list(p.dir), list(p.up)] == [[float(x) for x in c[k]] for k in ("eye", "dir", "up")]
    assert (p.fov, p.aspect) == (float(c["fov"]), float(c["aspect"]))


def test_scene_loader_accepts_toml_forms(tmp_path):
    """The same Cornell box written with dotted keys, an inline table, literal strings, comments,
    '_' separators and a multi-line array loads to identical tables."""
    for ext in ("obj", "mtl"):
        (tmp_path / f"cornell_box.{ext}").write_bytes(open(os.path.join(FIXTURES, f"cornell_box.{ext}"), "rb").read())
    (tmp_path / "s.toml").write_text(
        "# variant\nroot = 'Box'\ncamera.type = \"perspective\"\n"
        "camera.eye = [\n  0,     # x\n  0.9,\n  2_5e-1_0,\n]\n"
        "camera.dir = [0, 0, -1]\ncamera.up = [0, 1, 0]\ncamera.fov = 6_0\n"
        "objects = [ { name = \"Box\", type = 'import', file = \"cornell_box.obj\" } ]\n")
    a = sol_amd.Scene.load(str(tmp_path / "s.toml"), width=64, height=64)
    want = SF.load_scene(str(tmp_path / "s.toml"), aspect=1.0)
    assert_same_tables(a, want)
    assert list(a.mesh_view().camera.eye) == [0.0, float(np.float32(0.9)), float(np.float32(25e-10))]
    assert list(a.mesh_view().camera.eye) == [float(x) for x in want["camera"]["eye"]]


@pytest.mark.parametrize("text,needle", [
    ("root = 'X'\n[camera]\ntype = 'orthographic'\n", "Unknown camera type 'orthographic'"),
    ("root = 'nothing'\n", "Root geometry named 'nothing' cannot be found"),
    ("root = 'A'\n[[objects]]\nname = 'A'\ntype = 'sphere'\n", "Unknown node type 'sphere'"),
    ("root = 'A'\n[[objects]]\nname = 'A'\ntype = 'import'\nfile = 'x.ply'\n", "Unknown file format for 'x.ply'"),
    ("root = 'A'\n[[objects]]\nname = 'A'\ntype = 'import'\nfile = 'cornell_box.obj'\n"
     "[[objects]]\nname = 'A'\ntype = 'import'\nfile = 'cornell_box.obj'\n", "Duplicate geometry found with name 'A'"),
    ("root = 'A'\n[[objects]]\nname = 'A'\ntype = 'import'\nfile = 'missing.obj'\n", "Cannot open OBJ file"),
    ("root = 'A'\nroot = 'B'\n", ":2:"),                  # duplicate key, with its line
    ("root = 'A'\ncamera = { type = 'perspective', fov = 1979-05-27 }\n", "date-time"),
    ("root = \"A\n", "newline in a single-line string"),
])
def test_scene_loader_errors(tmp_path, text, needle):
    for ext in ("obj", "mtl"):
        (tmp_path / f"cornell_box.{ext}").write_bytes(open(os.path.join(FIXTURES, f"cornell_box.{ext}"), "rb").read())
    (tmp_path / "bad.toml").write_text(text)
    with pytest.raises(F.SolError) as e:
        sol_amd.Scene.load(str(tmp_path / "bad.toml"))
    assert e.value.status == F.SOL_E_LOAD and needle in str(e.value)


def test_missing_scene_file_and_image_maps(tmp_path):
    with pytest.raises(F.SolError) as e:
        sol_amd.Scene.load(str(tmp_path / "none.toml"))
    assert e.value.status == F.SOL_E_LOAD and "Cannot open scene file" in str(e.value)
    # a texture map whose image cannot be opened falls back to the constant colour (as sol does
    # when Image::load fails) but still counts as present: weight 1 for the specular part
    path = write_case(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl m\nf 1 2 3\n",
                      "newmtl m\nKd 0.5 0.5 0.5\nmap_Ks no_such_image.png\n")
    s = sol_amd.Scene.load_obj(path, **CAM)
    want = SF.load_obj(path)
    assert_same_tables(s, want)
    mix = [r for r in want["materials"] if r["type"] == F.BSDF_INTERP][0]
    assert abs(float(mix["mix"]) - 1.0 / (0.5 + 1.0)) < 1e-6
    # an image that exists is refused: only constant textures reach the GPU
    (tmp_path / "tex.png").write_bytes(b"\x89PNG")
    path = write_case(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl m\nf 1 2 3\n",
                      f"newmtl m\nKd 0.5 0.5 0.5\nmap_Kd {tmp_path / 'tex.png'}\n", name="img")
    with pytest.raises(F.SolError) as e:
        sol_amd.Scene.load_obj(path, **CAM)
    assert "constant textures only" in str(e.value)


def test_mixed_material_survives_the_array_round_trip():
    """A scene built from arrays with a SOL_BSDF_INTERP material keeps the caller's indices; the
    two mixed BSDFs are appended to the table."""
    s = sol_amd.Scene.load(os.path.join(FIXTURES, "cornell_box_glossy.toml"))
    a = s.arrays()
    mats = s.materials()
    cp = s.mesh_view().camera
    t = sol_amd.Scene.from_arrays(a["positions"], a["normals"], a["indices"], a["tri_material"], a["tri_light"],
                                  mats, s.lights(), cp)
    got = t.materials()
    n_mix = sum(m.type == F.BSDF_INTERP for m in mats)
    assert len(got) == len(mats) + 2 * n_mix
    assert [bytes(m) for m in got[:len(mats)]] == [bytes(m) for m in mats]
    assert np.array_equal(t.arrays()["tri_material"], a["tri_material"])
