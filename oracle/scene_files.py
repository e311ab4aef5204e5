"""Oracle of the scene-file loaders. TEST INFRASTRUCTURE ONLY: the checker that tests/ use; the
product never imports it.

A plain-Python restatement of sol's OBJ / MTL import and TOML scene loading, written from the
reference sources (paths relative to the sol tree), producing the flat tables the renderers read:
positions, normals, triangle indices, per-triangle material / light, the material table in the
order sol's loader creates the BSDFs, the light table in creation order, and the camera.

    function            follows                              pinned by (tests/test_scene_files.py)
    parse_obj           src/formats/obj.cpp:159-269          hand-written OBJ cases (negative and
                                                             v/t/n indices, fans, invalid faces,
                                                             usemtl / g / o / s), fixture counts
    parse_mtl           src/formats/obj.cpp:281-355          hand-written MTL cases, fixture values
    convert_material    src/formats/obj.cpp:374-411          illum 5 / 7 / default mix weights by hand
    build_mesh          src/formats/obj.cpp:424-510          vertex sharing, missing normals of a planar
                                                             face = its unit normal, light per triangle
    load_scene          src/scene_loader.cpp:44-102          fixture cameras, defaults, error cases
    camera_params       src/scene_loader.cpp:80-91           (the fixture tables)

Arithmetic: fp32 throughout (numpy float32 scalars, one rounding per operation, in sol's order),
as the product compiles with -ffp-contract=off. OBJ numbers are read as strtof reads them (the
longest numeric prefix); the value goes through a Python float (double) before float32, which
matches strtof for all but decimal strings within 2^-53 of a float32 rounding midpoint (none in
the fixtures). TOML numbers are doubles converted to float32, as toml++'s value_or<float> does.
Parsed with tomli (TOML 1.0), independent of the product's reader.
"""
import os
import re

import numpy as np
import tomli

F32 = np.float32
LINE_BUFFER = 1024          # obj.cpp:26: getline into char[1024]

BSDF_NONE, BSDF_DIFFUSE, BSDF_PHONG, BSDF_GLASS, BSDF_MIRROR, BSDF_INTERP = -1, 0, 1, 2, 3, 4
_FLOAT = re.compile(r"\s*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))", re.I)
_INT = re.compile(r"\s*([+-]?\d+)")


def strtof(s, pos):
    """(value, new position) as C strtof reads s[pos:]: 0 and pos unchanged without a number."""
    m = _FLOAT.match(s, pos)
    if not m:
        return F32(0.0), pos
    return F32(float(m.group(1))), m.end()


def strtol(s, pos):
    m = _INT.match(s, pos)
    if not m:
        return 0, pos
    return int(m.group(1)), m.end()


def _lines(path):
    """Lines as sol reads them: stop at a line that overflows the buffer; skip blank and '#'
    lines; strip leading and trailing white space (obj.cpp:93-107, 172-182)."""
    with open(path, "rb") as f:
        data = f.read().decode("latin-1")
    for raw in data.split("\n"):
        if len(raw) > LINE_BUFFER - 1:
            return
        line = raw.lstrip(" \t\r\f\v")
        if not line or line[0] == "#":
            continue
        yield line.rstrip(" \t\r\n\f\v")


def _vec(s, pos, n):
    out = []
    for _ in range(n):
        v, pos = strtof(s, pos)
        out.append(v)
    return out


def _space(c):
    return c in " \t\r\n\f\v"


def _at(s, i):
    return s[i] if i < len(s) else "\0"


def parse_obj(path):
    """parse_obj (obj.cpp:159-269): vertices / normals / texture coordinates with the zero dummy at
    index 0, valid faces in file order with their material index, usemtl names, mtllib names."""
    verts, norms, texs = [[F32(0)] * 3], [[F32(0)] * 3], [[F32(0)] * 2]
    faces, materials, mtl_files = [], ["#dummy"], []
    cur = 0
    for s in _lines(path):
        c1 = _at(s, 1)
        if s[0] == "v":
            if c1 in " \t":
                verts.append(_vec(s, 1, 3))
            elif c1 == "n":
                norms.append(_vec(s, 2, 3))
            elif c1 == "t":
                texs.append(_vec(s, 2, 2))
        elif s[0] == "f" and _space(c1):
            idx, pos = [], 2
            while True:
                # parse_index (obj.cpp:111-138)
                p = pos
                while p < len(s) and _space(s[p]):
                    p += 1
                if p >= len(s) or not (s[p].isdigit() or s[p] == "-"):
                    break
                v, q = strtol(s, p)
                if q == p:
                    break
                t = n = 0
                while q < len(s) and _space(s[q]):
                    q += 1
                if _at(s, q) == "/":
                    q += 1
                    if _at(s, q) != "/":
                        t, q = strtol(s, q)
                    while q < len(s) and _space(s[q]):
                        q += 1
                    if _at(s, q) == "/":
                        q += 1
                        n, q = strtol(s, q)
                idx.append([v, t, n])
                pos = q
            ok = len(idx) >= 3
            for i in idx:
                i[0] = len(verts) + i[0] if i[0] < 0 else i[0]
                i[1] = len(texs) + i[1] if i[1] < 0 else i[1]
                i[2] = len(norms) + i[2] if i[2] < 0 else i[2]
                ok = ok and 0 < i[0] < len(verts) and 0 <= i[1] < len(texs) and 0 <= i[2] < len(norms)
            if ok:
                faces.append(([tuple(i) for i in idx], cur))
        elif s.startswith("usemtl") and _space(_at(s, 6)):
            name = s[6:].lstrip().split()[0]
            if name not in materials:
                materials.append(name)
            cur = materials.index(name)
        elif s.startswith("mtllib") and _space(_at(s, 6)):
            name = s[6:].lstrip().split()[0]
            if name not in mtl_files:
                mtl_files.append(name)
    return {"v": verts, "vn": norms, "vt": texs, "faces": faces, "materials": materials, "mtl_files": mtl_files}


def _material():
    z = [F32(0)] * 3
    return {"Ka": z, "Kd": z, "Ks": z, "Ke": z, "Tf": z, "Ns": F32(0), "Ni": F32(0), "Tr": F32(0), "d": F32(0),
            "illum": 0, "map_Ka": "", "map_Kd": "", "map_Ks": "", "map_Ns": "", "map_Ke": "", "map_d": "", "bump": ""}


def _color(s, pos):
    return [max(v, F32(0.0)) for v in _vec(s, pos, 3)]


def parse_mtl(path, lib):
    """parse_mtl (obj.cpp:281-355); a missing file is skipped (non-strict mode)."""
    if not os.path.exists(path):
        return
    cur = lib.setdefault("#dummy", _material())
    for s in _lines(path):
        if s.startswith("newmtl") and _space(_at(s, 6)):
            name = s[7:].lstrip().split()[0] if s[7:].strip() else ""
            cur = lib.setdefault(name, _material())
        elif len(s) >= 3 and s[0] == "K" and _space(s[2]) and s[1] in "adse":
            cur["K" + s[1]] = _color(s, 3)
        elif len(s) >= 3 and s[0] == "N" and _space(s[2]) and s[1] in "si":
            cur["N" + s[1]] = strtof(s, 3)[0]
        elif len(s) >= 3 and s[0] == "T" and _space(s[2]) and s[1] in "fr":
            cur["T" + s[1]] = _color(s, 3) if s[1] == "f" else strtof(s, 3)[0]
        elif s[0] == "d" and _space(_at(s, 1)):
            cur["d"] = strtof(s, 2)[0]
        elif s.startswith("illum") and _space(_at(s, 5)):
            cur["illum"] = int(float(strtof(s, 6)[0]))
        else:
            for key, n in (("map_Ka", 6), ("map_Kd", 6), ("map_Ks", 6), ("map_Ke", 6), ("map_Ns", 6), ("map_d", 5)):
                if s.startswith(key) and _space(_at(s, n)):
                    cur[key] = s[n + 1:].lstrip()
                    break
            else:
                if s.startswith("bump") and _space(_at(s, 4)):
                    cur["bump"] = s[4:].lstrip()


def luminance(c):
    """RgbColor::luminance (color.h:58)."""
    return c[0] * F32(0.2126) + c[1] * F32(0.7152) + c[2] * F32(0.0722)


class _Tables:
    """Materials and lights in creation order, shared when equal (scene_loader.h:57-105)."""

    def __init__(self):
        self.materials, self.lights, self._bsdf, self._light = [], [], {}, {}

    def bsdf(self, key, record):
        if key not in self._bsdf:
            self._bsdf[key] = len(self.materials)
            self.materials.append(record)
        return self._bsdf[key]

    def light(self, tri, ke):
        key = tuple(float(x) for v in tri for x in v) + tuple(float(x) for x in ke)
        if key not in self._light:
            self._light[key] = len(self.lights)
            self.lights.append(triangle_light(tri, ke))
        return self._light[key]


def _black(c):
    return all(float(x) == 0.0 for x in c)


def convert_material(tables, m):
    """convert_material (obj.cpp:374-411): index into tables.materials, or -1 (no BSDF).
    Texture maps count as present (weight 1) but only constant colours are used."""
    c3 = lambda c: tuple(float(x) for x in c)
    if m["illum"] == 5:
        return tables.bsdf(("mirror", c3(m["Ks"])), {"type": BSDF_MIRROR, "ks": m["Ks"]})
    if m["illum"] == 7:
        eta = F32(1.0) / m["Ni"] if m["Ni"] != 0 else F32(np.inf)
        return tables.bsdf(("glass", c3(m["Ks"]), c3(m["Tf"]), float(eta)),
                           {"type": BSDF_GLASS, "ks": m["Ks"], "kt": m["Tf"], "eta": eta})
    spec = diff = None
    spec_k = diff_k = F32(0)
    if not _black(m["Ks"]) or m["map_Ks"]:
        spec = tables.bsdf(("phong", c3(m["Ks"]), float(m["Ns"])), {"type": BSDF_PHONG, "ks": m["Ks"], "ns": m["Ns"]})
        spec_k = F32(1.0) if m["map_Ks"] else luminance(m["Ks"])
    if not _black(m["Kd"]) or m["map_Kd"]:
        diff = tables.bsdf(("diffuse", c3(m["Kd"])), {"type": BSDF_DIFFUSE, "kd": m["Kd"]})
        diff_k = F32(1.0) if m["map_Kd"] else luminance(m["Kd"])
    if spec is not None and diff is not None:
        k = spec_k / (diff_k + spec_k)
        rec = {"type": BSDF_INTERP, "kd": m["Kd"], "ks": m["Ks"], "ns": m["Ns"], "mix": k}
        return tables.bsdf(("interp", diff, spec, float(k)), rec)
    if diff is not None:
        return diff
    return spec if spec is not None else -1


def sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def length(a):
    return np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def normalize(a):
    s = F32(1.0) / length(a)
    return [a[0] * s, a[1] * s, a[2] * s]


def unit_normal(tri):
    """Trianglef::normal: normalize(cross(v1 - v0, v2 - v0)) (shapes.h:55, the light's normal)."""
    return normalize(cross(sub(tri[1], tri[0]), sub(tri[2], tri[0])))


def triangle_light(tri, ke):
    """UniformTriangleLight (lights.cpp:59-111, shapes.h:49-60): unit normal, 1 / area."""
    raw = cross(sub(tri[1], tri[0]), sub(tri[2], tri[0]))
    return {"v": tri, "normal": normalize(raw), "inv_area": F32(1.0) / (F32(0.5) * length(raw)), "intensity": ke}


def build_mesh(obj, lib, tables):
    """build_mesh (obj.cpp:424-510): flat mesh arrays; tables receive materials and lights."""
    mats = [lib[name] if name in lib else lib["#dummy"] for name in obj["materials"]]   # check_materials
    ids, verts, norms, fix = {}, [], [], []
    indices, tri_mat, tri_light = [], [], []
    for face, mi in obj["faces"]:
        vid = []
        for (v, t, n) in face:
            key = (v, n, t)
            if key not in ids:
                ids[key] = len(verts)
                if n == 0:
                    fix.append(len(verts))
                verts.append(obj["v"][v])
                norms.append(obj["vn"][n])
            vid.append(ids[key])
        m = mats[mi]
        bsdf = convert_material(tables, m)
        emissive = not _black(m["Ke"]) or m["map_Ke"]
        for k in range(1, len(vid) - 1):
            tri = (verts[vid[0]], verts[vid[k]], verts[vid[k + 1]])
            tri_light.append(tables.light(tri, m["Ke"]) if emissive else -1)
            indices.append((vid[0], vid[k], vid[k + 1]))
            tri_mat.append(bsdf)
    if fix:
        zero = [F32(0)] * 3
        smooth = [list(zero) for _ in verts]
        for (a, b, c) in indices:
            n = unit_normal((verts[a], verts[b], verts[c]))
            for i in (a, b, c):
                smooth[i] = [smooth[i][j] + n[j] for j in range(3)]
        for i in fix:
            norms[i] = normalize(smooth[i])
    return {"positions": np.array(verts, np.float32).reshape(-1, 3),
            "normals": np.array(norms, np.float32).reshape(-1, 3),
            "indices": np.array(indices, np.uint32).reshape(-1, 3),
            "tri_material": np.array(tri_mat, np.int32), "tri_light": np.array(tri_light, np.int32)}


def load_obj(path, tables=None):
    """obj::load (obj.cpp:512-531)."""
    tables = tables if tables is not None else _Tables()
    obj = parse_obj(path)
    lib = {"#dummy": _material()}
    base = os.path.dirname(os.path.abspath(path))
    for name in obj["mtl_files"]:
        parse_mtl(base + "/" + name, lib)
    mesh = build_mesh(obj, lib, tables)
    mesh["materials"], mesh["lights"] = tables.materials, tables.lights
    return mesh


DEFAULTS = {"fov": 60.0, "aspect": 1.0, "eye": (0.0, 0.0, 0.0), "dir": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0)}


def _vec3_or(node, d):
    """parse_vec3 (scene_loader.cpp:15-23)."""
    if not isinstance(node, list):
        return [F32(x) for x in d]
    out = []
    for i in range(3):
        e = node[i] if i < len(node) else None
        ok = isinstance(e, (int, float)) and not isinstance(e, bool)
        out.append(F32(e) if ok else F32(d[i]))
    return out


def _num_or(node, d):
    ok = isinstance(node, (int, float)) and not isinstance(node, bool)
    return F32(node) if ok else F32(d)


def load_scene(path, aspect=None, **defaults):
    """Scene::load (scene_loader.cpp:44-64, 80-102): the root object's mesh and tables, and the
    camera parameters (eye, dir, up, fov, aspect). Raises ValueError like sol's errors."""
    d = dict(DEFAULTS)
    d.update(defaults)
    if aspect is not None:
        d["aspect"] = aspect
    with open(path, "rb") as f:
        doc = tomli.load(f)
    base = os.path.dirname(os.path.abspath(path))
    camera = None
    cam = doc.get("camera")
    if isinstance(cam, dict):
        if cam.get("type") != "perspective":
            raise ValueError(f"Unknown camera type '{cam.get('type', '')}'")
        camera = {"eye": _vec3_or(cam.get("eye"), d["eye"]), "dir": _vec3_or(cam.get("dir"), d["dir"]),
                  "up": _vec3_or(cam.get("up"), d["up"]), "fov": _num_or(cam.get("fov"), d["fov"]),
                  "aspect": _num_or(cam.get("aspect"), d["aspect"])}
    tables = _Tables()
    geoms = {}
    objects = doc.get("objects")
    for obj in objects if isinstance(objects, list) else []:
        if not isinstance(obj, dict):
            continue
        name, kind = obj.get("name", ""), obj.get("type", "")
        if kind != "import":
            raise ValueError(f"Unknown node type '{kind}'")
        file = obj.get("file", "")
        if not file.endswith(".obj"):
            raise ValueError(f"Unknown file format for '{file}'")
        if name in geoms:
            raise ValueError(f"Duplicate geometry found with name '{name}'")
        geoms[name] = load_obj(base + "/" + file, tables)
    root = doc.get("root", "")
    if root not in geoms:
        raise ValueError(f"Root geometry named '{root}' cannot be found")
    mesh = geoms[root]
    mesh["camera"] = camera
    return mesh
