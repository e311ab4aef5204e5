"""Writes tests/golden/gpu_build_digests.json: SHA-256 digests of the arrays solhip_scene_build
makes (binary nodes, 4-wide nodes, leaf-order triangles, shading records) for five small scenes,
with their node counts and depths. Run on a GPU box:
    python tests/golden/make_gpu_build_digests.py [output path]

The committed digests were written by the builder whose Morton sort and prefix sums came from
hipCUB; the builder's own radix sort and scans reproduce them bit for bit.
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

SCENES = {
    "cornell": dict(kind="cornell_box"),
    "spheres": dict(kind="sphere_field", num_spheres=60, subdivision=2, seed=9, extent=4.0, radius_max=1.0),
    "soup": dict(kind="triangle_soup", num_tris=20000, seed=4),
    "soup5": dict(kind="triangle_soup", num_tris=5, seed=4),
    "water": dict(kind="load", path="tests/golden/cornell/cornell_box_water.toml", width=64, height=48),
}


def make_scene(sol_amd, spec):
    s = dict(spec)
    kind = s.pop("kind")
    if kind == "cornell_box":
        return sol_amd.Scene.cornell_box()
    if kind == "sphere_field":
        return sol_amd.Scene.sphere_field(**s)
    if kind == "triangle_soup":
        return sol_amd.Scene.triangle_soup(s["num_tris"], seed=s["seed"])
    return sol_amd.Scene.load(os.path.join(ROOT, s["path"]), s["width"], s["height"])


def digest(built):
    """Digest of the built arrays. The box of an unused 4-wide slot is a marker (the inverted box
    today, zeros when the digests were recorded), so it is hashed as zeros; so is the kind word of
    a triangle record (pad0: the material's kind, kept by the library for the traversal kernel,
    zero when the digests were recorded)."""
    d = built.download()
    h = hashlib.sha256()
    for k in ("nodes", "nodes4", "tris", "shade"):
        for x in d[k]:
            if k == "nodes4":
                x = type(x).from_buffer_copy(bytes(x))
                for c in range(4):
                    if x.child[c] == 0x7FFFFFFF:
                        for f in ("lo_x", "hi_x", "lo_y", "hi_y", "lo_z", "hi_z"):
                            getattr(x, f)[c] = 0.0
            elif k == "tris":
                x = type(x).from_buffer_copy(bytes(x))
                x.pad0 = 0
            h.update(bytes(x))
    i = d["info"]
    return {"sha256": h.hexdigest(), "num_nodes": i.num_nodes, "num_nodes4": i.num_nodes4, "depth": i.bvh_depth,
            "depth4": i.bvh4_depth, "top_nodes": i.top_nodes, "top_nodes4": i.top_nodes4}


def main():
    import sol_amd
    ctx = sol_amd.Context(0)
    out = {}
    for name, spec in SCENES.items():
        b = ctx.build(make_scene(sol_amd, spec))
        out[name] = digest(b)
        b.close()
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(ROOT, "tests", "golden", "gpu_build_digests.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
