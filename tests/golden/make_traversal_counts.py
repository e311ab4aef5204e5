"""Writes traversal_counts.json: inner-node visits and triangle tests per ray, counted by the
oracle's instrumented traversal of the product BVH on the path tracer's own rays.

These per-ray figures turn measured ray rates into algorithmic bytes for the traversal kernel
(ray 32 B + result + 64 B per inner-node visit + 48 B per triangle test). They are deterministic:
the scene generators are seeded, the BVH builder is deterministic and the rays are the ones the
renderer traces for the chosen (pixel, sample) set.

Workloads:
  sphere_field_1080p  the 400-sphere field (2,048,008 triangles), 1920x1080, max path length 16,
                      every 8th row at sample 0 (all bounces of those paths)
  soup_1m             the 1M-triangle soup, the first 262,144 rays of each intersection batch

Run from the repository root: python tests/golden/make_traversal_counts.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import sol_amd  # noqa: E402
from tests import oracle_lib  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "traversal_counts.json")


def bvh_signature(scene):
    st = scene.build_stats()
    return {"triangles": scene.num_tris, "nodes": st.num_nodes, "leaves": st.num_leaves, "depth": st.depth,
            "sah_cost": round(float(st.sah_cost_final), 4)}


def per_ray(st, prefix):
    rays = getattr(st, f"{prefix}_rays")
    kind = "ext" if prefix == "extension" else "shadow"
    return {"rays": rays, "inner_visits_per_ray": getattr(st, f"{kind}_inner_visits") / rays,
            "tri_tests_per_ray": getattr(st, f"{kind}_tri_tests") / rays}


def sphere_field():
    w, h, row_step = 1920, 1080, 8
    scene = sol_amd.Scene.sphere_field()
    cfg = sol_amd.PathTracerConfig(max_path_len=16)
    osc = oracle_lib.OracleScene(scene, use_product_bvh=True)
    st = oracle_lib.OracleStats()
    img = sol_amd.new_image(w, h)
    for y in range(0, h, row_step):
        osc.render(img, 0, 1, cfg, rows=(y, y + 1), stats=st)
    return {"scene": "sphere field, 400 icospheres, subdivision 4, seed 2", "bvh": bvh_signature(scene),
            "rays": f"{w}x{h}, rows y % {row_step} == 0, sample 0, max_path_len 16",
            "extension": per_ray(st, "extension"), "shadow": per_ray(st, "shadow"),
            "samples": st.samples}


def soup():
    n = 1 << 18
    scene = sol_amd.Scene.triangle_soup(1_000_000)
    osc = oracle_lib.OracleScene(scene, use_product_bvh=True)
    out = {"scene": "triangle soup, 1,000,000 triangles, seed 1", "bvh": bvh_signature(scene),
           "rays": f"first {n} rays of each intersection batch (seed 7)"}
    for kind, name, any_hit in ((0, "extension", False), (1, "shadow", True)):
        st = oracle_lib.OracleStats()
        osc.intersect(sol_amd.gen_rays(kind, n), any_hit=any_hit, stats=st)
        out[name] = per_ray(st, name)
    return out


def main():
    data = {"sphere_field_1080p": sphere_field(), "soup_1m": soup()}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
        f.write("\n")
    print(json.dumps(data, indent=1))


if __name__ == "__main__":
    main()
