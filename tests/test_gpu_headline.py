"""The headline frame at its full size: the procedural sphere field (2,048,008 triangles),
1920x1080, 64 samples per pixel, paths up to 16 segments long.

The wavefront renderer runs with its default tuning (on a 288 GB MI355X every one of the frame's
132,710,400 paths is in flight at once, the configuration the throughput figure is measured in)
and must equal the megakernel bit for bit. Both are then held to the oracle rendering the same
frame over the same BVH on the host's cores (about 11 s on 16 threads): every pixel identical.
"""
import gc
import os

import numpy as np
import pytest

import sol_amd
from tests import oracle_lib as O
from tests.test_gpu_parity import compare_images

pytestmark = pytest.mark.gpu

W, H, SPP, DEPTH = 1920, 1080, 64, 16


def test_headline_frame_matches_megakernel_and_oracle():
    ctx = sol_amd.Context(0)
    scene = sol_amd.Scene.sphere_field()
    assert scene.num_tris == 2_048_008
    cfg = sol_amd.PathTracerConfig(max_path_len=DEPTH)
    images = {}
    for kind in ("wavefront", "megakernel"):
        print(f"render {kind} {W}x{H} {SPP} spp, path slots {ctx.tuning().path_slots}", flush=True)
        img = sol_amd.new_image(W, H)
        sol_amd.GpuPathTracer(scene, cfg, kind=kind, ctx=ctx).render(img, 0, SPP)
        images[kind] = img
    assert images["wavefront"].any()
    np.testing.assert_array_equal(images["wavefront"], images["megakernel"])
    del ctx
    gc.collect()   # the wavefront state of a whole frame in flight (tens of GB) goes with the context
    print("oracle render", flush=True)
    ref = sol_amd.new_image(W, H)
    O.OracleScene(scene, use_product_bvh=True).render(ref, 0, SPP, cfg, threads=min(16, os.cpu_count() or 1))
    compare_images(images["wavefront"], ref)
