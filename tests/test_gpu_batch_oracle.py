"""Batched rendering (solhip_render_batch) against the oracle, job by job, bit for bit.

A batch runs many scenes and images in one wavefront run: tiles of every job feed one global work
queue, paths of different scenes share the queues and the sample buffer, and per-pixel sums are
reduced chunk by chunk. None of that may change a pixel: each job must equal the oracle's
PathTracer::render of its own scene, image and sample range (src/algorithms/path_tracer.cpp:9-23),
whatever the tile shape, the paths in flight, the chunking or the queue split (segments per queue,
material-split or merged shading queues). The full 64-scene batched workload
is run as well, with properties that hold at any size.
"""
import os

import numpy as np
import pytest
import torch

import sol_amd
from sol_amd import _ffi as F
from sol_amd.workloads import BATCH_DEPTH, batch_specs
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell")


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


def _scenes():
    return {
        "cornell": sol_amd.Scene.cornell_box(aspect=1.5),
        "spheres": sol_amd.Scene.sphere_field(num_spheres=16, subdivision=1, seed=3, extent=3.0, aspect=1.0),
        "glossy": sol_amd.Scene.load(os.path.join(GOLDEN, "cornell_box_glossy.toml"), width=45, height=17),
        "specular": sol_amd.Scene.load(os.path.join(GOLDEN, "cornell_box_specular.toml"), width=31, height=29),
    }


# (scene, width, height, first sample, sample count): distinct scenes, uneven sizes (none a multiple
# of the 16 x 8 tile), a scene used twice, sample indices far from zero, two jobs adding into one image.
JOBS = [
    ("cornell", 37, 25, 0, 24),
    ("spheres", 64, 48, 7, 24),
    ("glossy", 45, 17, 3, 20),
    ("spheres", 20, 33, 1_000_003, 8),
    ("specular", 31, 29, 11, 16),
    ("cornell", 37, 25, 24, 6),          # adds into job 0's image: samples 24..29 after 0..23
]
SHARED = {5: 0}


# Chunks larger than the paths in flight restart paths from the work items; chunks that fit (chunk
# 0: as many items as paths) run in place, several of them here.
@pytest.mark.parametrize("slots,chunk,tiles,segments,shade_queues", [
    (5000, 65536, (16, 8), 8, 3),
    (1 << 20, 1 << 20, (32, 32), 8, 3),
    (5000, 65536, (16, 8), 1, 1),
    (1 << 20, 1 << 20, (32, 32), 2, 1),
    (1 << 16, 0, (16, 8), 1, 3),
])
def test_batch_matches_oracle_per_job(ctx, slots, chunk, tiles, segments, shade_queues):
    scenes = _scenes()
    gs = {k: ctx.upload(s) for k, s in scenes.items()}
    cfg = sol_amd.PathTracerConfig(max_path_len=BATCH_DEPTH)
    bufs = {}
    jobs = []
    for k, (name, w, h, si, sc) in enumerate(JOBS):
        if k in SHARED:
            buf = bufs[SHARED[k]]
        else:
            buf = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
            bufs[k] = buf
        jobs.append((gs[name], w, h, si, sc, buf.data_ptr()))
    torch.cuda.synchronize()
    total = sum(w * h * sc for _, w, h, _, sc in JOBS)
    print(f"batch of {len(JOBS)} jobs, {total} samples, slots {slots}, chunk {chunk}, tiles {tiles}, "
          f"{segments} segments, {shade_queues} shade queues", flush=True)
    ctx.set_tuning(path_slots=slots, chunk_items=chunk, tile_w=tiles[0], tile_h=tiles[1], graph_iters=4,
                   queue_segments=segments, shade_queues=shade_queues)
    try:
        ctx.reset_stats()
        ctx.render_batch(jobs, cfg)
        st = ctx.stats()
    finally:
        ctx.set_tuning(path_slots=0, chunk_items=0, tile_w=0, tile_h=0, graph_iters=0, queue_segments=0,
                       shade_queues=0)
    assert st.samples == total
    if chunk in (0, 65536):
        assert total > 2 * 65536                 # at least three chunks of work
    # the oracle renders each job on its own; a shared image receives its jobs in job order
    want = {}
    for k, (name, w, h, si, sc) in enumerate(JOBS):
        img = want[SHARED.get(k, k)] if k in SHARED else np.zeros((3, h, w), np.float32)
        O.OracleScene(scenes[name]).render(img, si, sc, cfg)
        want[SHARED.get(k, k)] = img
    for k, buf in bufs.items():
        got = buf.cpu().numpy()
        assert want[k].any()
        print(f"job {k} ({JOBS[k][0]}): identical pixels {np.all(got == want[k], axis=0).mean():.6f}", flush=True)
        np.testing.assert_array_equal(got, want[k])


def test_full_batched_workload_properties(ctx):
    # The 64-scene batched workload at its full size (sphere counts log-uniform 4-2000, images from
    # 64^2 to 1024^2, 4 / 16 / 64 spp, depth 12) in one render_batch with the default tuning: every
    # image finite and lit, every requested sample taken, and the smallest jobs equal to their
    # single renders and to the oracle.
    specs = batch_specs()
    cfg = sol_amd.PathTracerConfig(max_path_len=BATCH_DEPTH)
    scenes = [sp.scene() for sp in specs]
    gs = [ctx.upload(s) for s in scenes]
    bufs = [torch.zeros((3, sp.height, sp.width), dtype=torch.float32, device="cuda") for sp in specs]
    torch.cuda.synchronize()
    total = sum(sp.width * sp.height * sp.spp for sp in specs)
    print(f"64-scene batch: {total} samples, {sum(s.num_tris for s in scenes)} triangles", flush=True)
    ctx.reset_stats()
    ctx.render_batch([(g, sp.width, sp.height, 0, sp.spp, b.data_ptr()) for sp, g, b in zip(specs, gs, bufs)], cfg)
    st = ctx.stats()
    assert st.samples == total
    assert st.extension_rays >= total
    for b in bufs:
        img = b.cpu().numpy()
        assert np.isfinite(img).all() and img.mean() > 0
    smallest = sorted(range(len(specs)), key=lambda i: specs[i].width * specs[i].height * specs[i].spp)[:3]
    for i in smallest:
        sp = specs[i]
        single = sol_amd.new_image(sp.width, sp.height)
        gs[i].render(single, 0, sp.spp, cfg, renderer=F.RENDERER_MEGAKERNEL)
        got = bufs[i].cpu().numpy()
        np.testing.assert_array_equal(got, single)
        ref = sol_amd.new_image(sp.width, sp.height)
        O.OracleScene(scenes[i], use_product_bvh=True).render(ref, 0, sp.spp, cfg)
        print(f"job {i}: {sp.width}x{sp.height} {sp.spp} spp, {scenes[i].num_tris} triangles", flush=True)
        np.testing.assert_array_equal(got, ref)
