"""Batched rendering of the uneven multi-scene workload and the launch-parameter autotuner.

Images do not depend on launch parameters or on batching: every sample follows its own
(pixel, sample index) random stream and the per-pixel sums run in sample order, so a batch equals
the single renders and a tuned context renders the same image bit for bit.
"""
import os

import numpy as np
import pytest
import torch

import sol_amd
from sol_amd import _ffi as F
from sol_amd.workloads import BATCH_DEPTH, batch_specs

pytestmark = pytest.mark.gpu

SPECULAR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell", "cornell_box_specular.toml")


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


@pytest.mark.parametrize("block", [256, 128, 64])
def test_batch_workload_equals_single_renders(ctx, block):
    # The multi-scene traversal kernel splits its workgroup's LDS copy of top BVH nodes between
    # its waves (20, 40 or 80 nodes a wave for 256-, 128- or 64-thread workgroups).
    specs = batch_specs(count=8, seed=7, max_spheres=300, sides=(64, 128, 256), spps=(1, 4))
    cfg = sol_amd.PathTracerConfig(max_path_len=BATCH_DEPTH)
    gs = [ctx.upload(sp.scene()) for sp in specs]
    want = []
    for sp, g in zip(specs, gs):
        img = sol_amd.new_image(sp.width, sp.height)
        g.render(img, 0, sp.spp, cfg, renderer=F.RENDERER_MEGAKERNEL)
        want.append(img)
    bufs = [torch.zeros((3, sp.height, sp.width), dtype=torch.float32, device="cuda") for sp in specs]
    torch.cuda.synchronize()
    ctx.set_tuning(block_size=block)
    try:
        ctx.render_batch([(g, sp.width, sp.height, 0, sp.spp, b.data_ptr()) for sp, g, b in zip(specs, gs, bufs)], cfg)
    finally:
        ctx.set_tuning(block_size=0)
    for sp, b, wimg in zip(specs, bufs, want):
        got = b.cpu().numpy()
        assert got.any()
        np.testing.assert_array_equal(got, wimg)


def test_autotune_returns_valid_parameters_and_keeps_the_image(ctx):
    scene = sol_amd.Scene.sphere_field(num_spheres=40, subdivision=2, seed=3, extent=4.0, radius_max=1.0)
    g = ctx.upload(scene)
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    before = sol_amd.new_image(160, 90)
    g.render(before, 0, 4, cfg)
    ctx.reset_stats()
    t = g.autotune(160, 90, budget_spp=2, config=cfg)
    assert t.block_size in (128, 256) and t.lds_stack in (8, 16)
    assert 4096 <= t.path_slots <= (1 << 27) and t.graph_iters in (8, 16, 32)
    assert (t.tile_w, t.tile_h) in ((16, 16), (32, 32), (64, 16), (128, 8)) and t.bvh_width in (2, 4)
    assert t.queue_segments in (1, 2, 4, 8) and t.shade_queues in (1, 3)
    cur = ctx.tuning()
    fields = ("block_size", "lds_stack", "path_slots", "graph_iters", "tile_w", "tile_h", "bvh_width",
              "queue_segments", "shade_queues")
    assert [getattr(cur, f) for f in fields] == [getattr(t, f) for f in fields]
    assert ctx.stats().samples == 0          # trial renders are not counted
    after = sol_amd.new_image(160, 90)
    g.render(after, 0, 4, cfg)
    np.testing.assert_array_equal(after, before)
    ctx.set_tuning(**{k: 0 for k in fields})


@pytest.mark.parametrize("name", ["spheres", "specular"])
def test_queue_split_keeps_the_image(ctx, name):
    # Every queue split (1, 2, 4 or 8 segments per queue; three material queues or one merged
    # queue) renders the megakernel's image bit for bit. Few paths in flight and small chunks make
    # paths restart from the other segments' work items and cross chunk boundaries.
    scene = {"spheres": lambda: sol_amd.Scene.sphere_field(num_spheres=40, subdivision=2, seed=3, extent=4.0,
                                                           radius_max=1.0),
             "specular": lambda: sol_amd.Scene.load(SPECULAR, width=72, height=60)}[name]()
    g = ctx.upload(scene)
    cfg = sol_amd.PathTracerConfig(max_path_len=10)
    want = sol_amd.new_image(72, 60)
    g.render(want, 5, 24, cfg, renderer=F.RENDERER_MEGAKERNEL)
    assert want.any()
    try:
        for segments in (1, 2, 4, 8):
            for shade_queues in (1, 3):
                ctx.set_tuning(path_slots=4096, chunk_items=65536, graph_iters=4, queue_segments=segments,
                               shade_queues=shade_queues)
                got = sol_amd.new_image(72, 60)
                g.render(got, 5, 24, cfg)
                print(f"{name}: {segments} segments, {shade_queues} shade queues", flush=True)
                np.testing.assert_array_equal(got, want)
    finally:
        ctx.set_tuning(path_slots=0, chunk_items=0, graph_iters=0, queue_segments=0, shade_queues=0)


def test_default_paths_in_flight_follow_device_memory(ctx):
    # the largest power of two up to 2^27 whose 348 bytes a slot fit in a quarter of the memory
    mem = ctx.device_info().total_memory
    ctx.set_tuning(path_slots=0)
    want = 1 << 27
    while want > 4096 and want * 348 > mem // 4:
        want //= 2
    assert ctx.tuning().path_slots == want
    if mem >= 4 * 348 * (1 << 27):          # an MI355X (288 GB): a 1080p 64-spp frame in flight at once
        assert want == 1 << 27


def test_tuning_rejects_unsupported_values(ctx):
    with pytest.raises(F.SolError):
        ctx.set_tuning(block_size=96)
    with pytest.raises(F.SolError):
        ctx.set_tuning(path_slots=10)
    with pytest.raises(F.SolError):
        ctx.set_tuning(path_slots=(1 << 27) + 1)
    with pytest.raises(F.SolError):
        ctx.set_tuning(bvh_width=3)
    for segments in (3, 16):
        with pytest.raises(F.SolError):
            ctx.set_tuning(queue_segments=segments)
    with pytest.raises(F.SolError):
        ctx.set_tuning(shade_queues=2)
    g = ctx.upload(sol_amd.Scene.cornell_box())
    with pytest.raises(F.SolError):
        g.autotune(0, 16)
