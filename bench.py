"""Headline benchmark: Msamples/s of the wavefront renderer at 1920x1080, 64 spp (the
400-sphere field, ~2.05M triangles, max path length 16) against the plain GPU megakernel on the
same frame, on one MI355X.

A step is one full frame: 1920 x 1080 x 64 = 132,710,400 path samples, scene resident in HBM,
image accumulated in device memory. The two renderers share the per-(pixel, sample) random streams
and arithmetic, so their images are identical (equal RMSE by construction; checked below).

Also reported:
  roofline      the traversal kernel k_trace (closest hits of the extension rays and any hits of the
                shadow rays), timed with HIP events around each launch over a second timed pass
                of the same frames; its algorithmic bytes per ray come from the oracle's
                instrumented traversal of the same BVH.
  cpu_baseline  the CPU oracle (C++ restatement of sol's path tracer, OpenMP) on the host's cores,
                on a bounded sample of the same frame (all pixels, fewer samples per pixel).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WIDTH, HEIGHT, SPP, DEPTH = 1920, 1080, 64, 16
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of a ray query (implementation independent): ray 32 B, result 16 B (closest
# hit) or 4 B (any hit), 64 B per inner-node visit (both child boxes + child words), 48 B per
# triangle test.
B_RAY, B_HIT, B_FLAG, B_NODE, B_TRI = 32, 16, 4, 64, 48


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--no-megakernel", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-spp", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (each frame renders on one "
              "MI355X; see DESIGN.md for how the path shards)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import sol_amd
    from sol_amd import _ffi as F

    t0 = time.time()
    scene = sol_amd.Scene.sphere_field()
    bst = scene.build_stats()
    log(f"scene: {scene.num_tris} triangles, BVH build {bst.build_seconds:.2f}s + reinsertion "
        f"{bst.optimize_seconds:.2f}s (SAH {bst.sah_cost_built:.3f} -> {bst.sah_cost_final:.3f}), depth {bst.depth}")
    ctx = sol_amd.Context(0)
    info = ctx.device_info()
    g = ctx.upload(scene)
    cfg = sol_amd.PathTracerConfig(max_path_len=DEPTH)
    samples_per_step = WIDTH * HEIGHT * args.spp
    img = {k: torch.zeros((3, HEIGHT, WIDTH), dtype=torch.float32, device="cuda") for k in ("wave", "mega")}
    torch.cuda.synchronize()

    def frame(kind):
        r = F.RENDERER_WAVEFRONT if kind == "wave" else F.RENDERER_MEGAKERNEL
        g.render_device(img[kind].data_ptr(), WIDTH, HEIGHT, 0, args.spp, cfg, renderer=r)

    def timed(kind, profile):
        for _ in range(args.warmup):
            frame(kind)
        img[kind].zero_()
        torch.cuda.synchronize()
        ctx.set_profiling(profile)
        ctx.reset_stats()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(args.steps):
            frame(kind)
            log(f"{kind} step {i + 1}/{args.steps} done")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        st = ctx.stats()
        ctx.set_profiling(False)
        return dt / args.steps, st

    # Headline: the production path (bounce loop replayed as a captured graph).
    wave_s, gst = timed("wave", profile=False)
    value = samples_per_step / wave_s / 1e6
    mrays = (gst.extension_rays + gst.shadow_rays) / args.steps / wave_s / 1e6
    log(f"wavefront: {wave_s * 1e3:.1f} ms/frame, {value:.1f} Msamples/s, {mrays:.0f} Mrays/s, "
        f"graph replays {gst.graph_replays}")
    # Kernel durations: the same frames again with HIP events recorded around every kernel on the
    # renderer's stream (events recorded by a graph cannot be timed, so this pass launches the
    # iterations eagerly).
    prof_s, wst = timed("wave", profile=True)
    log(f"wavefront (HIP-event profiled): {prof_s * 1e3:.1f} ms/frame; per frame trace {wst.trace_ms / args.steps:.1f} ms, "
        f"sort {wst.sort_ms / args.steps:.1f} ms, shade {wst.shade_ms / args.steps:.1f} ms")

    mega = None
    if not args.no_megakernel:
        mega_s, mst = timed("mega", profile=False)
        mega = samples_per_step / mega_s / 1e6
        log(f"megakernel: {mega_s * 1e3:.1f} ms/frame, {mega:.1f} Msamples/s")
        a, b = img["wave"].cpu().numpy(), img["mega"].cpu().numpy()
        identical = float(np.all(a == b, axis=0).mean())
        rmse_wm = float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) / args.steps / args.spp)
        log(f"wavefront vs megakernel: identical pixels {identical:.6f}, RMSE between them {rmse_wm:.3e}")

    # Algorithmic bytes per extension ray from the oracle's instrumented traversal of the same BVH
    # (every 8th row, 1 spp: the same rays the GPU traces for those samples).
    from tests import oracle_lib
    osc = oracle_lib.OracleScene(scene, use_product_bvh=True)
    ost = oracle_lib.OracleStats()
    probe = sol_amd.new_image(WIDTH, HEIGHT)
    for y in range(0, HEIGHT, 8):
        osc.render(probe, 0, 1, cfg, rows=(y, y + 1), stats=ost)
    # k_trace processes the extension rays (closest hit, 16 B result) and the shadow rays (any
    # hit, 4 B result) of one bounce of every path in flight.
    ext_rays, shadow_rays = ost.extension_rays, ost.shadow_rays
    b_ext = B_RAY + B_HIT + B_NODE * ost.ext_inner_visits / ext_rays + B_TRI * ost.ext_tri_tests / ext_rays
    b_shadow = B_RAY + B_FLAG + B_NODE * ost.shadow_inner_visits / shadow_rays + B_TRI * ost.shadow_tri_tests / shadow_rays
    launches = max(1, wst.trace_launches)
    bytes_per_launch = (b_ext * wst.extension_rays + b_shadow * wst.shadow_rays) / launches
    launch_ms = wst.trace_ms / launches
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    log(f"trace kernel: {launches} launches, {launch_ms:.3f} ms avg, {bytes_per_launch / 1e6:.1f} MB algorithmic per launch "
        f"(extension {b_ext:.0f} B/ray: {ost.ext_inner_visits / ext_rays:.1f} nodes {ost.ext_tri_tests / ext_rays:.2f} tris; "
        f"shadow {b_shadow:.0f} B/ray) -> {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cimg = sol_amd.new_image(WIDTH, HEIGHT)
        t = time.perf_counter()
        osc.render(cimg, 0, args.cpu_spp, cfg, threads=threads)
        dt = time.perf_counter() - t
        cpu = {"value": WIDTH * HEIGHT * args.cpu_spp / dt / 1e6, "unit": "Msamples/s", "cores": threads,
               "kind": "port",
               "sample": f"1920x1080 at {args.cpu_spp} spp (samples 0..{args.cpu_spp - 1}) of the same frame, "
                         f"oracle over the same BVH, {dt:.1f} s"}
        log(f"cpu oracle: {cpu['value']:.2f} Msamples/s on {threads} threads")

    out = {
        "metric": "Msamples/s at 1080p 64spp, wavefront-batched vs plain GPU, at equal RMSE",
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wave_s * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: procedural 400-icosphere field (subdivision 4, 2,048,008 triangles), seeded",
        "config": {"workload": f"sphere field path tracing {WIDTH}x{HEIGHT} {args.spp}spp depth {DEPTH}",
                   "width": WIDTH, "height": HEIGHT, "spp": args.spp, "max_path_len": DEPTH,
                   "triangles": scene.num_tris, "renderer": "wavefront"},
        "megakernel_msamples_s": round(mega, 2) if mega else None,
        "speedup_vs_megakernel": round(value / mega, 3) if mega else None,
        "mrays_s": round(mrays, 1),
        "device": info.name.decode(errors="replace"),
        "roofline": {"bound": "hbm", "kernel": "k_trace", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": round(bytes_per_launch), "avg_launch_ms": round(launch_ms, 4),
                     "launches": int(launches), "profiled_ms_per_step": round(prof_s * 1e3, 2),
                     "kernel_ms_per_step": {"trace": round(wst.trace_ms / args.steps, 2),
                                            "sort": round(wst.sort_ms / args.steps, 2),
                                            "shade": round(wst.shade_ms / args.steps, 2)}},
        "cpu_baseline": cpu,
        "setup_s": round(time.time() - t0, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
