"""Headline benchmark: Msamples/s of the wavefront renderer at 1920x1080, 64 spp (the
400-sphere field, ~2.05M triangles, max path length 16) against the plain GPU megakernel on the
same frame, on one MI355X.

A step is one full frame: 1920 x 1080 x 64 = 132,710,400 path samples, scene resident in HBM,
image accumulated in device memory. The two renderers share the per-(pixel, sample) random streams
and arithmetic, so their images are identical (equal RMSE by construction; checked below).

Also reported:
  roofline      the traversal kernel k_trace (closest hits of the extension rays and any hits of the
                shadow rays), timed with HIP events around each launch over a second timed pass
                of the same frames. "achieved" is its HBM traffic per launch ("traffic", FETCH_SIZE
                + WRITE_SIZE from the committed rocprofv3 counter summary of this workload, scaled
                by the counters' reading of a copy of known size) over the measured launch time;
                "frac" = achieved / 8 TB/s. "request" is the rate of the survey's algorithmic bytes
                per ray (64 B per inner-node visit and 48 B per triangle test of the binary tree,
                counted on the same BVH and rays in tests/golden/traversal_counts.json): a request
                rate that caches serve mostly, so it can exceed HBM peak. "issue" names the binding
                limit: the busier of the VALU (vector ALU instruction issue) and the texture-address
                unit (vector memory instruction issue), with both busy fractions, from the committed
                counter summary. "copy_gbs" is a device-to-device copy timed here.
  cpu_baseline  the CPU oracle (C++ restatement of sol's path tracer, OpenMP) on the host's cores,
                on a bounded sample of the same frame (all pixels; by default all 64 samples, about
                11 s on 16 threads, and then the oracle's image is also compared with the wavefront
                image: "oracle_identical_pixels", as "megakernel_identical_pixels" for the megakernel).
  intersect     the intersection microbenchmark: 1M-triangle soup, 4,194,304 closest-hit rays and
                4,194,304 any-hit rays (device buffers), Mrays/s; the oracle on 262,144 of each.
  cornell       the Cornell box at 256x256, 16 spp, depth 8: both GPU renderers and the oracle.
  batched       64 uneven sphere-field scenes (sol_amd.workloads.batch_specs) rendered by one
                solhip_render_batch call, with the default launch parameters and after
                solhip_autotune, against the same jobs rendered one by one with the megakernel
                (all three timed alike: one warm-up, then the mean of the timed steps).
  host_frames   frames of 1 and 8 samples of the 1080p frame rendered into a host image (the sums
                return to host memory once per frame, as for sol::RenderJob over GpuPathTracer)
                against the same frames into a device image.
  adjoint       solhip_backward on the sphere field with 256 material groups at 512x512, 32 spp
                (Russian roulette off), with a central-difference check (h = 1e-3, common random
                numbers) of the largest albedo, Phong log-exponent and emission gradients.
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
COUNTS = os.path.join(ROOT, "tests", "golden", "traversal_counts.json")
TRAFFIC = os.path.join(ROOT, "profiles", "k_trace_sphere_field_1080p.pmc.txt")
COUNTERS = os.path.join(ROOT, "profiles", "wavefront_sphere_field_1080p.counters.md")
COPY_BYTES = 1 << 31


def copy_bandwidth(torch):
    """A plain streaming kernel over COPY_BYTES (reads COPY_BYTES, writes COPY_BYTES: torch.neg into
    a second buffer), best of 5, in GB/s. The same kernel calibrates the HBM byte counters."""
    src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(5):
        e0.record()
        torch.neg(src, out=dst)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    return 2 * COPY_BYTES / (best * 1e-3) / 1e9


def read_key(path, key):
    """The value of a "key: value" line of a committed profile summary, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    return None


def issue_limit(ta_busy, valu_busy):
    """The instruction-issue limit that binds k_trace: the busier of the two units."""
    if ta_busy is None or valu_busy is None:
        return None
    if valu_busy >= ta_busy:
        return "VALU (vector ALU instruction issue)"
    return "texture-address unit (vector memory instruction issue)"


def read_traffic():
    """HBM bytes per k_trace launch from the committed counter summary."""
    return read_key(TRAFFIC, "traffic_bytes_per_launch")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_intersect(sol_amd, torch, with_cpu):
    n = 1 << 22
    t = time.perf_counter()
    scene = sol_amd.Scene.triangle_soup(1_000_000)
    build_s = time.perf_counter() - t
    ctx = sol_amd.Context(0)
    g = ctx.upload(scene)
    out = {"triangles": scene.num_tris, "rays": n, "build_s": round(build_s, 2)}
    for kind, name in ((0, "closest"), (1, "any")):
        rays = sol_amd.gen_rays(kind, n)
        d_rays = torch.from_numpy(rays).cuda()
        d_out = torch.empty((n, 4) if kind == 0 else (n,), dtype=torch.float32 if kind == 0 else torch.int32,
                            device="cuda")
        torch.cuda.synchronize()
        g.intersect_device(d_rays.data_ptr(), n, d_out.data_ptr(), any_hit=kind == 1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 5
        for _ in range(reps):
            g.intersect_device(d_rays.data_ptr(), n, d_out.data_ptr(), any_hit=kind == 1)
        torch.cuda.synchronize()
        out[f"{name}_mrays_s"] = round(n * reps / (time.perf_counter() - t) / 1e6, 1)
        if with_cpu:
            from tests import oracle_lib
            osc = oracle_lib.OracleScene(scene, use_product_bvh=True)
            m = 1 << 18
            t = time.perf_counter()
            osc.intersect(rays[:m], any_hit=kind == 1)
            out[f"cpu_{name}_mrays_s"] = round(m / (time.perf_counter() - t) / 1e6, 2)
    log(f"intersect: {out}")
    g.close()
    ctx.close()
    return out


def bench_cornell(sol_amd, F, torch, with_cpu):
    w = h = 256
    spp = 16
    scene = sol_amd.Scene.cornell_box(aspect=1.0)
    ctx = sol_amd.Context(0)
    g = ctx.upload(scene)
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    img = torch.zeros((3, h, w), dtype=torch.float32, device="cuda")
    out = {"width": w, "height": h, "spp": spp, "max_path_len": 8}
    for name, r in (("wavefront", F.RENDERER_WAVEFRONT), ("megakernel", F.RENDERER_MEGAKERNEL)):
        g.render_device(img.data_ptr(), w, h, 0, spp, cfg, renderer=r)
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 5
        for _ in range(reps):
            g.render_device(img.data_ptr(), w, h, 0, spp, cfg, renderer=r)
        torch.cuda.synchronize()
        out[f"{name}_msamples_s"] = round(w * h * spp * reps / (time.perf_counter() - t) / 1e6, 1)
    if with_cpu:
        from tests import oracle_lib
        threads = min(16, os.cpu_count() or 1)
        cimg = sol_amd.new_image(w, h)
        t = time.perf_counter()
        oracle_lib.OracleScene(scene).render(cimg, 0, spp, cfg, threads=threads)
        out["cpu_msamples_s"] = round(w * h * spp / (time.perf_counter() - t) / 1e6, 2)
        out["cpu_threads"] = threads
    log(f"cornell: {out}")
    g.close()
    ctx.close()
    return out


def bench_batched(sol_amd, F, torch, steps):
    from sol_amd.workloads import BATCH_DEPTH, batch_specs
    specs = batch_specs()
    t = time.perf_counter()
    scenes = [sp.scene() for sp in specs]
    build_s = time.perf_counter() - t
    ctx = sol_amd.Context(0)
    gs = [ctx.upload(sc) for sc in scenes]
    cfg = sol_amd.PathTracerConfig(max_path_len=BATCH_DEPTH)
    bufs = [torch.zeros((3, sp.height, sp.width), dtype=torch.float32, device="cuda") for sp in specs]
    jobs = [(g, sp.width, sp.height, 0, sp.spp, b.data_ptr()) for sp, g, b in zip(specs, gs, bufs)]
    samples = sum(sp.width * sp.height * sp.spp for sp in specs)
    torch.cuda.synchronize()

    def timed(fn):
        fn()                                   # warm-up: arenas, graph capture, code objects
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    default_s = timed(lambda: ctx.render_batch(jobs, cfg))
    t = time.perf_counter()
    tuned = ctx.autotune_batch(jobs, cfg)
    tune_s = time.perf_counter() - t
    batch_s = timed(lambda: ctx.render_batch(jobs, cfg))

    def one_by_one():
        for sp, g, b in zip(specs, gs, bufs):
            g.render_device(b.data_ptr(), sp.width, sp.height, 0, sp.spp, cfg, renderer=F.RENDERER_MEGAKERNEL)

    mega_s = timed(one_by_one)
    # the same batch with every BVH built on the GPU
    torch.cuda.synchronize()
    t = time.perf_counter()
    gbs = [ctx.build(sc) for sc in scenes]
    torch.cuda.synchronize()
    gpu_build_s = time.perf_counter() - t
    host_build_s = sum(sc.build_stats().build_seconds + sc.build_stats().optimize_seconds for sc in scenes)
    gjobs = [(g, sp.width, sp.height, 0, sp.spp, b.data_ptr()) for sp, g, b in zip(specs, gbs, bufs)]
    ctx.render_batch(gjobs, cfg)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        ctx.render_batch(gjobs, cfg)
    torch.cuda.synchronize()
    gbatch_s = (time.perf_counter() - t) / steps
    for g in gbs:
        g.close()
    out = {"scenes": len(specs), "triangles": int(sum(sc.num_tris for sc in scenes)), "samples": samples,
           "msamples_s": round(samples / batch_s / 1e6, 2), "ms": round(batch_s * 1e3, 2),
           "default_msamples_s": round(samples / default_s / 1e6, 2), "default_ms": round(default_s * 1e3, 2),
           "megakernel_one_by_one_msamples_s": round(samples / mega_s / 1e6, 2),
           "speedup_vs_megakernel_one_by_one": round(mega_s / batch_s, 3),
           "timing": f"each: one warm-up, then the mean of {steps} calls",
           "tuned": {k: getattr(tuned, k) for k, _ in tuned._fields_},
           "build_s": round(build_s, 1), "autotune_s": round(tune_s, 1),
           "host_bvh_build_s": round(host_build_s, 2), "gpu_bvh_build_s": round(gpu_build_s, 3),
           "gpu_bvh_msamples_s": round(samples / gbatch_s / 1e6, 2)}
    log(f"batched: {out}")
    for g in gs:
        g.close()
    ctx.close()
    return out


def bench_gpu_build(sol_amd, F, torch, ctx, scene, cfg, spp):
    """The headline scene's BVH built on the GPU (solhip_scene_build) against the host build, and
    the frame rendered on it."""
    bst = scene.build_stats()
    warm = sol_amd.Scene.triangle_soup(4096, seed=1)
    ctx.build(warm).close()                      # first-use costs (code objects) out of the timing
    torch.cuda.synchronize()
    t = time.perf_counter()
    gb = ctx.build(scene)
    build_s = time.perf_counter() - t
    st = gb.build_stats
    img = torch.zeros((3, HEIGHT, WIDTH), dtype=torch.float32, device="cuda")
    gb.render_device(img.data_ptr(), WIDTH, HEIGHT, 0, spp, cfg, renderer=F.RENDERER_WAVEFRONT)
    torch.cuda.synchronize()
    t = time.perf_counter()
    gb.render_device(img.data_ptr(), WIDTH, HEIGHT, 0, spp, cfg, renderer=F.RENDERER_WAVEFRONT)
    torch.cuda.synchronize()
    frame_s = time.perf_counter() - t
    out = {"build_ms": round(build_s * 1e3, 1), "host_build_s": round(bst.build_seconds + bst.optimize_seconds, 2),
           "sah": round(st.sah_cost, 3), "host_sah": round(bst.sah_cost_final, 3), "passes": st.iterations,
           "depth": st.depth, "depth4": st.depth4,
           "wavefront_msamples_s": round(WIDTH * HEIGHT * spp / frame_s / 1e6, 2)}
    log(f"gpu build: {out}")
    gb.close()
    return out


def bench_host_frames(sol_amd, F, torch, g, cfg):
    """Frames of a few samples rendered into a host image (what sol::RenderJob does through
    GpuPathTracer: each frame's sums come back to host memory and are added there) against the
    same frames into a device image, at 1920x1080 on the wavefront renderer."""
    import numpy as np
    out = {"width": WIDTH, "height": HEIGHT}
    host = np.zeros((3, HEIGHT, WIDTH), np.float32)
    dev = torch.zeros((3, HEIGHT, WIDTH), dtype=torch.float32, device="cuda")
    for spf in (1, 8):
        frames = 4
        g.render(host, 0, spf, cfg)                       # warm-up
        g.render_device(dev.data_ptr(), WIDTH, HEIGHT, 0, spf, cfg)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for f in range(frames):
            g.render(host, f * spf, spf, cfg)
        host_s = (time.perf_counter() - t) / frames
        torch.cuda.synchronize()
        t = time.perf_counter()
        for f in range(frames):
            g.render_device(dev.data_ptr(), WIDTH, HEIGHT, f * spf, spf, cfg)
        torch.cuda.synchronize()
        dev_s = (time.perf_counter() - t) / frames
        out[f"spf{spf}"] = {"host_image_ms": round(host_s * 1e3, 2), "device_image_ms": round(dev_s * 1e3, 2),
                            "host_image_msamples_s": round(WIDTH * HEIGHT * spf / host_s / 1e6, 1),
                            "device_image_msamples_s": round(WIDTH * HEIGHT * spf / dev_s / 1e6, 1)}
    log(f"host frames: {out}")
    return out


def bench_adjoint(sol_amd, torch):
    import numpy as np
    from sol_amd.diff import DiffRenderer
    w = h = 512
    spp = 32
    scene = sol_amd.Scene.sphere_field(num_groups=256, aspect=1.0)
    ctx = sol_amd.Context(0)
    cfg = sol_amd.PathTracerConfig(max_path_len=DEPTH, disable_rr=True)
    dr = DiffRenderer(ctx.upload(scene), w, h, 0, spp, cfg)
    albedo, log_ns, le = dr.parameters()
    rng = np.random.default_rng(0)
    target = dr.render_with(np.clip(albedo + rng.normal(0, 0.1, albedo.shape), 0.02, 0.98).astype(np.float32),
                            (log_ns + rng.normal(0, 0.1, log_ns.shape)).astype(np.float32),
                            np.maximum(le + rng.normal(0, 0.1, le.shape), 0).astype(np.float32)).double()
    def timed(fn, reps=3):
        # median wall time of `reps` calls (one call is a few tens of ms)
        times, out = [], None
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        return sorted(times)[reps // 2], out

    fwd_s, img = timed(lambda: dr.render_with(albedo, log_ns, le))
    d_image = (2 * (img.double() - target)).float()
    dr.gradients(d_image, albedo, log_ns, le)
    bwd_s, (da, dn, de) = timed(lambda: dr.gradients(d_image, albedo, log_ns, le))
    da, dn, de = da.cpu().numpy(), dn.cpu().numpy(), de.cpu().numpy()

    def loss(a, n, e):
        return float(((dr.render_with(a, n, e).double() - target) ** 2).sum())

    checks = []
    for k, grad in ((0, da), (1, dn), (2, de)):
        idx = np.unravel_index(int(np.argmax(np.abs(grad))), grad.shape)
        plus = [albedo.copy(), log_ns.copy(), le.copy()]
        minus = [albedo.copy(), log_ns.copy(), le.copy()]
        plus[k][idx] += 1e-3
        minus[k][idx] -= 1e-3
        fd = (loss(*plus) - loss(*minus)) / 2e-3
        checks.append(abs(float(grad[idx]) - fd) / abs(fd))
    samples = w * h * spp
    out = {"workload": "sphere field, 256 material groups, 512x512, 32 spp, depth 16, Russian roulette off",
           "backward_ms": round(bwd_s * 1e3, 2), "backward_msamples_s": round(samples / bwd_s / 1e6, 2),
           "forward_ms": round(fwd_s * 1e3, 2), "timing": "median of 3 calls each", "fd_params_checked": len(checks),
           "fd_max_rel_err": float(f"{max(checks):.3g}")}
    log(f"adjoint: {out}")
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--no-megakernel", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-spp", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the intersection and Cornell box legs")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--only-cornell", action="store_true",
                    help="run only the Cornell box leg (for kernel traces of a small frame) and print its object")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the 64-scene batch leg (for kernel traces of the batch) and print its object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported. A frame renders on one MI355X; "
              "across GPUs the work splits by sample ranges or image tiles with one final image sum, which is "
              "not implemented", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import sol_amd
    from sol_amd import _ffi as F

    if args.only_cornell:
        print(json.dumps(bench_cornell(sol_amd, F, torch, not args.no_cpu)), flush=True)
        return
    if args.only_batch:
        print(json.dumps(bench_batched(sol_amd, F, torch, args.steps)), flush=True)
        return

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

    mega, identical = None, None
    if not args.no_megakernel:
        mega_s, mst = timed("mega", profile=False)
        mega = samples_per_step / mega_s / 1e6
        log(f"megakernel: {mega_s * 1e3:.1f} ms/frame, {mega:.1f} Msamples/s")
        a, b = img["wave"].cpu().numpy(), img["mega"].cpu().numpy()
        identical = float(np.all(a == b, axis=0).mean())
        rmse_wm = float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) / args.steps / args.spp)
        log(f"wavefront vs megakernel: identical pixels {identical:.6f}, RMSE between them {rmse_wm:.3e}")

    # Algorithmic bytes per ray: inner-node visits and triangle tests per ray of the same BVH and
    # rays (committed fixture). k_trace processes the extension rays (closest hit, 16 B result)
    # and the shadow rays (any hit, 4 B result) of one bounce of every path in flight.
    with open(COUNTS) as f:
        counts = json.load(f)["sphere_field_1080p"]
    sig = counts["bvh"]
    if (sig["triangles"], sig["nodes"], sig["depth"]) != (scene.num_tris, bst.num_nodes, bst.depth):
        log(f"warning: BVH differs from the one traversal_counts.json was counted on ({sig}); regenerate it")
    ce, cs = counts["extension"], counts["shadow"]
    b_ext = B_RAY + B_HIT + B_NODE * ce["inner_visits_per_ray"] + B_TRI * ce["tri_tests_per_ray"]
    b_shadow = B_RAY + B_FLAG + B_NODE * cs["inner_visits_per_ray"] + B_TRI * cs["tri_tests_per_ray"]
    launches = max(1, wst.trace_launches)
    bytes_per_launch = (b_ext * wst.extension_rays + b_shadow * wst.shadow_rays) / launches
    launch_ms = wst.trace_ms / launches
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    log(f"trace kernel: {launches} launches, {launch_ms:.3f} ms avg, {bytes_per_launch / 1e6:.1f} MB algorithmic per launch "
        f"(extension {b_ext:.0f} B/ray: {ce['inner_visits_per_ray']:.1f} nodes {ce['tri_tests_per_ray']:.2f} tris; "
        f"shadow {b_shadow:.0f} B/ray) -> {achieved:.0f} GB/s")
    copy_gbs = copy_bandwidth(torch)
    traffic = read_traffic()
    dram_gbs = traffic / (launch_ms * 1e6) if traffic and launch_ms > 0 else None
    ta_busy = read_key(COUNTERS, "k_trace_ta_busy_frac")
    valu_busy = read_key(COUNTERS, "k_trace_valu_busy_frac")
    log(f"device copy: {copy_gbs:.0f} GB/s; k_trace HBM traffic per launch: {traffic} -> {dram_gbs} GB/s; "
        f"TA busy {ta_busy}, VALU busy {valu_busy}")

    gpu_build = None if args.no_small else bench_gpu_build(sol_amd, F, torch, ctx, scene, cfg, args.spp)
    host_frames = None if args.no_small else bench_host_frames(sol_amd, F, torch, g, cfg)
    intersect = None if args.no_small else bench_intersect(sol_amd, torch, not args.no_cpu)
    cornell = None if args.no_small else bench_cornell(sol_amd, F, torch, not args.no_cpu)
    batched = None if args.no_batch else bench_batched(sol_amd, F, torch, args.steps)
    adjoint = None if args.no_adjoint else bench_adjoint(sol_amd, torch)

    cpu = None
    if not args.no_cpu:
        from tests import oracle_lib
        osc = oracle_lib.OracleScene(scene, use_product_bvh=True)
        threads = min(16, os.cpu_count() or 1)
        cimg = sol_amd.new_image(WIDTH, HEIGHT)
        t = time.perf_counter()
        osc.render(cimg, 0, args.cpu_spp, cfg, threads=threads)
        dt = time.perf_counter() - t
        # one thread on one sample of the frame (a secondary figure)
        one = sol_amd.new_image(WIDTH, HEIGHT)
        t = time.perf_counter()
        osc.render(one, 0, 1, cfg, threads=1)
        dt1 = time.perf_counter() - t
        # The oracle's frame, accumulated as often as the timed wavefront frames were (the same fp32
        # additions in the same order), against the wavefront image: the checker at full size.
        oracle_identical = None
        if args.cpu_spp == args.spp:
            acc = np.zeros_like(cimg)
            for _ in range(args.steps):
                acc = acc + cimg
            oracle_identical = float(np.all(img["wave"].cpu().numpy() == acc, axis=0).mean())
            log(f"wavefront vs oracle: identical pixels {oracle_identical:.6f}")
        cpu = {"value": WIDTH * HEIGHT * args.cpu_spp / dt / 1e6, "unit": "Msamples/s", "cores": threads,
               "kind": "port",
               "sample": f"1920x1080 at {args.cpu_spp} spp (samples 0..{args.cpu_spp - 1}) of the same frame, "
                         f"oracle over the same BVH, {dt:.1f} s",
               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
               "oracle_identical_pixels": oracle_identical,
               "one_thread": {"value": round(WIDTH * HEIGHT / dt1 / 1e6, 3),
                              "sample": f"1920x1080 at 1 spp (sample 0), {dt1:.1f} s"}}
        log(f"cpu oracle: {cpu['value']:.2f} Msamples/s on {threads} threads; one thread "
            f"{cpu['one_thread']['value']:.3f} Msamples/s")

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
        "megakernel_identical_pixels": identical,
        "mrays_s": round(mrays, 1),
        "device": info.name.decode(errors="replace"),
        "roofline": {"bound": "hbm", "kernel": "k_trace", "achieved": round(dram_gbs, 1) if dram_gbs else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dram_gbs / HBM_PEAK_GBS, 4) if dram_gbs else None,
                     "traffic": round(traffic) if traffic else None, "copy_gbs": round(copy_gbs, 1),
                     "request": {"gbs": round(achieved, 1), "bytes_per_launch": round(bytes_per_launch),
                                 "frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 4),
                                 "model": "survey bytes per ray: 32 ray + 16 / 4 result + 64 per binary inner-node "
                                          "visit + 48 per triangle test; served mostly by LDS, L2 and MALL"},
                     "issue": {"limit": issue_limit(ta_busy, valu_busy),
                               "ta_busy_frac": ta_busy, "valu_busy_frac": valu_busy,
                               "source": os.path.relpath(COUNTERS, ROOT)},
                     "avg_launch_ms": round(launch_ms, 4),
                     "launches": int(launches), "profiled_ms_per_step": round(prof_s * 1e3, 2),
                     "kernel_ms_per_step": {"trace": round(wst.trace_ms / args.steps, 2),
                                            "sort": round(wst.sort_ms / args.steps, 2),
                                            "shade": round(wst.shade_ms / args.steps, 2)}},
        "cpu_baseline": cpu,
        "gpu_build": gpu_build,
        "host_frames": host_frames,
        "intersect": intersect,
        "cornell": cornell,
        "batched": batched,
        "adjoint": adjoint,
        "setup_s": round(time.time() - t0, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
