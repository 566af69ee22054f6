"""Benchmark of the ray-tracer hot path on one MI355X.

Workload (BVH build + traversal microbenchmark): 1,000,000 random float32 triangles
(centres uniform in the unit cube, Gaussian edges with lognormal(-5, 0.5) scale, seed 1),
built on the GPU (63-bit Morton LBVH + binned SAH with spatial splits, bit-exact with the
reference builder); 4,194,304 incoherent rays (origins uniform in the cube, directions
uniform on the sphere) plus 4,194,304 coherent primary rays of a 2048x2048 pinhole.

A step = one closest-hit pass over the 4M incoherent rays (inputs resident in HBM).
value = incoherent closest-hit Mrays/s. Coherent closest-hit, any-hit and build rates
are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "ray-tracer_amd")]

HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, a wave64 VALU instruction every 4 cycles per SIMD, at
# the 2.4 GHz peak engine clock (MI355X's published 157.3 TFLOP/s vector fp32 is this rate
# x 64 lanes x 2 flops per FMA x 2 for packed fp32).
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 4

# Counter passes of the measured kernels (each pass fits the per-block counter limits:
# FETCH_SIZE takes 3 TCC counters, WRITE_SIZE 2).
PMC_PASSES = [["FETCH_SIZE"],
              ["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum"],
              ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES",
               "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR"]]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-render", action="store_true",
                    help="skip the render legs (batched 64-scene, differentiable, Cornell box)")
    ap.add_argument("--batched-scenes", type=int, default=64)
    ap.add_argument("--no-inverse", action="store_true", help="skip the terrain inverse-render leg")
    ap.add_argument("--inverse-steps", type=int, default=3,
                    help="Adam iterations of the terrain inverse render (the full fit is 200)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (the reference point that scales HBM "
                         "counters under the profiler)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time with the library's default launch shapes instead of tuning them here first")
    ap.add_argument("--tuning", default=None,
                    help='JSON launch shapes applied to the headline after tuning, e.g. \'{"trace_lds_depth": 12}\' '
                         "(with --no-autotune: exactly these on top of the defaults)")
    ap.add_argument("--skip-extras", action="store_true",
                    help="time only the headline launch (used under the profiler so k_trace's average is "
                         "the incoherent launch alone)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (HBM traffic, L2 hit rate, VALU issue)")
    ap.add_argument("--counter-legs", default="trace,c3,c4,c5",
                    help="legs measured under rocprofv3: trace (headline), c3 (batched render, plain and "
                         "batched), c4 (differentiable render), c5 (terrain inverse render)")
    ap.add_argument("--counters-dir", default=None, help="keep the raw rocprofv3 output here")
    ap.add_argument("--counter-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--only-leg", default=None, choices=["c1", "c3", "c4", "c5"],
                    help="run one render leg alone (for profiling it) and print its figures")
    args = ap.parse_args()
    if args.counter_child:
        counter_child(args)
        return
    if args.only_leg:
        import torch
        import rtx
        if args.tuning:
            rtx.set_tuning(**json.loads(args.tuning))
        stream = torch.cuda.current_stream()
        leg = {"c1": lambda: bench_cornell(stream, cpu=not args.no_cpu, tune=not args.no_autotune),
               "c3": lambda: bench_batched(stream, args.batched_scenes, tune=not args.no_autotune),
               "c4": lambda: bench_differentiable(stream),
               "c5": lambda: bench_inverse(stream, steps=args.inverse_steps)}[args.only_leg]()
        print(json.dumps({"leg": args.only_leg, "result": leg}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the traversal shards by rays with the "
              "BVH replicated, and no multi-GPU figure is produced here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import rtx
    from rtx import autotune
    from rtx import scenes as S

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    v, t = S.random_triangles(args.tris, seed=1)
    vd = torch.from_numpy(v).to(dev)
    td = torch.from_numpy(t.view(np.int32)).to(dev)
    torch.cuda.synchronize()
    # build: warm once (allocations, code objects), then time
    rtx.build_device(vd, td, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bvh = rtx.build_device(vd, td, stream=stream)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    info = bvh.info()

    n = args.rays
    inc = torch.from_numpy(S.incoherent_rays(n, seed=1)).to(dev)
    side = int(round(n ** 0.5))
    cam = S.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 45.0)
    coh = torch.from_numpy(S.pinhole_rays(cam, side, side)).to(dev)
    hits = torch.empty((n, 4), dtype=torch.float32, device=dev)
    occ = torch.empty(n, dtype=torch.uint8, device=dev)
    counts = torch.empty((n, 2), dtype=torch.int32, device=dev)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t_wall = time.perf_counter()
        with rtx.KernelTimer() as kt:
            for a, b in evs:
                a.record(stream)
                fn()
                b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_wall
        per = [a.elapsed_time(b) for a, b in evs]
        kernel_ms = kt.ms / kt.launches if kt.launches else float("nan")
        return float(np.mean(per)), wall, kernel_ms

    # Launch shapes for the headline: the project autotuner (coordinate descent over LDS
    # stack depth, loop structure, ray sorting, leaf-phase wait and persistent grid size).
    # The headline walks the binary tree (trace_tree 0): the reference's visiting order, so
    # hit ids equal the reference's bit for bit, exact-t ties included; the 4-wide layout is
    # timed beside it below.
    headline_keys = [k for k in autotune.TRACE_KEYS if k != "trace_tree"]
    rtx.set_tuning(trace_tree=0)
    trace_tuning, trace_tune_s = rtx.get_tuning(), 0.0
    if not args.no_autotune:
        t_tune = time.perf_counter()
        trace_tuning, _, _ = autotune.tune(lambda: (lambda: bvh.trace_closest_device(inc, hits, stream=stream)),
                                           keys=headline_keys, reps=3)
        trace_tune_s = time.perf_counter() - t_tune
    if args.tuning:
        trace_tuning = dict(trace_tuning, **json.loads(args.tuning))
    rtx.set_tuning(**trace_tuning)

    ms_inc, wall, kms_inc = timed(lambda: bvh.trace_closest_device(inc, hits, stream=stream), args.steps,
                                  args.warmup)
    ms_coh = ms_any = float("nan")
    coh_tuning = any_tuning = trace_tuning
    host = None
    if not args.skip_extras:
        # host buffers through the C ABI (PCIe-inclusive): rays in, hits out
        inc_h = inc.cpu().numpy()
        bvh.trace(inc_h)
        t_h = time.perf_counter()
        bvh.trace(inc_h)
        host_s = time.perf_counter() - t_h
        one = inc_h[:1].copy()
        for _ in range(20):
            bvh.trace(one)
        t_h = time.perf_counter()
        for _ in range(200):
            bvh.trace(one)
        host = {"pcie_inclusive_mrays_s": round(n / host_s / 1e6, 2),
                "single_ray_call_us": round((time.perf_counter() - t_h) / 200 * 1e6, 1)}
        # one call per ray over 1,000 different rays, and calls of 256 rays (calls of at most
        # 16,384 rays take the one-wave-per-ray kernel)
        t_h = time.perf_counter()
        for i in range(1000):
            bvh.trace(inc_h[i:i + 1])
        host["single_ray_call_us_1000_rays"] = round((time.perf_counter() - t_h) / 1000 * 1e6, 1)
        t_h = time.perf_counter()
        for i in range(50):
            bvh.trace(inc_h[256 * i:256 * (i + 1)])
        host["call_256_rays_us"] = round((time.perf_counter() - t_h) / 50 * 1e6, 1)
        # the coherent and any-hit launches get launch shapes of their own (sorting rays
        # that are already in camera order only costs the sort, for instance)
        if not args.no_autotune:
            coh_tuning, _, _ = autotune.tune(lambda: (lambda: bvh.trace_closest_device(coh, hits, stream=stream)),
                                             keys=headline_keys, reps=3)
            rtx.set_tuning(**trace_tuning)
            # any-hit answers do not depend on the tree walked, so the layout is tuned too
            any_tuning, _, _ = autotune.tune(lambda: (lambda: bvh.trace_any_device(inc, occ, stream=stream)),
                                             keys=autotune.TRACE_KEYS, reps=3)
        rtx.set_tuning(**coh_tuning)
        ms_coh, _, _ = timed(lambda: bvh.trace_closest_device(coh, hits, stream=stream), max(3, args.steps // 4), 1)
        rtx.set_tuning(**any_tuning)
        ms_any, _, _ = timed(lambda: bvh.trace_any_device(inc, occ, stream=stream), max(3, args.steps // 4), 1)
        rtx.set_tuning(**trace_tuning)

    # Algorithmic bytes (the per-ray model: 32 B ray + 16 B hit + 64 B per internal node
    # visited + 48 B per triangle tested), counted on the binary walk's visit sequence.
    bvh.trace_counts_device(inc, counts, stream=stream)
    torch.cuda.synchronize()
    c = counts.to(torch.int64).sum(0).cpu().numpy()
    visits, tests = int(c[0]), int(c[1])
    alg_bytes = 48 * n + 64 * visits + 48 * tests
    # The 4-wide layout with the same launch shapes: its rate, and its hits against the
    # headline's (t bit-identical for every ray, triangle ids equal except at exact-t ties).
    wide = None
    if not args.skip_extras:
        bvh.trace_closest_device(inc, hits, stream=stream)
        hits_w = torch.empty_like(hits)
        rtx.set_tuning(**dict(trace_tuning, trace_tree=1))
        ms_w, _, kms_w = timed(lambda: bvh.trace_closest_device(inc, hits_w, stream=stream), max(5, args.steps // 2), 2)
        rtx.set_tuning(**trace_tuning)
        torch.cuda.synchronize()
        hb, hw = hits.view(torch.int32), hits_w.view(torch.int32)
        diff_id = hb[:, 3] != hw[:, 3]
        wide = {"mrays_s": round(n / (ms_w * 1e-3) / 1e6, 2), "kernel_ms": round(kms_w, 4),
                "vs_headline": {"rays": n, "t_bit_identical": bool(torch.equal(hb[:, 0], hw[:, 0])),
                                "id_differences": int(diff_id.sum()),
                                "id_differences_at_exact_t_ties": int((diff_id & (hb[:, 0] == hw[:, 0])).sum())}}
        del hits_w
    bvh.trace_closest_device(inc, hits, stream=stream)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(bvh, v, t, inc[: 1 << 21].cpu().numpy(), coh[: 1 << 20].cpu().numpy(), args.cpu_seconds,
                           gpu_hits=hits[: 1 << 21].cpu().numpy())

    # HBM traffic, L2 hit rate and VALU issue of the headline kernel from rocprofv3 counters
    # of this same launch (child processes after the timed region), priced on its live
    # kernel time above
    counters, roofline = {}, None
    if not args.no_counters:
        # the C3 leg runs with the launch shapes its own autotuner picks, below
        for leg in [x for x in args.counter_legs.split(",") if x and x != "c3"]:
            tun = trace_tuning if leg == "trace" else None
            counters[leg] = measure_counters(leg, tun, args)
        roofline = trace_roofline(counters.get("trace"), kms_inc)

    render = None
    if not (args.no_render or args.skip_extras):
        del inc, coh, hits, occ, counts
        torch.cuda.empty_cache()
        # the batched leg tunes launch shapes for its own mixed-scene launches; the other
        # render legs run with the launch shapes that were set before it
        before = rtx.get_tuning()
        batched = bench_batched(stream, args.batched_scenes, tune=not args.no_autotune)
        rtx.set_tuning(**before)
        if not args.no_counters and "c3" in args.counter_legs.split(","):
            counters["c3"] = measure_counters("c3", batched["tuning"], args)
            batched["counters_plain_vs_batched"] = c3_comparison(counters["c3"])
        render = {"differentiable": bench_differentiable(stream), "batched": batched,
                  "cornell": bench_cornell(stream, cpu=not args.no_cpu, tune=not args.no_autotune)}
        if "c4" in counters:
            render["differentiable"]["roofline"] = render_roofline(counters["c4"], "k_wtrace_persist")
        if not args.no_inverse:
            render["inverse_terrain"] = bench_inverse(stream, steps=args.inverse_steps)
            if "c5" in counters:
                render["inverse_terrain"]["roofline"] = render_roofline(counters["c5"], "k_wtrace_persist")

    mrays = n / (ms_inc * 1e-3) / 1e6
    out = {
        "metric": "Mrays/s incoherent closest-hit (batched vs plain); fwd+adjoint ms/iter",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_inc, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "BVH build + traversal microbenchmark: 1M random fp32 triangles (seed 1), "
                               "4,194,304 incoherent rays, closest hit",
                   "triangles": args.tris, "rays": n, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": {
            "wide_tree": wide,
            "algorithmic_model": {"bytes_per_launch": alg_bytes,
                                  "rate_gbs": round(alg_bytes / (kms_inc * 1e-3) / 1e9, 1),
                                  "note": "counts every node and triangle read as memory traffic, cache hits "
                                          "included; not an HBM figure"},
            "counters": counters or None,
            "coherent_closest_mrays_s": None if args.skip_extras else round(n / (ms_coh * 1e-3) / 1e6, 2),
            "host_buffers": host,
            "incoherent_any_hit_mrays_s": None if args.skip_extras else round(n / (ms_any * 1e-3) / 1e6, 2),
            "build_ms": round(build_ms, 2),
            "build_mtris_s": round(args.tris / build_ms / 1e3, 3),
            "tree": {k: info[k] for k in ("node_count", "tri_count", "max_depth", "levels")},
            "avg_internal_visits_per_ray": round(visits / n, 2),
            "avg_triangle_tests_per_ray": round(tests / n, 2),
            "trace_kernel_ms": round(kms_inc, 4),
            "ray_sort_and_launch_ms": round(ms_inc - kms_inc, 4),
            "tuning": trace_tuning,
            "coherent_tuning": {k: v for k, v in coh_tuning.items() if k in autotune.TRACE_KEYS},
            "any_hit_tuning": {k: v for k, v in any_tuning.items() if k in autotune.TRACE_KEYS},
            "autotune": None if args.no_autotune else {"seconds": round(trace_tune_s, 2)},
            "timed_wall_s": round(wall, 3),
        },
        "render": render,
    }
    print(json.dumps(out), flush=True)


def _cuda_ms(fn, reps, stream):
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_differentiable(stream, reps=6):
    """Differentiable render: Cornell box + 4096-triangle displaced sphere, per-triangle
    albedo, 2-triangle emitter; 512x512, 32 spp, max_bounce 4, no Russian roulette (the
    adjoint needs a path set independent of albedo). Loss = sum (I - T)^2."""
    import numpy as np
    import torch

    import rtx
    from rtx import scenes as S
    v, t, albedo_true, emission, cam = S.diff_scene(seed=3)
    g = rtx.build(v, t)
    p = rtx.make_params(512, 512, 32, 4, 99, seed=3)
    e_t = torch.from_numpy(emission).cuda()
    target = rtx.render(g, cam, p, torch.from_numpy(albedo_true).cuda(), e_t, stream=stream)
    rng = np.random.default_rng(0)
    albedo = torch.from_numpy(np.clip(albedo_true + rng.normal(0, 0.1, albedo_true.shape), 0.05, 0.95)
                              .astype(np.float32)).cuda()
    img = torch.empty_like(target)

    def fwd():
        rtx.render(g, cam, p, albedo, e_t, out=img, stream=stream)

    def fwd_adj():
        a = albedo.detach().clone().requires_grad_(True)
        e = e_t.detach().clone().requires_grad_(True)
        loss = ((rtx.differentiable_render(a, e, g, cam, p) - target) ** 2).sum()
        loss.backward()
    fwd()
    fwd_adj()
    ms_f = _cuda_ms(fwd, reps, stream)
    ms_fa = _cuda_ms(fwd_adj, reps, stream)
    return {"workload": "512x512, 32 spp, max_bounce 4, 4118 triangles", "forward_ms": round(ms_f, 3),
            "forward_adjoint_ms_per_iter": round(ms_fa, 3), "ratio": round(ms_fa / ms_f, 3),
            "paths": 512 * 512 * 32}


def bench_inverse(stream, steps=3):
    """Terrain inverse render: 2048x2048-vertex fBm heightfield (8,380,418 triangles, seed 4),
    sun + constant sky, 1920x1080, 64 spp, max_bounce 8, no roulette; the 512x512 albedo
    texture starts at 0.5 and Adam (lr 1e-2) fits it to the image of the seed-4 texture.
    Reports the GPU build, one forward render, and the time per Adam iteration (forward +
    path-replay adjoint + update), with the losses of the first iterations."""
    import numpy as np
    import torch

    import rtx
    from rtx import scenes as S
    v, t, uv = S.terrain(2048, seed=4)
    vd = torch.from_numpy(v).cuda()
    td = torch.from_numpy(t.view(np.int32)).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = rtx.build_device(vd, td, stream=stream)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    del vd, td
    cam, sky, sun = S.terrain_setup()
    uv_t = torch.from_numpy(uv).cuda()
    tex_true = torch.from_numpy(S.albedo_texture(512, seed=4)).cuda()
    p = rtx.make_params(1920, 1080, 64, 8, 99, seed=4, env_mode=3, env_color=sky, uv=uv_t, tex=tex_true, sun=sun)
    target = rtx.render(g, cam, p, stream=stream).clone()
    x = torch.full_like(tex_true, 0.5).requires_grad_(True)
    opt = torch.optim.Adam([x], lr=1e-2)
    img = torch.empty_like(target)

    def forward():
        rtx.set_texture(p, uv_t, x.detach())
        rtx.render(g, cam, p, out=img, stream=stream)
    fwd_ms = _cuda_ms(forward, 1, stream)
    def tex_rmse():
        return round(float(((x.detach() - tex_true) ** 2).mean().sqrt()), 5)
    rmse0 = tex_rmse()
    losses, step_ms = [], []
    for it in range(steps):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        opt.zero_grad()
        loss = ((rtx.differentiable_render_texture(x, uv_t, g, cam, p) - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            x.clamp_(1e-3, 1.0)   # albedo stays positive: the replay gradient needs it
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - t1) * 1e3)
        losses.append(round(float(loss.detach()), 2))
        if (it + 1) % 10 == 0:
            print(f"[inverse] step {it + 1}/{steps} loss {losses[-1]} {step_ms[-1]:.0f} ms", file=sys.stderr, flush=True)
    shown = losses if steps <= 10 else {str(i + 1): losses[i] for i in range(steps) if i < 3 or (i + 1) % 20 == 0}
    return {"workload": "2048^2 fBm terrain (8,380,418 triangles), 1920x1080, 64 spp, max_bounce 8, "
                        "512x512 texture, Adam lr 1e-2",
            "paths": 1920 * 1080 * 64, "build_ms": round(build_ms, 1), "forward_ms": round(fwd_ms, 1),
            "adam_iteration_ms": round(min(step_ms[1:] or step_ms), 1), "adam_steps": steps,
            "total_fit_s": round(sum(step_ms) / 1e3, 1), "losses": shown,
            "texture_rmse": [rmse0, tex_rmse()]}


def bench_cornell(stream, cpu=True, reps=3, tune=True):
    """Cornell box: open box (20 wall + 2 light + 1280 icosphere triangles), 256x256,
    16 spp, max_bounce 5, roulette from bounce 3; GPU render time, and the oracle's
    render of the same image on the host at 1 thread and at all cores. The renderer's
    launch shapes are autotuned for this small (1M-path) launch, in two rounds so that the
    loop structure is revisited after the grid changes, and restored after it."""
    import torch

    import rtx
    from rtx import scenes as S
    v, t, albedo, emission, cam = S.cornell_box(seed=0)
    g = rtx.build(v, t)
    p = rtx.make_params(256, 256, 16, 5, 3, seed=1)
    a_t, e_t = torch.from_numpy(albedo).cuda(), torch.from_numpy(emission).cuda()
    img = torch.empty((256, 256, 3), device="cuda")
    saved = rtx.get_tuning()
    tuning = saved
    try:
        if tune:
            from rtx import autotune
            tuning, _, _ = autotune.tune(lambda: (lambda: rtx.render(g, cam, p, a_t, e_t, out=img, stream=stream)),
                                         keys=autotune.RENDER_KEYS, reps=2, rounds=2)
            rtx.set_tuning(**tuning)
        ms = _cuda_ms(lambda: rtx.render(g, cam, p, a_t, e_t, out=img, stream=stream), reps, stream)
    finally:
        rtx.set_tuning(**saved)
    out = {"workload": "256x256, 16 spp, max_bounce 5, roulette from bounce 3, 1302 triangles",
           "paths": 256 * 256 * 16, "gpu_ms": round(ms, 3),
           "tuning": {k: tuning[k] for k in ("render_loop", "render_leaf_coop", "grid_blocks", "shade_threads",
                                            "batch_paths", "render_lds_depth", "render_wait_lanes")}}
    if cpu:
        from oracle import oracle as O
        o = O.build(v, t)
        op = O.make_params(256, 256, 16, 5, 3, seed=1)
        threads = min(16, os.cpu_count() or 1)
        for name, nt in (("cpu_1thread_s", 1), ("cpu_allcore_s", threads)):
            t0 = time.perf_counter()
            O.render(o, cam, op, albedo, emission, nthreads=nt)
            out[name] = round(time.perf_counter() - t0, 3)
        out["cpu_threads"] = threads
        out["cpu_kind"] = "port"
    return out


def bench_batched(stream, nscenes, reps=2, tune=True):
    """Batched wavefront render of fBm-displaced spheres of uneven size (seed 2): one
    launch sequence per scene (plain) vs all scenes in shared launches replayed from one
    hipGraph (batched). The queue size (paths per wavefront batch), block sizes, grid and
    stack depth are first chosen by the autotuner on the batched launch; both sides are
    then timed with that tuning."""
    import torch

    import rtx
    from rtx import autotune
    from rtx import scenes as S
    sc = S.batched_scenes(nscenes, seed=2)
    bvhs = [rtx.build(s["verts"], s["tris"]) for s in sc]
    params = [rtx.make_params(s["width"], s["height"], s["spp"], s["max_bounce"], 3, seed=2 + i)
              for i, s in enumerate(sc)]
    imgs = [torch.empty((s["height"], s["width"], 3), device="cuda") for s in sc]
    paths = sum(s["width"] * s["height"] * s["spp"] for s in sc)
    out = {"workload": f"{nscenes} scenes, {sum(len(s['tris']) for s in sc)} triangles, {paths} paths, "
                       "max_bounce 6", "paths": paths}
    if tune:
        cams = [s["cam"] for s in sc]

        def make_step():
            b = rtx.Batch(bvhs, cams, params, [None] * nscenes, [None] * nscenes, imgs, 1)
            return lambda: b.launch(stream)
        t0 = time.perf_counter()
        best, _, hist = autotune.tune(make_step, keys=autotune.RENDER_KEYS, reps=2)
        rtx.set_tuning(**best)
        out["autotune"] = {"seconds": round(time.perf_counter() - t0, 2), "candidates": len(hist)}
    out["tuning"] = rtx.get_tuning()
    for mode, name in ((0, "plain"), (1, "batched_graph")):
        b = rtx.Batch(bvhs, [s["cam"] for s in sc], params, [None] * nscenes, [None] * nscenes, imgs, mode)
        ms = _cuda_ms(lambda: b.launch(stream), reps, stream)
        out[name + "_ms"] = round(ms, 2)
        out[name + "_mpaths_s"] = round(paths / ms / 1e3, 2)
        del b
    out["speedup"] = round(out["plain_ms"] / out["batched_graph_ms"], 3)
    return out


def counter_child(args):
    """One leg of the workload, run by measure_counters under rocprofv3: the 1 GiB device copy
    that calibrates the HBM counters, then the leg's kernels with the given launch shapes."""
    import numpy as np
    import torch

    import rtx
    from rtx import scenes as S
    dev = torch.device("cuda:0")
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    if args.tuning:
        rtx.set_tuning(**json.loads(args.tuning))
    leg = args.counter_child
    if leg == "trace":
        v, t = S.random_triangles(args.tris, seed=1)
        bvh = rtx.build_device(torch.from_numpy(v).to(dev), torch.from_numpy(t.view(np.int32)).to(dev))
        inc = torch.from_numpy(S.incoherent_rays(args.rays, seed=1)).to(dev)
        hits = torch.empty((args.rays, 4), dtype=torch.float32, device=dev)
        for _ in range(14):   # the median dispatch is past the first, cold-cache launches
            bvh.trace_closest_device(inc, hits)
    elif leg == "c4":
        v, t, albedo_true, emission, cam = S.diff_scene(seed=3)
        g = rtx.build(v, t)
        p = rtx.make_params(512, 512, 32, 4, 99, seed=3)
        a_t, e_t = torch.from_numpy(albedo_true).cuda(), torch.from_numpy(emission).cuda()
        target = rtx.render(g, cam, p, a_t, e_t)
        for _ in range(2):
            a = (a_t * 0.9).requires_grad_(True)
            e = e_t.detach().clone().requires_grad_(True)
            ((rtx.differentiable_render(a, e, g, cam, p) - target) ** 2).sum().backward()
    elif leg == "c3":
        sc = S.batched_scenes(64, seed=2)
        bvhs = [rtx.build(s["verts"], s["tris"]) for s in sc]
        params = [rtx.make_params(s["width"], s["height"], s["spp"], s["max_bounce"], 3, seed=2 + i)
                  for i, s in enumerate(sc)]
        imgs = [torch.empty((s["height"], s["width"], 3), device="cuda") for s in sc]
        for mode in (0, 1):   # plain, then the graph-batched launch
            b = rtx.Batch(bvhs, [s["cam"] for s in sc], params, [None] * 64, [None] * 64, imgs, mode)
            b.launch()
            torch.cuda.synchronize()
            del b
    elif leg == "c5":
        v, t, uv = S.terrain(2048, seed=4)
        g = rtx.build_device(torch.from_numpy(v).cuda(), torch.from_numpy(t.view(np.int32)).cuda())
        cam, sky, sun = S.terrain_setup()
        uv_t = torch.from_numpy(uv).cuda()
        tex_true = torch.from_numpy(S.albedo_texture(512, seed=4)).cuda()
        p = rtx.make_params(1920, 1080, 64, 8, 99, seed=4, env_mode=3, env_color=sky, uv=uv_t, tex=tex_true, sun=sun)
        target = rtx.render(g, cam, p).clone()
        x = torch.full_like(tex_true, 0.5).requires_grad_(True)
        ((rtx.differentiable_render_texture(x, uv_t, g, cam, p) - target) ** 2).sum().backward()
    else:
        raise SystemExit(f"unknown counter leg {leg}")
    torch.cuda.synchronize()


def _kernel_family(name):
    """Kernel name without namespaces and parameter list: k_trace_persist<false, false, 12, 2, 1>."""
    n = name.replace("rtx::(anonymous namespace)::", "").replace("rtx::", "").replace("void ", "")
    depth, end = 0, len(n)
    for i, ch in enumerate(n):   # the parameter list starts at the first '(' outside <...>
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            end = i
            break
    return n[:end].strip()


def measure_counters(leg, tuning, args):
    """Runs `bench.py --counter-child leg` under rocprofv3: three --pmc passes (PMC_PASSES)
    and one --kernel-trace pass, each in its own process (counters never combined with a
    trace domain), then summarize_counters. Per kernel: calibrated HBM bytes read and
    written (FETCH_SIZE and WRITE_SIZE, each scaled by what the same counter reads for the
    1 GiB copy that starts the child) and VALU instructions of the median dispatch, its
    duration, and over all dispatches the HBM rate, L2 hit rate, VALU issue rate, lanes
    active and the fraction of wave time spent waiting."""
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    rp = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    base = os.path.abspath(args.counters_dir) if args.counters_dir else tempfile.mkdtemp(prefix="rtx_counters_")
    child = [sys.executable, os.path.abspath(__file__), "--counter-child", leg, "--tris", str(args.tris),
             "--rays", str(args.rays)]
    if tuning:
        child += ["--tuning", json.dumps(tuning)]
    env = dict(os.environ, TMPDIR="/tmp")

    def run(opts, tag):
        d = os.path.join(base, f"{leg}_{tag}")
        cmd = [rp, *opts, "-d", d, "-o", "run", "--output-format", "csv", "--", *child]
        pr = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                              start_new_session=True)
        try:
            out, _ = pr.communicate(timeout=240)
        except subprocess.TimeoutExpired:
            os.killpg(pr.pid, signal.SIGKILL)
            pr.communicate()
            raise RuntimeError(f"rocprofv3 {tag} pass of {leg} timed out")
        files = glob.glob(os.path.join(d, "**", "*.csv"), recursive=True)
        # rocprofv3 may end with a segmentation fault after the program has finished and
        # its results are complete
        if pr.returncode not in (0, 139, -11) or not files:
            raise RuntimeError(f"rocprofv3 {tag} pass of {leg} failed ({pr.returncode}): "
                               f"{out.decode(errors='replace')[-400:]}")
        return d

    try:
        for k, cs in enumerate(PMC_PASSES):
            run(["--pmc", *cs], f"pmc{k}")
        run(["--kernel-trace"], "trace")
    except (RuntimeError, OSError) as exc:
        return {"error": str(exc)}
    return summarize_counters(base, leg)


def summarize_counters(base, leg):
    """Reads the rocprofv3 output of measure_counters (directories <leg>_pmc<k> and
    <leg>_trace under `base`) into the per-kernel summary it returns."""
    import csv
    import glob

    import numpy as np
    per, calib = {}, {}
    for k in range(len(PMC_PASSES)):
        d = os.path.join(base, f"{leg}_pmc{k}")
        # per dispatch: a counter's value summed over the rows of that dispatch
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                fam, cn, val = _kernel_family(r["Kernel_Name"]), r["Counter_Name"], float(r["Counter_Value"])
                if "copyBuffer" in fam and cn in ("FETCH_SIZE", "WRITE_SIZE"):
                    calib[cn] = max(calib.get(cn, 0.0), val)
                    continue
                dv = vals.setdefault(fam, {}).setdefault(cn, {})
                key = r.get("Dispatch_Id", str(len(dv)))
                dv[key] = dv.get(key, 0.0) + val
        for fam, cv in vals.items():
            e = per.setdefault(fam, {})
            for cn, dv in cv.items():
                e[cn] = list(dv.values())
    for f in glob.glob(os.path.join(base, f"{leg}_trace", "**", "*kernel_trace.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            fam = _kernel_family(r["Kernel_Name"])
            if fam in per:
                per[fam].setdefault("_ns", []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    if "FETCH_SIZE" not in calib or "WRITE_SIZE" not in calib:
        return {"error": "no calibration copy in the counter output"}
    ff, fw = 1048576.0 / calib["FETCH_SIZE"], 1048576.0 / calib["WRITE_SIZE"]
    out = {"calibration": {"copy_kib": 1048576, "fetch_factor": round(ff, 4), "write_factor": round(fw, 4)},
           "kernels": {}}
    # Per kernel: the median dispatch (the steady state of a repeated launch, not its first,
    # cold-cache dispatch: per-launch bytes and instructions, duration), and rates over all
    # of its dispatches together (total bytes or instructions / total duration: the figure
    # for kernels launched at many sizes, such as one launch per bounce).
    for fam, e in per.items():
        ns = e.get("_ns", [])
        if not ns or "FETCH_SIZE" not in e or "SQ_INSTS_VALU" not in e:
            continue
        med = lambda cn: float(np.median(e[cn])) if cn in e else 0.0
        tot = lambda cn: float(np.sum(e[cn])) if cn in e else 0.0
        us, tot_s = float(np.median(ns)) / 1e3, float(np.sum(ns)) * 1e-9
        fetch, write = med("FETCH_SIZE") * 1024 * ff, med("WRITE_SIZE") * 1024 * fw
        tbytes = (tot("FETCH_SIZE") * ff + tot("WRITE_SIZE") * fw) * 1024
        hit, miss = tot("TCC_HIT_sum"), tot("TCC_MISS_sum")
        out["kernels"][fam] = {
            "dispatches": len(ns), "avg_us": round(us, 2), "total_us": round(tot_s * 1e6, 1),
            "hbm_read_bytes": round(fetch), "hbm_write_bytes": round(write),
            "hbm_gbs": round(tbytes / tot_s / 1e9, 1),
            "l2_hit_rate": round(hit / (hit + miss), 4) if hit + miss else None,
            "valu_insts": round(med("SQ_INSTS_VALU")),
            "valu_issue_frac": round(tot("SQ_INSTS_VALU") / tot_s / VALU_PEAK_PER_S, 4),
            "lanes_active": round(tot("SQ_THREAD_CYCLES_VALU") / (64 * tot("SQ_ACTIVE_INST_VALU")), 3)
            if tot("SQ_ACTIVE_INST_VALU") else None,
            "wave_wait_frac": round(tot("SQ_WAIT_ANY") / tot("SQ_WAVE_CYCLES"), 3) if tot("SQ_WAVE_CYCLES") else None,
            "lds_insts": round(med("SQ_INSTS_LDS")), "vmem_rd_insts": round(med("SQ_INSTS_VMEM_RD")),
            "vmem_wr_insts": round(med("SQ_INSTS_VMEM_WR"))}
    # the busiest kernels first
    out["kernels"] = dict(sorted(out["kernels"].items(), key=lambda kv: -kv[1]["total_us"])[:12])
    return out


def trace_roofline(ctr, kernel_ms):
    """Roofline of the headline traversal kernel: HBM bytes per launch from the counters,
    priced on the kernel's live HIP-event duration of the timed launches; the VALU issue
    rate beside it (the resource that binds this kernel)."""
    if not ctr or "kernels" not in ctr:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None,
                "error": (ctr or {}).get("error", "counters not measured")}
    fam = [k for k in ctr["kernels"] if k.startswith("k_trace_persist") or k.startswith("k_trace<")]
    if not fam:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None,
                "error": "traversal kernel not in the counter output"}
    k = ctr["kernels"][fam[0]]
    traffic = k["hbm_read_bytes"] + k["hbm_write_bytes"]
    gbs = traffic / (kernel_ms * 1e-3) / 1e9
    valu = k["valu_insts"] / (kernel_ms * 1e-3)
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": fam[0],
            "kernel_ms": round(kernel_ms, 4), "limiter": "valu_issue",
            "valu_issue": {"achieved": round(valu / 1e9, 1), "peak": round(VALU_PEAK_PER_S / 1e9, 1),
                           "unit": "G wave64 VALU instructions/s", "frac": round(valu / VALU_PEAK_PER_S, 4)},
            "l2_hit_rate": k["l2_hit_rate"], "lanes_active": k["lanes_active"],
            "profiled_kernel_us": k["avg_us"]}


def render_roofline(ctr, prefix):
    """Roofline of a render leg's dominant kernel (the first kernel family named `prefix`),
    all figures from its counter and kernel-trace passes: rates over all of its dispatches
    (one per bounce and batch, of falling sizes); `traffic` is the median dispatch's."""
    if not ctr or "kernels" not in ctr:
        return None
    fam = [k for k in ctr["kernels"] if k.startswith(prefix)]
    if not fam:
        return None
    k = ctr["kernels"][fam[0]]
    traffic = k["hbm_read_bytes"] + k["hbm_write_bytes"]
    return {"bound": "hbm", "achieved": k["hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(k["hbm_gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": fam[0],
            "kernel_us": k["avg_us"], "dispatches": k["dispatches"], "total_us": k["total_us"],
            "limiter": "valu_issue" if k["valu_issue_frac"] > k["hbm_gbs"] / HBM_PEAK_GBS else "hbm",
            "valu_issue": {"frac": k["valu_issue_frac"]}, "l2_hit_rate": k["l2_hit_rate"],
            "lanes_active": k["lanes_active"]}


def c3_comparison(ctr):
    """The C3 counter leg's traversal and shade kernels, plain (one scene per launch
    sequence: scene form 0, the last template argument of k_wtrace, and shading with a
    uniform descriptor, the second of k_shade) against batched (mixed-scene queues: scene
    form 1, per-lane descriptors): time over all dispatches, VALU issue, lanes active and
    HBM rate of each."""
    if not ctr or "kernels" not in ctr:
        return None
    out = {}
    for fam, k in ctr["kernels"].items():
        args = [a.strip() for a in fam[fam.find("<") + 1:fam.rfind(">")].split(",")]
        if fam.startswith("k_wtrace") and args:
            side = "plain" if args[-1] == "0" else "batched"
        elif fam.startswith("k_shade") and len(args) > 1:
            side = "plain" if args[1] == "true" else "batched"
        else:
            continue
        out.setdefault(side, {})[fam] = {x: k[x] for x in ("dispatches", "total_us", "valu_issue_frac", "lanes_active",
                                                           "hbm_gbs", "l2_hit_rate", "wave_wait_frac")}
    return out


def cpu_baseline(bvh, verts, tris, rays, coh_rays, seconds, gpu_hits=None):
    """The compiled reference's trace() (single thread, as the reference runs) on a bounded
    sample of the same incoherent rays over the same tree; falls back to the oracle port.
    Beside it: the all-core port, a coherent-ray sample, and the reference's serial build
    of the same 1M triangles."""
    import numpy as np
    from oracle import oracle as O

    nodes, rtris, prims = bvh.export()
    tree = O.Tree(nodes, rtris, prims, verts)
    kind = "reference" if O.ref_available() else "port"

    ids = []

    def rate(rs, budget, keep=False):
        if kind == "reference":
            O.ref_trace(tree, rs[:1])   # load the tree into the reference tracer once
        done, t0 = 0, time.perf_counter()
        chunk = 2048
        while time.perf_counter() - t0 < budget and done < len(rs):
            r = rs[done: done + chunk]
            h = O.ref_trace(None, r) if kind == "reference" else O.trace(tree, r, nthreads=1)
            if keep:
                ids.append(np.stack([h["t"].view(np.uint32), h["tri"]], 1))
            done += len(r)
        el = time.perf_counter() - t0
        return done, el
    done, el = rate(rays, seconds, keep=gpu_hits is not None)
    match = None
    if gpu_hits is not None:
        # the timed GPU launch's hits against the CPU reference's on the traced sample
        ref = np.concatenate(ids)
        g = gpu_hits[:len(ref)].view(np.uint32)
        same_t = ref[:, 0] == g[:, 0]
        diff = ref[:, 1] != g[:, 3]
        match = {"rays": int(len(ref)), "t_bit_identical": bool(same_t.all()), "id_differences": int(diff.sum()),
                 "id_differences_at_exact_t_ties": int((diff & same_t).sum())}
    value = done / el / 1e6
    cdone, cel = rate(coh_rays, seconds / 4)
    cores = os.cpu_count() or 1
    threads = min(16, cores)
    t1 = time.perf_counter()
    m = min(len(rays), 1 << 18)
    O.trace(tree, rays[:m], nthreads=threads)
    allcore = m / (time.perf_counter() - t1) / 1e6
    build_s = None
    if kind == "reference":
        t2 = time.perf_counter()
        O.ref_build(verts, tris)
        build_s = round(time.perf_counter() - t2, 3)
    return {"value": round(value, 4), "unit": "Mrays/s", "cores": 1, "kind": kind,
            "sample": f"{done} of the 4,194,304 incoherent rays, same 1M-triangle tree, single thread "
                      f"({el:.1f} s)",
            "coherent_mrays_s": round(cdone / cel / 1e6, 4), "coherent_sample": f"{cdone} coherent rays ({cel:.1f} s)",
            "allcore_port_mrays_s": round(allcore, 4), "allcore_threads": threads, "host_cores": cores,
            "build_1m_serial_s": build_s, "cpu_model": cpu_model(), "gpu_hits_vs_cpu": match}


if __name__ == "__main__":
    main()
