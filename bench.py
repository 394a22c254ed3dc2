"""Benchmark: GPU wavefront PT-NEE-MIS on the procedural ~2.07M-triangle terrain
scene (1024x512 fBm heightfield + 200 icospheres, 1920x1080, max depth 12).

One step = one 16-spp pass over the whole 1920x1080 frame (spp_axis 4); the
default 16 steps accumulate the workload's 256 spp.  Inputs (scene, BVH) are
resident in HBM before timing starts; cpt_render returns after the device has
finished (it synchronises its stream).

Prints one JSON line.  `roofline` describes the dominant kernel (closest-hit
traversal, `extend`): algorithmic bytes per extension ray = 32 (ray) + 16 (hit)
+ 128 x BVH4 nodes visited + 48 x triangles tested, counted by a counting run of
the same kernel on the same rays, divided by that kernel's summed launch time
measured with HIP events on the stream it runs on.  `cpu_baseline` renders a
row sample of the same workload with the CPU oracle (a restatement of the
reference's PathtracerMIS with its mt19937 column sampler and binary BVH).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cpp-pt_amd"))
sys.path.insert(0, ROOT)

METRIC = "samples/s + Mrays/s @1080p; equal-time relMSE guided vs plain PT (1 GPU)"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-axis", type=int, default=4, help="per-axis samples of one step (4 -> 16 spp)")
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spheres", type=int, default=200)
    ap.add_argument("--grid", type=str, default="1024x512")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--grid-blocks", type=int, default=0)
    ap.add_argument("--max-paths", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU baseline sample duration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip cpt_autotune (default launch configuration)")
    ap.add_argument("--gpu-bvh", action="store_true", help="build the BVH on the device (CPT_SCENE_GPU_BVH)")
    ap.add_argument("--workload", choices=["terrain", "views", "guided", "bvh", "box"], default="terrain",
                    help="terrain: the headline 1080p line; views: 64-view batch of the box room; "
                         "guided: equal-time relMSE of guided vs plain PT on the hard-lighting interior")
    ap.add_argument("--budgets", type=str, default="10,30,60", help="guided: equal-time budgets in seconds")
    ap.add_argument("--ref-spp-axis", type=int, default=64, help="guided: reference spp per axis (64 -> 4096 spp)")
    ap.add_argument("--boxes", type=int, default=40000, help="guided: furniture boxes (48 triangles each)")
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--no-ppg", action="store_true", help="guided: skip the SD-tree (PPG) arm")
    ap.add_argument("--fp8", action="store_true", help="guided: run the guide network in FP8 (training stays BF16)")
    ap.add_argument("--train-paths", type=int, default=0,
                    help="guided: paths per pass that record training samples (0: min(pixels, 2^18))")
    ap.add_argument("--guide-width", type=int, choices=[64, 128], default=64,
                    help="guided: hidden width of the guide network (64: 8 lobes, 128: 16 lobes)")
    ap.add_argument("--guide-depth", type=int, default=0,
                    help="guided: guide the vertices below this depth only (0: every vertex)")
    ap.add_argument("--scene", choices=["interior", "grid"], default="interior",
                    help="guided: the hard-lighting interior, or the 4x4 grid of interiors (7.7M triangles, "
                         "32 emitters, mixed GGX; device BVH build)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (scales HBM counters)")
    return ap.parse_args()


def copy_calibration(mib):
    """Plain copy kernel of known size: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    it reads and writes mib MiB per dispatch, which scales the counters."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "checksum": float(y[:16].sum())}))


def bench_guided(args):
    """Equal-time comparison on the hard-lighting interior (one 0.2x0.2 light outside
    a cluttered room, light entering through a 0.3-wide door gap): plain PT-NEE-MIS
    vs guided PT (learned MLP guide trained online, training time included), both
    accumulated one sample per pixel at a time; relMSE = mean((I-R)^2/(R^2+0.01))
    against an FP64-accumulated plain reference."""
    import numpy as np

    import cpt

    W, H = args.width, args.height
    budgets = [float(b) for b in args.budgets.split(",")]
    ctx = cpt.Context(0)
    if args.scene == "grid":
        scene = cpt.scenes.interior_grid_scene(res=(W, H))
        t0 = time.perf_counter()
        # the device build keeps set-up short at 7.7M triangles; the tie-rank table only
        # decides exact-t ties between coincident triangles, which this scene has none of
        ctx.upload(scene, gpu_bvh=True, tie_rank=False)
        print(f"bench: grid scene {scene.n_triangles} triangles uploaded in {time.perf_counter() - t0:.1f} s "
              f"(BVH {ctx.bvh_info()['build_ms']:.0f} ms on the device)", file=sys.stderr, flush=True)
    else:
        scene = cpt.scenes.interior_scene(res=(W, H), n_boxes=args.boxes)
        ctx.upload(scene)
    depth = args.depth
    t0 = time.perf_counter()
    # reference in chunks of at most 64x64 spp (progress lines; distinct seeds per chunk)
    chunk = min(args.ref_spp_axis, 64 if W * H <= (1 << 21) else 16)
    n_chunks = max(1, (args.ref_spp_axis // chunk) ** 2)
    for k in range(n_chunks):
        ctx.render(W, H, chunk, max_depth=depth, seed=99991 + 7919 * k,
                   flags=cpt.CPT_RENDER_ACCUMULATE if k else 0)
        print(f"bench: reference chunk {k + 1}/{n_chunks}", file=sys.stderr, flush=True)
    _, ref_sums = ctx.read_film(sums=True)
    ref_spp = n_chunks * chunk * chunk
    ref = ref_sums / float(ref_spp)
    ref_s = time.perf_counter() - t0
    print(f"bench: reference {ref_spp} spp in {ref_s:.1f} s", file=sys.stderr, flush=True)

    def relmse(img):
        return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))

    def relmse_trim(img):
        # without the worst 1% of pixels: full relMSE at these sample counts is
        # decided by a few fireflies
        e = np.sort(((img - ref) ** 2 / (ref ** 2 + 1e-2)).mean(axis=2).ravel())
        return float(e[:int(0.99 * e.size)].mean())

    def run(guided):
        if guided:
            ctx.guide_init(seed=11, lr=3e-3, train_paths=args.train_paths or min(W * H, 1 << 18), train_steps=1,
                           guide_depth=args.guide_depth, infer_fp8=args.fp8, width=args.guide_width)
        flags = (cpt.CPT_RENDER_GUIDED | cpt.CPT_RENDER_GUIDE_TRAIN) if guided else 0
        curve = []
        elapsed, spp, k = 0.0, 0, 0
        gsamples = gsteps = 0
        next_check = 0.25
        while elapsed < budgets[-1]:
            t = time.perf_counter()
            st = ctx.render(W, H, 1, max_depth=depth, seed=1 + k, flags=flags | (cpt.CPT_RENDER_ACCUMULATE if k else 0))
            elapsed += time.perf_counter() - t
            gsamples += st["guide_samples"]
            gsteps += st["guide_steps"]
            spp += 1
            k += 1
            if elapsed >= next_check or elapsed >= budgets[-1]:
                _, sums = ctx.read_film(sums=True)
                curve.append((elapsed, spp, relmse(sums / spp), relmse_trim(sums / spp)))
                next_check = elapsed * 1.25
                print(f"bench: {'guided' if guided else 'plain'} t={elapsed:.2f}s spp={spp} relMSE={curve[-1][2]:.4f} "
                      f"trimmed={curve[-1][3]:.5f}" +
                      (f" training samples/step {gsamples / max(gsteps, 1):.0f}" if guided else ""),
                      file=sys.stderr, flush=True)
        return curve

    plain = run(False)
    guided = run(True)

    # SD-tree arm: the reference's PPG schedule (iterations of 1, 2, 4, ... spp, the
    # last one kept; no NEE), sized to each budget from a short calibration run
    ppg = {}
    if not args.no_ppg:
        t = time.perf_counter()
        ctx.render(W, H, 8, max_depth=depth, seed=7, flags=cpt.CPT_RENDER_PPG)
        per_spp = (time.perf_counter() - t) / 63.0
        for T in budgets:
            # the schedule renders about 2x its last iteration: total ~ sa^2 spp
            sa = max(2, int(math.sqrt(T / per_spp)))
            t = time.perf_counter()
            ctx.render(W, H, sa, max_depth=depth, seed=11, flags=cpt.CPT_RENDER_PPG)
            el = time.perf_counter() - t
            last = 1
            while last * 2 < sa * sa:
                last *= 2
            _, sums = ctx.read_film(sums=True)
            ppg[T] = (el, last, relmse(sums / last), relmse_trim(sums / last))
            print(f"bench: ppg budget {T:.0f}s: {el:.1f}s, last iteration {last} spp, relMSE={ppg[T][2]:.4f}",
                  file=sys.stderr, flush=True)

    # per-kernel time of a 1-spp frame from HIP events (eager launches), with the
    # guide as trained by the guided run
    def frame_ms(flags):
        acc = {}
        for k in range(4):
            st = ctx.render(W, H, 1, max_depth=depth, seed=900 + k, flags=flags | cpt.CPT_RENDER_KERNEL_TIMING)
            for key in ("ms_extend", "ms_shade", "ms_shadow", "ms_guide", "ms_other"):
                acc[key] = acc.get(key, 0.0) + st[key] / 4
        return {k: round(v, 3) for k, v in acc.items()}

    kernel_ms = {"plain": frame_ms(0), "guided": frame_ms(cpt.CPT_RENDER_GUIDED)}
    print(f"bench: per-frame kernel ms {kernel_ms}", file=sys.stderr, flush=True)

    def at(curve, T):
        best = curve[0]
        for c in curve:
            if c[0] <= T * 1.0001:
                best = c
        return best

    rows = []
    for T in budgets:
        p_ = at(plain, T)
        g_ = at(guided, T)
        reach = next((c[0] for c in guided if c[2] <= p_[2]), None)
        reach_t = next((c[0] for c in guided if c[3] <= p_[3]), None)
        rows.append({"budget_s": T, "plain_relmse": p_[2], "plain_spp": p_[1], "guided_relmse": g_[2],
                     "guided_spp": g_[1], "guided_time_to_plain_relmse_s": reach,
                     "time_ratio": (reach / p_[0]) if reach else None,
                     "plain_relmse_trimmed": p_[3], "guided_relmse_trimmed": g_[3],
                     "time_ratio_trimmed": (reach_t / p_[0]) if reach_t else None})
        if T in ppg:
            rows[-1].update({"ppg_seconds": ppg[T][0], "ppg_last_spp": ppg[T][1], "ppg_relmse": ppg[T][2],
                             "ppg_relmse_trimmed": ppg[T][3]})
        print(f"bench: T={T:.0f}s plain {p_[2]:.4f} ({p_[1]} spp) guided {g_[2]:.4f} ({g_[1]} spp) "
              f"guided reaches plain's error at {reach}", file=sys.stderr, flush=True)
    out = {
        "metric": "equal-time relMSE guided vs plain PT (1 GPU)",
        "value": rows[-1]["plain_relmse"] / max(rows[-1]["guided_relmse"], 1e-30),
        "unit": "relMSE ratio plain/guided at the largest budget",
        "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "dtype": "f32 (guide: " + ("fp8 e4m3 MFMA inference, bf16 training" if args.fp8 else "bf16 MFMA") + ", fp32 acc)",
        "guide": {"width": args.guide_width, "lobes": 8 if args.guide_width == 64 else 16, "infer_fp8": bool(args.fp8),
                  "train_paths": args.train_paths or min(W * H, 1 << 18)},
        "config": {"workload": f"{'4x4 grid of hard-lighting interiors, 32 emitters, mixed GGX' if args.scene == 'grid' else 'hard-lighting interior'}, "
                               f"{scene.n_triangles} triangles, {W}x{H}, max depth {depth}, "
                               f"reference {ref_spp} spp FP64", "budgets_s": budgets},
        "budgets": rows,
        "frame_kernel_ms": kernel_ms,
        "reference_seconds": ref_s,
    }
    ctx.close()
    print(json.dumps(out))


def bench_views(args):
    """64 camera views of the box room, resolutions from {256^2, 512x384, 1024x768,
    1920x1080}, spp from {16, 64, 256}, rendered as one streaming batch."""
    import numpy as np

    import cpt

    rng = np.random.default_rng(7)
    sizes = [(256, 256), (512, 384), (1024, 768), (1920, 1080)]
    cams, whs = [], []
    for v in range(args.views):
        w, h = sizes[rng.integers(4)]
        sa = int(rng.choice([4, 8, 16]))
        ang = rng.uniform(-0.35, 0.35)
        o = (3.9 * np.sin(ang), rng.uniform(-0.25, 0.25), 3.9 * np.cos(ang))
        cams.append(cpt.scenes.camera_frame(o, (0, 0, 0), (0, 1, 0), 0.35411, w, h))
        whs.append((w, h, sa))
    cams, whs = np.stack(cams), np.array(whs, np.int32)
    ctx = cpt.Context(0)
    ctx.upload(cpt.scenes.box_room())
    ctx.render_views(cams[:2], np.minimum(whs[:2], [256, 256, 2]), max_depth=args.depth)  # warm-up
    t0 = time.perf_counter()
    _, ms, st = ctx.render_views(cams, whs, max_depth=args.depth, seed=1)
    total = time.perf_counter() - t0
    out = {
        "metric": "aggregate samples/s, multi-view batch",
        "value": st["paths"] / total,
        "unit": "samples/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic",
        "dtype": "f32",
        "config": {"workload": f"{args.views} views of the 36-triangle box room, uneven resolution/spp, "
                               f"max depth {args.depth}, one batch of multi-view passes",
                   "paths": st["paths"]},
        "seconds": total,
        "mrays_per_s": (st["extension_rays"] + st["shadow_rays"]) / total / 1e6,
        "view_ms_p50": float(np.percentile(ms, 50)),
        "view_ms_p99": float(np.percentile(ms, 99)),
        "view_ms_max": float(ms.max()),
    }
    ctx.close()
    print(json.dumps(out))


def bench_bvh(args):
    """Host binned-SAH build vs the device PLOC build on the terrain workload:
    build time, BVH shape, and the traversal kernels' time and node visits over
    the same rays (renders are identical by construction)."""
    import cpt

    W, H, S, depth = args.width, args.height, args.spp_axis, args.depth
    gx, gz = (int(v) for v in args.grid.split("x"))
    scene = cpt.scenes.terrain_scene(res=(W, H), n_spheres=args.spheres, grid=(gx, gz))
    ctx = cpt.Context(0)
    res = {}
    for name, gpu in (("host_sah", False), ("device_ploc", True)):
        t0 = time.perf_counter()
        ctx.upload(scene, gpu_bvh=gpu)
        up = time.perf_counter() - t0
        info = ctx.bvh_info()
        ctx.render(W, H, S, max_depth=depth, seed=1)
        tm = ctx.render(W, H, S, max_depth=depth, seed=1, flags=cpt.CPT_RENDER_KERNEL_TIMING)
        ct = ctx.render(W, H, S, max_depth=depth, seed=1, flags=cpt.CPT_RENDER_COUNT)
        t0 = time.perf_counter()
        st = ctx.render(W, H, S, max_depth=depth, seed=1)
        frame = time.perf_counter() - t0
        res[name] = {"build_ms": info["build_ms"], "upload_s": up, "nodes": info["nodes"], "depth": info["depth"],
                     "leaves": info["leaves"], "ms_extend": tm["ms_extend"], "ms_shadow": tm["ms_shadow"],
                     "ms_frame": frame * 1e3, "samples_per_s": st["paths"] / frame,
                     "nodes_per_ray": ct["node_visits"] / max(1, ct["extension_rays"]),
                     "tris_per_ray": ct["triangle_tests"] / max(1, ct["extension_rays"])}
        print(f"bench: {name} {res[name]}", file=sys.stderr, flush=True)
    ctx.close()
    h, d = res["host_sah"], res["device_ploc"]
    print(json.dumps({
        "metric": "BVH build time (device PLOC vs host binned SAH) and traversal cost",
        "value": d["build_ms"], "unit": "ms", "higher_is_better": False, "n_gpus": 1, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"procedural terrain {gx}x{gz} + {args.spheres} icospheres, "
                               f"{scene.n_triangles} triangles, {W}x{H}, {S * S} spp, max depth {depth}"},
        "build_speedup": h["build_ms"] / d["build_ms"], "extend_time_ratio": d["ms_extend"] / h["ms_extend"],
        "host_sah": h, "device_ploc": d}))


def bench_box(args):
    """The reference's own CPU-runnable case: the 36-triangle box room, 512x512, 64 spp
    (settings spp 8), max depth 8; also at the reference's fixed 32 bounces.  One
    step = one 64-spp frame."""
    import cpt

    W = H = 512
    S = 8
    scene = cpt.scenes.box_room(res=(W, H))
    ctx = cpt.Context(0)
    ctx.upload(scene)
    res = {}
    for depth in (8, 32):
        ctx.autotune(W, H, max_depth=depth)
        for w in range(args.warmup):
            ctx.render(W, H, S, max_depth=depth, seed=1000 + w)
        t0 = time.perf_counter()
        paths = rays = 0
        for k in range(args.steps):
            st = ctx.render(W, H, S, max_depth=depth, seed=1 + k, flags=cpt.CPT_RENDER_ACCUMULATE if k else 0)
            paths += st["paths"]
            rays += st["extension_rays"] + st["shadow_rays"]
        total = time.perf_counter() - t0
        res[depth] = {"samples_per_s": paths / total, "ms_per_step": total / args.steps * 1e3,
                      "mrays_per_s": rays / total / 1e6}
        print(f"bench: box room depth {depth}: {res[depth]}", file=sys.stderr, flush=True)
    ctx.close()
    out = {
        "metric": "samples/s (box room, the reference's CPU-runnable case)",
        "value": res[8]["samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res[8]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "box room, 36 triangles, 512x512, 64 spp per step, max depth 8"},
        "mrays_per_s": res[8]["mrays_per_s"], "depth32": res[32],
    }
    if not args.no_cpu:
        cb = cpu_baseline(scene, W, H, 8, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out))


def cpu_baseline(scene, W, H, depth, target_s):
    import numpy as np

    import oracle

    threads = min(16, os.cpu_count() or 1)
    o = oracle.OracleScene(scene)
    # spread single rows over the frame; grow the sample until it takes ~target_s
    rows_done, paths, rays, elapsed = 0, 0, 0, 0.0
    order = np.random.default_rng(0).permutation(H)
    while elapsed < target_s and rows_done < 4 * H:
        r = int(order[rows_done % H])
        t0 = time.perf_counter()
        _, _, counts = o.render(W, H, 1, max_depth=depth, sampler="mt", seed=0, rows=(r, r + 1), threads=threads)
        elapsed += time.perf_counter() - t0
        paths += int(counts[0])
        rays += int(counts[1] + counts[2])
        rows_done += 1
    o.close()
    return {
        "value": paths / elapsed,
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{rows_done} random rows (cycled) x {W} px x 1 spp of the same {W}x{H} scene, depth {depth}, "
                  f"mt19937 column sampler, reference binary BVH; {elapsed:.1f} s",
        "mrays_per_s": rays / elapsed / 1e6,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding across GPUs is not built", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.workload == "views":
        bench_views(args)
        return
    if args.workload == "guided":
        bench_guided(args)
        return
    if args.workload == "bvh":
        bench_bvh(args)
        return
    if args.workload == "box":
        bench_box(args)
        return
    import cpt

    W, H, S, depth = args.width, args.height, args.spp_axis, args.depth
    gx, gz = (int(v) for v in args.grid.split("x"))
    t0 = time.perf_counter()
    scene = cpt.scenes.terrain_scene(res=(W, H), n_spheres=args.spheres, grid=(gx, gz))
    ctx = cpt.Context(0)
    ctx.upload(scene, gpu_bvh=args.gpu_bvh)
    setup_s = time.perf_counter() - t0
    bvh = ctx.bvh_info()
    print(f"bench: scene {bvh['triangles']} triangles, {bvh['nodes']} BVH4 nodes "
          f"({'device' if bvh['gpu_built'] else 'host'} build {bvh['build_ms']:.0f} ms), setup {setup_s:.1f} s",
          file=sys.stderr, flush=True)
    kw = dict(max_depth=depth, block_size=args.block, grid_blocks=args.grid_blocks, max_paths=args.max_paths)
    tune = None
    if not args.no_tune:
        # launch configuration search (part of setup, outside the timed steps)
        tune = ctx.autotune(W, H, max_depth=depth, max_paths=args.max_paths)
        print(f"bench: autotune {tune}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        ctx.render(W, H, S, seed=1000 + w, **kw)
    t0 = time.perf_counter()
    ext = shadow = paths = 0
    for k in range(args.steps):
        st = ctx.render(W, H, S, seed=1 + k, flags=cpt.CPT_RENDER_ACCUMULATE if k else 0, **kw)
        ext += st["extension_rays"]
        shadow += st["shadow_rays"]
        paths += st["paths"]
        print(f"bench: step {k + 1}/{args.steps} {st['seconds'] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    total = time.perf_counter() - t0
    samples_per_s = paths / total
    out = {
        "metric": METRIC,
        "value": samples_per_s,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"GPU wavefront PT-NEE-MIS, procedural terrain {gx}x{gz} fBm heightfield + "
                        f"{args.spheres} icospheres (subdiv 4), {bvh['triangles']} triangles, {W}x{H}, "
                        f"{S * S} spp per step ({S * S * args.steps} spp total), max depth {depth}",
            "width": W, "height": H, "spp_per_step": S * S, "max_depth": depth,
            "triangles": bvh["triangles"],
            "bvh_build": "device PLOC" if bvh["gpu_built"] else "host binned SAH",
            "launch": tune,
        },
        "mrays_per_s": (ext + shadow) / total / 1e6,
        "extension_mrays_per_s": ext / total / 1e6,
        "shadow_mrays_per_s": shadow / total / 1e6,
        "setup_s": setup_s,
    }
    if not args.no_roofline:
        # same seed -> same rays: timing run (HIP events around every launch) and counting run
        tm = ctx.render(W, H, S, seed=1, flags=cpt.CPT_RENDER_KERNEL_TIMING, **kw)
        ct = ctx.render(W, H, S, seed=1, flags=cpt.CPT_RENDER_COUNT, **kw)
        alg_bytes = 48 * ct["extension_rays"] + 128 * ct["node_visits"] + 48 * ct["triangle_tests"]
        ext_s = tm["ms_extend"] / 1e3
        achieved = alg_bytes / ext_s / 1e9
        traffic = None
        tp = os.path.join(ROOT, "profiles", "extend_traffic.json")
        if os.path.exists(tp):
            traffic = json.load(open(tp)).get("hbm_bytes_per_launch")
        out["roofline"] = {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_extend (closest-hit BVH4 traversal)",
            "launches": tm["n_extend"],
            "avg_launch_us": tm["ms_extend"] / max(1, tm["n_extend"]) * 1e3,
            "bytes_per_launch": alg_bytes / max(1, tm["n_extend"]),
            "nodes_per_ray": ct["node_visits"] / max(1, ct["extension_rays"]),
            "tris_per_ray": ct["triangle_tests"] / max(1, ct["extension_rays"]),
            "ms_shade": tm["ms_shade"],
            "ms_shadow": tm["ms_shadow"],
            "ms_extend": tm["ms_extend"],
        }
    if not args.no_cpu:
        cb = cpu_baseline(scene, W, H, depth, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = samples_per_s / cb["value"]
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
