"""Benchmark: GPU wavefront PT-NEE-MIS on the procedural ~2.07M-triangle terrain
scene (1024x512 fBm heightfield + 200 icospheres, 1920x1080, max depth 12).

One step = one 16-spp pass over the whole 1920x1080 frame (spp_axis 4); the
default 16 steps accumulate the workload's 256 spp.  Inputs (scene, BVH) are
resident in HBM before timing starts; cpt_render returns after the device has
finished (it synchronises its stream).

Prints one JSON line.  `roofline` describes the dominant kernel (closest-hit
traversal, `k_extend`): achieved = HBM bytes per launch (two rocprofv3 counter
passes, TCC_EA0_RDREQ_{32B,64B,128B} x size and WRITE_SIZE, over a child process rendering the same frame
with the same launch configuration and seed, scaled by device copies of known
size) / that kernel's average launch time (HIP events on the stream it runs on).
The bytes the traversal asks of the whole memory hierarchy (32 B ray + 16 B hit +
64 B per quantized BVH4 node visit + 48 B per triangle test, counted by a counting run of
the same kernel on the same rays) are reported beside it as `cache_level`.
`cpu_baseline` renders whole frames of the same workload with the CPU oracle (a
restatement of the reference's PathtracerMIS with its mt19937 column sampler and
binary BVH) over the host's cores.  `guided` is the equal-time comparison of the
learned guide against plain PT on the hard-lighting interior (1280x720, budgets
10/30/60 s, 65,536-spp FP64 reference).
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
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline duration (whole frames)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip cpt_autotune (default launch configuration)")
    ap.add_argument("--gpu-bvh", action="store_true", help="build the BVH on the device (CPT_SCENE_GPU_BVH)")
    ap.add_argument("--workload", choices=["terrain", "views", "guided", "bvh", "box", "mixtures", "train"],
                    default="terrain",
                    help="terrain: the headline 1080p line; views: 64-view batch of the box room; "
                         "guided: equal-time relMSE of guided vs plain PT on the hard-lighting interior; "
                         "mixtures: the guide's per-bounce inference (cpt_guide_mixtures) on random inputs; "
                         "train: Adam steps of the guide (cpt_guide_train_step) on random samples")
    ap.add_argument("--budgets", type=str, default="10,30,60", help="guided: equal-time budgets in seconds")
    ap.add_argument("--repeats", type=int, default=1,
                    help="guided: runs of each arm with distinct seeds; budget rows give medians")
    ap.add_argument("--ref-spp-axis", type=int, default=64, help="guided: reference spp per axis (64 -> 4096 spp)")
    ap.add_argument("--boxes", type=int, default=40000, help="guided: furniture boxes (48 triangles each)")
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--no-ppg", action="store_true", help="guided: skip the SD-tree (PPG) arm")
    ap.add_argument("--fp8", action="store_true", help="guided: run the guide network in FP8 (training stays BF16)")
    ap.add_argument("--train-paths", type=int, default=0,
                    help="guided: paths per pass that record training samples, spread over the image "
                         "(0: 2^18, or 2^21 above 4M pixels)")
    ap.add_argument("--guide-width", type=int, choices=[64, 128], default=64,
                    help="guided: hidden width of the guide network (64: 8 lobes, 128: 16 lobes)")
    ap.add_argument("--fraction-ramp", type=int, default=100,
                    help="guided: Adam steps over which the guide fraction rises from 0 to 0.5 (0: constant)")
    ap.add_argument("--guide-depth", type=int, default=0,
                    help="guided: guide the vertices below this depth only (0: every vertex)")
    ap.add_argument("--scene", choices=["interior", "grid"], default="interior",
                    help="guided: the hard-lighting interior, or the 4x4 grid of interiors (7.7M triangles, "
                         "32 emitters, mixed GGX; device BVH build)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (scales HBM counters)")
    ap.add_argument("--shade-block", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--no-traffic", action="store_true",
                    help="terrain: HBM traffic from the committed counter measurement instead of two rocprofv3 passes")
    ap.add_argument("--no-guided", action="store_true", help="terrain: leave out the equal-time guided comparison")
    ap.add_argument("--guided-res", type=str, default="1280x720", help="terrain line: resolution of the guided comparison")
    ap.add_argument("--guided-ref-spp-axis", type=int, default=256,
                    help="terrain line: reference spp per axis of the guided comparison (256 -> 65,536 spp)")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
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


def guided_equal_time(W, H, budgets, depth=12, fp8=False, width=64, train_paths=0, guide_depth=0, scene_kind="interior",
                      boxes=40000, ref_spp_axis=256, ppg=False, log=None, repeats=1, fraction_ramp=0):
    """Equal-time comparison on the hard-lighting interior (one 0.2x0.2 light outside
    a cluttered room, light entering through a 0.3-wide door gap): plain PT-NEE-MIS
    vs guided PT (learned MLP guide trained online, training time included), both
    accumulated one sample per pixel at a time; relMSE = mean((I-R)^2/(R^2+0.01))
    against an FP64-accumulated plain reference of ref_spp_axis^2 samples per pixel.
    With repeats > 1 both arms run that many times with distinct seeds and the
    budget rows give medians over the runs (full relMSE at these sample counts is
    decided by a few fireflies and varies from run to run).  Returns the result
    record (the `guided` object of the bench line)."""
    import numpy as np

    import cpt

    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    ctx = cpt.Context(0)
    t0 = time.perf_counter()
    if scene_kind == "grid":
        scene = cpt.scenes.interior_grid_scene(res=(W, H))
        # the device build keeps set-up short at 7.7M triangles; the tie-rank table only
        # decides exact-t ties between coincident triangles, which this scene has none of
        ctx.upload(scene, gpu_bvh=True, tie_rank=False)
    else:
        scene = cpt.scenes.interior_scene(res=(W, H), n_boxes=boxes)
        ctx.upload(scene)
    log(f"bench: guided scene {scene.n_triangles} triangles ready in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    # reference in chunks of at most 64x64 spp (progress lines; distinct seeds per chunk)
    chunk = min(ref_spp_axis, 64 if W * H <= (1 << 21) else 16)
    n_chunks = max(1, (ref_spp_axis // chunk) ** 2)
    for k in range(n_chunks):
        ctx.render(W, H, chunk, max_depth=depth, seed=99991 + 7919 * k,
                   flags=cpt.CPT_RENDER_ACCUMULATE if k else 0)
        log(f"bench: reference chunk {k + 1}/{n_chunks}")
    _, ref_sums = ctx.read_film(sums=True)
    ref_spp = n_chunks * chunk * chunk
    ref = ref_sums / float(ref_spp)
    ref_s = time.perf_counter() - t0
    log(f"bench: reference {ref_spp} spp in {ref_s:.1f} s")

    def relmse(img):
        return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))

    def relmse_trim(img):
        # without the worst 1% of pixels: full relMSE at these sample counts is
        # decided by a few fireflies
        e = np.sort(((img - ref) ** 2 / (ref ** 2 + 1e-2)).mean(axis=2).ravel())
        return float(e[:int(0.99 * e.size)].mean())

    # recording paths per pass (spread over the image): 2^18 at 720p, 2^21 at 4K
    tp = train_paths or min(W * H, 1 << 21 if W * H > (1 << 22) else 1 << 18)

    def run(guided, rep):
        seed0 = 1 + 1_000_003 * rep + (500_000 if guided else 0)
        if guided:
            ctx.guide_init(seed=11 + rep, lr=3e-3, train_paths=tp, train_steps=1, guide_depth=guide_depth,
                           infer_fp8=fp8, width=width, fraction_ramp=fraction_ramp)
        flags = (cpt.CPT_RENDER_GUIDED | cpt.CPT_RENDER_GUIDE_TRAIN) if guided else 0
        curve = []
        elapsed, spp, k = 0.0, 0, 0
        gsamples = gsteps = 0
        next_check = 0.25
        while elapsed < budgets[-1]:
            t = time.perf_counter()
            st = ctx.render(W, H, 1, max_depth=depth, seed=seed0 + k,
                            flags=flags | (cpt.CPT_RENDER_ACCUMULATE if k else 0))
            elapsed += time.perf_counter() - t
            gsamples += st["guide_samples"]
            gsteps += st["guide_steps"]
            spp += 1
            k += 1
            if elapsed >= next_check or elapsed >= budgets[-1]:
                _, sums = ctx.read_film(sums=True)
                curve.append((elapsed, spp, relmse(sums / spp), relmse_trim(sums / spp)))
                next_check = elapsed * 1.25
                log(f"bench: run {rep} {'guided' if guided else 'plain'} t={elapsed:.2f}s spp={spp} relMSE={curve[-1][2]:.4f} "
                    f"trimmed={curve[-1][3]:.5f}" +
                    (f" training samples/step {gsamples / max(gsteps, 1):.0f}" if guided else ""))
        return curve, gsamples / max(gsteps, 1), gsteps

    plains, guideds, per_step, steps = [], [], [], []
    for rep in range(repeats):
        plains.append(run(False, rep)[0])
        g_, sps, ns = run(True, rep)
        guideds.append(g_)
        per_step.append(sps)
        steps.append(ns)

    # SD-tree arm: the reference's PPG schedule (iterations of 1, 2, 4, ... spp, the
    # last one kept; no NEE), sized to each budget from a short calibration run
    ppg_res = {}
    if ppg:
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
            ppg_res[T] = (el, last, relmse(sums / last), relmse_trim(sums / last))
            log(f"bench: ppg budget {T:.0f}s: {el:.1f}s, last iteration {last} spp, relMSE={ppg_res[T][2]:.4f}")

    # per-kernel time of a 1-spp frame from HIP events (eager launches), with the
    # guide as trained by the guided run; guide inference also per launch (one
    # launch per bounce)
    def frame_ms(flags):
        acc = {}
        n_guide = 0
        for k in range(4):
            st = ctx.render(W, H, 1, max_depth=depth, seed=900 + k, flags=flags | cpt.CPT_RENDER_KERNEL_TIMING)
            for key in ("ms_extend", "ms_shade", "ms_shadow", "ms_guide", "ms_train", "ms_other"):
                acc[key] = acc.get(key, 0.0) + st[key] / 4
            n_guide += st["n_guide"]
        res = {k: round(v, 3) for k, v in acc.items()}
        if n_guide:
            res["guide_launches"] = n_guide / 4
            res["ms_guide_per_launch"] = round(4 * acc["ms_guide"] / n_guide, 4)
        return res

    kernel_ms = {"plain": frame_ms(0), "guided": frame_ms(cpt.CPT_RENDER_GUIDED),
                 "guided_training": frame_ms(cpt.CPT_RENDER_GUIDED | cpt.CPT_RENDER_GUIDE_TRAIN)}
    log(f"bench: per-frame kernel ms {kernel_ms}")

    def at(curve, T):
        best = curve[0]
        for c in curve:
            if c[0] <= T * 1.0001:
                best = c
        return best

    def run_row(plain, guided, T):
        p_ = at(plain, T)
        g_ = at(guided, T)
        reach = next((c[0] for c in guided if c[2] <= p_[2]), None)
        reach_t = next((c[0] for c in guided if c[3] <= p_[3]), None)
        return {"budget_s": T, "plain_relmse": p_[2], "plain_spp": p_[1], "guided_relmse": g_[2],
                "guided_spp": g_[1], "guided_time_to_plain_relmse_s": reach,
                "time_ratio": (reach / p_[0]) if reach else None,
                "plain_relmse_trimmed": p_[3], "guided_relmse_trimmed": g_[3],
                "time_ratio_trimmed": (reach_t / p_[0]) if reach_t else None}

    def median(xs):
        xs = [x for x in xs if x is not None]
        return float(np.median(xs)) if xs else None

    runs = [[run_row(plains[r], guideds[r], T) for T in budgets] for r in range(repeats)]
    rows = []
    for i, T in enumerate(budgets):
        if repeats == 1:
            rows.append(dict(runs[0][i]))
        else:
            # a run that never reaches plain's error counts as beyond the budget
            reach = [runs[r][i]["guided_time_to_plain_relmse_s"] for r in range(repeats)]
            rows.append({k: median([runs[r][i][k] for r in range(repeats)])
                         for k in runs[0][i] if k not in ("guided_time_to_plain_relmse_s", "time_ratio")})
            rows[-1]["budget_s"] = T
            rows[-1]["runs_reaching_plain_relmse"] = sum(x is not None for x in reach)
            rows[-1]["guided_time_to_plain_relmse_s"] = (
                float(np.median([x if x is not None else float("inf") for x in reach]))
                if sum(x is not None for x in reach) * 2 > repeats else None)
            rows[-1]["time_ratio"] = (rows[-1]["guided_time_to_plain_relmse_s"] / T
                                      if rows[-1]["guided_time_to_plain_relmse_s"] is not None else None)
            rows[-1]["relmse_ratio_per_run"] = [runs[r][i]["plain_relmse"] / max(runs[r][i]["guided_relmse"], 1e-30)
                                                for r in range(repeats)]
        if T in ppg_res:
            rows[-1].update({"ppg_seconds": ppg_res[T][0], "ppg_last_spp": ppg_res[T][1], "ppg_relmse": ppg_res[T][2],
                             "ppg_relmse_trimmed": ppg_res[T][3]})
        log(f"bench: T={T:.0f}s plain {rows[-1]['plain_relmse']:.4f} guided {rows[-1]['guided_relmse']:.4f} "
            f"guided reaches plain's error at {rows[-1]['guided_time_to_plain_relmse_s']}")
    ctx.close()
    return {
        "metric": "equal-time relMSE guided vs plain PT (1 GPU)",
        "workload": f"{'4x4 grid of hard-lighting interiors, 32 emitters, mixed GGX' if scene_kind == 'grid' else 'hard-lighting interior'}, "
                    f"{scene.n_triangles} triangles, {W}x{H}, max depth {depth}, reference {ref_spp} spp FP64",
        "budgets_s": budgets,
        "time_ratio_at_largest_budget": rows[-1]["time_ratio"],
        "relmse_ratio_plain_over_guided": (median(rows[-1]["relmse_ratio_per_run"]) if repeats > 1 else
                                           rows[-1]["plain_relmse"] / max(rows[-1]["guided_relmse"], 1e-30)),
        "repeats": repeats,
        "guide": {"width": width, "lobes": 8 if width == 64 else 16, "inference": "fp8 e4m3 MFMA" if fp8 else "bf16 MFMA",
                  "training": "bf16 MFMA, fp32 accumulation and Adam", "train_paths": tp,
                  "training_samples_per_step": median(per_step), "adam_steps": median(steps),
                  "guide_fraction": 0.5, "fraction_ramp_steps": fraction_ramp},
        "budgets": rows,
        "runs": runs if repeats > 1 else None,
        "frame_kernel_ms": kernel_ms,
        "reference_seconds": ref_s,
    }


def bench_guided(args):
    """`--workload guided`: the equal-time comparison alone, as its own JSON line."""
    W, H = args.width, args.height
    budgets = [float(b) for b in args.budgets.split(",")]
    g = guided_equal_time(W, H, budgets, depth=args.depth, fp8=args.fp8, width=args.guide_width,
                          train_paths=args.train_paths, guide_depth=args.guide_depth, scene_kind=args.scene,
                          boxes=args.boxes, ref_spp_axis=args.ref_spp_axis, ppg=not args.no_ppg,
                          repeats=args.repeats, fraction_ramp=args.fraction_ramp)
    rows = g["budgets"]
    out = {
        "metric": g["metric"],
        "value": g["relmse_ratio_plain_over_guided"],
        "unit": "relMSE ratio plain/guided at the largest budget",
        "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
        "dtype": "f32 (guide: " + ("fp8 e4m3 MFMA inference, bf16 training" if args.fp8 else "bf16 MFMA") + ", fp32 acc)",
        "guide": g["guide"],
        "config": {"workload": g["workload"], "budgets_s": budgets},
        "budgets": rows,
        "repeats": g["repeats"],
        "runs": g["runs"],
        "frame_kernel_ms": g["frame_kernel_ms"],
        "reference_seconds": g["reference_seconds"],
    }
    print(json.dumps(out))


def bench_views(args):
    """64 camera views of the box room, resolutions from {256^2, 512x384, 1024x768,
    1920x1080}, spp from {16, 64, 256}, rendered as one batch of multi-view passes
    (views scheduled shortest first); per-view completion times from device
    timestamps after each view's last pass."""
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


def bench_mixtures(args, n=1 << 22, calls=3):
    """`--workload mixtures`: the guide network's inference with mixture decode
    (cpt_guide_mixtures, the device code of the per-bounce inference after its hit
    gather) on n random inputs, --guide-width / --fp8.  The call's time includes
    the host copies; the kernel's own time comes from a rocprofv3 kernel trace of
    this command."""
    import numpy as np
    import cpt

    X = np.random.default_rng(1).random((n, 7)).astype(np.float32)
    with cpt.Context(0) as ctx:
        ctx.guide_init(seed=3, infer_fp8=args.fp8, width=args.guide_width)
        ctx.guide_mixtures(X[:4096])
        times = []
        for _ in range(calls):
            t0 = time.perf_counter()
            ctx.guide_mixtures(X)
            times.append(time.perf_counter() - t0)
    t = min(times)
    print(json.dumps({
        "metric": "guide mixtures per second (host copies included)", "value": n / t, "unit": "samples/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "fp8 e4m3" if args.fp8 else "bf16", "data": "synthetic",
        "config": {"workload": f"cpt_guide_mixtures on {n} random inputs, width {args.guide_width}",
                   "calls": calls, "best_call_s": t},
    }))


def bench_train(args, n=204800, steps=8):
    """`--workload train`: Adam steps of the guide network (cpt_guide_train_step:
    forward, loss gradient, backward, weight gradients, Adam, the FP8 copy when
    --fp8) on n random training samples (the interior configuration's about 200k
    samples per step at width 64), --guide-width.  The call's time includes the
    samples' host copy; the kernels' own times come from a rocprofv3 kernel trace
    of this command."""
    import numpy as np
    import cpt

    rng = np.random.default_rng(2)
    S = np.zeros((n, 11), np.float32)
    S[:, :7] = rng.random((n, 7))
    d = rng.standard_normal((n, 3))
    S[:, 7:10] = d / np.linalg.norm(d, axis=1, keepdims=True)
    S[:, 10] = rng.exponential(1.0, n)
    with cpt.Context(0) as ctx:
        ctx.guide_init(seed=3, lr=3e-3, infer_fp8=args.fp8, width=args.guide_width)
        ctx.guide_train_step(S[:4096])
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            ctx.guide_train_step(S)
            times.append(time.perf_counter() - t0)
    t = min(times)
    print(json.dumps({
        "metric": "guide training samples per second (host copies included)", "value": n / t, "unit": "samples/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "bf16 (fp32 accumulation and Adam)", "data": "synthetic",
        "config": {"workload": f"cpt_guide_train_step on {n} random samples, width {args.guide_width}",
                   "steps": steps, "best_step_s": t},
    }))


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
        cb = cpu_baseline(scene, W, H, 8, args.cpu_seconds, S)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out))


def host_info():
    """The host's cores as the CPU baseline sees them: the machine's logical CPUs,
    the ones this process may run on, the OpenMP thread setting, and lscpu's model,
    sockets and threads per core."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        import subprocess

        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "cpu_model", "Socket(s)": "sockets", "Thread(s) per core": "threads_per_core",
                "Core(s) per socket": "cores_per_socket"}
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except Exception as e:  # lscpu missing: keep the counts
        info["lscpu_error"] = str(e)
    return info


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may use, capped by the
    host's OpenMP thread setting when it has one."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return n


def cpu_baseline(scene, W, H, depth, target_s, full_spp_axis):
    """The oracle's restatement of the reference (PathtracerMIS with its mt19937
    column sampler and binary BVH) rendering WHOLE frames of the same workload over
    the host's cores, OpenMP over columns as the reference does.  A 1-spp frame
    measures the rate; a second frame with spp_axis chosen for about target_s of
    work follows when the first was short.  samples/s = all paths / all seconds."""
    import oracle

    threads = cpu_threads()
    t0 = time.perf_counter()
    o = oracle.OracleScene(scene)
    build_s = time.perf_counter() - t0
    frames, paths, rays, elapsed = [], 0, 0, 0.0
    sa = 1
    while True:
        t0 = time.perf_counter()
        _, _, counts = o.render(W, H, sa, max_depth=depth, sampler="mt", seed=0, threads=threads)
        dt = time.perf_counter() - t0
        elapsed += dt
        paths += int(counts[0])
        rays += int(counts[1] + counts[2])
        frames.append(f"{W}x{H} x {sa * sa} spp in {dt:.2f} s")
        if elapsed >= 0.5 * target_s or sa >= full_spp_axis or len(frames) >= 2:
            break
        sa = max(1, min(full_spp_axis, int(math.sqrt((target_s - elapsed) / max(dt, 1e-3)))))
    o.close()
    return {
        "value": paths / elapsed,
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"whole frames of the same workload ({'; '.join(frames)}), depth {depth}, the oracle's "
                  f"restatement of PathtracerMIS (mt19937 column sampler, reference binary BVH), OpenMP dynamic "
                  f"schedule over columns; scene set-up ({build_s:.1f} s) not timed",
        "mrays_per_s": rays / elapsed / 1e6,
        "seconds": elapsed,
        "host": host_info(),
    }


def traffic_probe(args):
    """Child mode run under `rocprofv3 --pmc`: one frame of the bench's workload with
    the given launch configuration and seed 1 (the rays of the parent's timing run),
    then three 512 MiB device-to-device copies (a transfer of known size that scales
    the counters)."""
    import ctypes

    import cpt

    W, H, S, depth = args.width, args.height, args.spp_axis, args.depth
    gx, gz = (int(v) for v in args.grid.split("x"))
    scene = cpt.scenes.terrain_scene(res=(W, H), n_spheres=args.spheres, grid=(gx, gz))
    ctx = cpt.Context(0)
    ctx.upload(scene)
    ctx.render(W, H, S, seed=1, max_depth=depth, block_size=args.block, grid_blocks=args.grid_blocks,
               shade_block=args.shade_block, refill=args.refill, flags=cpt.CPT_RENDER_NO_GRAPH)
    ctx.close()
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    n = 512 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0 and hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("traffic probe done", flush=True)


# Read bytes from the L2's read requests to the fabric by size (gfx950 counts 32-,
# 64- and 128-byte requests separately; its FETCH_SIZE formula takes 128-byte
# requests from TCC_BUBBLE, which reads half of a plain copy's bytes here), write
# bytes from WRITE_SIZE (32- and 64-byte write requests).
READ_COUNTERS = {"TCC_EA0_RDREQ_32B_sum": 32, "TCC_EA0_RDREQ_64B_sum": 64, "TCC_EA0_RDREQ_128B_sum": 128}
WRITE_COUNTERS = {"WRITE_SIZE": 1024}


def measure_traffic(args, launch, log):
    """Bytes per k_extend launch between the L2 and the memory fabric (the Infinity
    Cache and HBM behind it) from two rocprofv3 counter passes over a child process
    (traffic_probe): the read requests by size, then WRITE_SIZE.  The child's three
    512 MiB copies check the scale: each pass must read (write) 1.5 GiB for them;
    the measured ratio is applied.  Returns a dict, or None with the reason logged
    (then the committed measurement is used)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    if not shutil.which("rocprofv3"):
        log("bench: rocprofv3 not found; HBM traffic from the committed measurement")
        return None
    child = [sys.executable, os.path.abspath(__file__), "--traffic-probe", "--width", str(args.width),
             "--height", str(args.height), "--spp-axis", str(args.spp_axis), "--depth", str(args.depth),
             "--spheres", str(args.spheres), "--grid", args.grid, "--block", str(launch["block_size"]),
             "--grid-blocks", str(launch["grid_blocks"]), "--shade-block", str(launch["shade_block"]),
             "--refill", str(launch["refill"])]
    res = {}
    copy_bytes = 3 * (512 << 20)
    env = dict(os.environ, TMPDIR="/tmp")
    for kind, counters in (("read", READ_COUNTERS), ("write", WRITE_COUNTERS)):
        d = tempfile.mkdtemp(prefix="cpt_pmc_", dir="/tmp")
        cmd = ["rocprofv3", "--pmc"] + list(counters) + ["-d", d, "-o", "probe", "--output-format", "csv", "--"] + child
        t0 = time.perf_counter()
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=240, env=env, cwd="/tmp")
        except subprocess.TimeoutExpired:
            log(f"bench: rocprofv3 --pmc ({kind}) timed out; HBM traffic from the committed measurement")
            shutil.rmtree(d, ignore_errors=True)
            return None
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        ext, cpy = {}, {}
        for f in files:
            with open(f, newline="") as fh:
                for row in csv.DictReader(fh):
                    name, cn = row.get("Kernel_Name", ""), row.get("Counter_Name", "")
                    if cn not in counters:
                        continue
                    b = float(row.get("Counter_Value", "nan")) * counters[cn]
                    did = row.get("Dispatch_Id", row.get("Correlation_Id", ""))
                    if "k_extend<" in name:
                        ext[did] = ext.get(did, 0.0) + b
                    elif "copy" in name.lower():
                        cpy[did] = cpy.get(did, 0.0) + b
        shutil.rmtree(d, ignore_errors=True)
        log(f"bench: rocprofv3 --pmc {' '.join(counters)}: rc {p.returncode}, {len(ext)} k_extend and "
            f"{len(cpy)} copy dispatches in {time.perf_counter() - t0:.1f} s")
        if not ext or not cpy or "traffic probe done" not in p.stdout:
            log(f"bench: no usable {kind} counts (stderr tail: {p.stderr[-400:]!r}); "
                "HBM traffic from the committed measurement")
            return None
        per = sum(ext.values()) / len(ext)
        scale = copy_bytes / sum(cpy.values())
        res[kind] = {"counters": list(counters), "bytes_per_launch": per, "launches": len(ext),
                     "copy_bytes_counted": sum(cpy.values()), "copy_bytes": copy_bytes, "scale": scale}
    hbm = res["read"]["bytes_per_launch"] * res["read"]["scale"] + res["write"]["bytes_per_launch"] * res["write"]["scale"]
    return {"hbm_bytes_per_launch": hbm, "counters": res,
            "source": "rocprofv3 --pmc over a child process of this run (same scene, launch configuration and "
                      "seed): TCC_EA0_RDREQ_{32B,64B,128B} x size and WRITE_SIZE per k_extend dispatch, each "
                      "pass scaled by its count of 3 x 512 MiB device copies"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding across GPUs is not built", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.traffic_probe:
        traffic_probe(args)
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
    if args.workload == "mixtures":
        bench_mixtures(args)
        return
    if args.workload == "train":
        bench_train(args)
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
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if not args.no_roofline:
        # same seed -> same rays: timing run (HIP events around every launch), counting run,
        # and the counter passes of measure_traffic
        tm = ctx.render(W, H, S, seed=1, flags=cpt.CPT_RENDER_KERNEL_TIMING, **kw)
        ct = ctx.render(W, H, S, seed=1, flags=cpt.CPT_RENDER_COUNT, **kw)
        launches = max(1, tm["n_extend"])
        launch_s = tm["ms_extend"] / 1e3 / launches
        # bytes the traversal asks of the memory system (every node and triangle fetch
        # counted, whether L2, Infinity Cache or HBM serves it)
        cache_bytes = (48 * ct["extension_rays"] + 64 * ct["node_visits"] + 48 * ct["triangle_tests"]) / launches
        launch = dict(tune) if tune else {"block_size": args.block, "grid_blocks": args.grid_blocks,
                                          "shade_block": args.shade_block, "refill": args.refill}
        if args.block:
            launch["block_size"] = args.block
        traffic = None if args.no_traffic else measure_traffic(args, launch, log)
        if traffic is None:
            tp = os.path.join(ROOT, "profiles", "extend_traffic.json")
            with open(tp) as fh:
                rec = json.load(fh)
            traffic = {"hbm_bytes_per_launch": rec["hbm_bytes_per_launch"],
                       "source": f"profiles/extend_traffic.json ({rec.get('measured', 'committed measurement')})"}
        hbm = traffic["hbm_bytes_per_launch"]
        achieved = hbm / launch_s / 1e9
        out["roofline"] = {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": hbm,
            "traffic_source": traffic["source"],
            "kernel": "k_extend (closest-hit BVH4 traversal)",
            "launches": launches,
            "avg_launch_us": launch_s * 1e6,
            "rays_per_launch": ct["extension_rays"] / launches,
            "nodes_per_ray": ct["node_visits"] / max(1, ct["extension_rays"]),
            "tris_per_ray": ct["triangle_tests"] / max(1, ct["extension_rays"]),
            "cache_level": {"bytes_per_launch": cache_bytes, "gbs": cache_bytes / launch_s / 1e9,
                            "note": "48 B per ray + 64 B per node visit + 48 B per triangle test, served by "
                                    "L2/Infinity Cache/HBM together; not an HBM figure"},
            "ms_extend": tm["ms_extend"],
            "ms_shadow": tm["ms_shadow"],
            "ms_shade": tm["ms_shade"],
            "ms_other": tm["ms_other"],
        }
        if "counters" in traffic:
            out["roofline"]["counters"] = traffic["counters"]
    ctx.close()
    if not args.no_cpu:
        cb = cpu_baseline(scene, W, H, depth, args.cpu_seconds, S)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = samples_per_s / cb["value"]
        log(f"bench: cpu baseline {cb['value']:.4g} samples/s on {cb['cores']} threads ({cb['sample']})")
    del scene
    if not args.no_guided:
        gw, gh = (int(v) for v in args.guided_res.split("x"))
        out["guided"] = guided_equal_time(gw, gh, [10.0, 30.0, 60.0], depth=12, fp8=False, width=64,
                                          ref_spp_axis=args.guided_ref_spp_axis, log=log,
                                          fraction_ramp=args.fraction_ramp)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
