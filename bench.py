"""Training-iteration throughput of the MI355X splat trainer at 1M Gaussians, 1920x1080.

    python bench.py [--steps K] [--warmup W] [--ordered] [--cpu-steps C]

Workload (one GPU): 1M Gaussians drawn as means U[-5,5]^3, log-scales N(-5,0.7), random
unit quaternions, opacity logits N(0,1), SH degree 3; one 1920x1080 pinhole camera with
fx = fy = 1600 at distance 8; 16x16 tiles.  The target image is rendered from a second
scene of the same generator (seed 1) and stored as RGB8.  One step = project -> bin ->
composite -> L1 + 0.2 D-SSIM loss -> backward -> Adam (src/train.rs:211-247), all
resident in HBM.  `value` is iterations/s over K timed steps with inputs resident.

A second pass of K steps runs with HIP events around every kernel (on the trainer's
stream) to attribute time per kernel; the roofline entry uses the dominant kernel's
algorithmic bytes per launch over its mean duration.  Counting instantiations of the
compositing kernels give the exact evaluated (Gaussian, pixel) pairs, which put the
compositing kernels on the FP32 roofline as well.  The CPU baseline runs the oracle's
train step (a scalar C++ restatement of the reference) on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "train iters/s @1M Gaussians 1080p (plain vs batched+learned) at PSNR parity"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32
VALU_PEAK_GINST = 256 * 2.4  # CUs x GHz: wave64 VALU quad-cycles per second per SIMD (G)
# Algorithmic FP32 operations per evaluated (Gaussian, pixel) pair of the compositing walks
# (the reference's evaluate / forward_tile / backward_tile arithmetic, src/rasterize.rs:
# 231-290, 394-448): about 30 in the forward and 70 in the backward.
FLOPS_PER_PAIR = {"raster_fwd": 30, "raster_bwd": 70}


def algorithmic_bytes(name, N, V, E, P):
    """Bytes a kernel must move per launch (reads + writes of its inputs/outputs)."""
    table = {
        "project_fwd": 304 * N,              # 236 params read; 68 record/depth/radius/rect/count
        "project_bwd": 508 * N,              # 236 params + 36 ProjGrad read; 236 grads written
        "adam": 1652 * N,                    # p, g, m, v read; p, m, v written (59 x 4 B each)
        "raster_fwd": 52 * E + 20 * P,       # 48 B record + 4 B id per entry; image, T, n
        # record + 2 ids per entry, 36 B slot record (44 B with the AbsGS sums); dL, T, n
        "raster_bwd": 92 * E + 20 * P,
        "raster_bwd_ordered": 92 * E + 20 * P,
        "reduce_slots": 36 * E + 36 * N,     # slots read; ProjGrad written
        "loss_ssim_stats": 51 * P,           # pred 12 + target 3 read; 3 partial maps written
        "loss_grad": 63 * P,                 # 3 maps + pred + target read; dL/dimage written
        "bin_keygen": 12 * E + 16 * N,
        "bin_scatter": 32 * E,               # gaussian, tile, slot, depth read; key + gaussian written
        "bin_sort_tiles": 24 * E,            # 16-byte record read; entry + gaussian written
        "stats": 24 * N,
        # slots + params + both moments read; params and moments written (no gradient planes)
        "grad_adam": 36 * E + 1416 * N,
    }
    return table.get(name)


def learned_compare(iters, log=print, learned_runs=((0.2, 0.0, 0, 0.5),), plain_thresholds=(),
                    credit_views=0, credit_ratio=1.0, prune_labels=(0.0, 0)):
    """Full training with densification on the ellipsoid scene (300k surface points, 100
    cameras 1297x840 f = 1100 on a radius-6 hemisphere, targets from a 2M-Gaussian ground
    truth, `iters` iterations, densify every 100 from 500 to 5000 with threshold 2e-4,
    opacity reset every 3000), run twice from the same initialisation:
      plain   -- the reference's threshold heuristic (with the scorer logging and fitting
                 online; the model is bitwise the plain method's),
      learned -- the scorer trained in the first run makes the decisions.
    credit_views > 0 labels the logged clones / splits by credit (the image share of the
    Gaussians they added at the outcome, over that many views, against credit_ratio times
    the median Gaussian's share) instead of survival.  prune_labels = (bar, from): the
    learn run also labels as prune the rows whose share of all views' images is below bar
    x the median share when logged (events from iteration `from` on).  learned_runs:
    (threshold, share_prune, share_prune_from, prune_threshold) of each mode-2 run: the
    scorer's prune head applies from iteration share_prune_from on (threshold
    prune_threshold), and share_prune > 0 also prunes, from then on, the Gaussians whose
    share of all views' images is below that fraction of the median share (measured).
    Returns PSNR over all views, Gaussian counts and wall-clock iterations/s of each run."""
    import numpy as np
    from splatrs_amd import abi, scenes
    from splatrs_amd import records as rec
    ctx = abi.Context(0)
    t0 = time.perf_counter()
    g_gt, pts, rgb = scenes.ellipsoid_scene(0)
    cams = scenes.hemisphere_cameras()
    gt = abi.Trainer(ctx, g_gt)
    blank = np.zeros((840, 1297, 3), np.uint8)
    for c in cams:
        gt.add_view(c, blank)
    views = [abi.to_rgb8(gt.render(i, 3)) for i in range(len(cams))]
    gt.close()
    del g_gt
    init = ctx.from_points(pts, rgb)
    extent = scenes.scene_extent(cams)
    log(f"scene ready in {time.perf_counter() - t0:.1f} s (extent {extent:.3f})")
    until = min(5000, iters)
    cfg = dict(iterations=iters, densify=1, densify_from_iter=500, densify_until_iter=until,
               densification_interval=100, opacity_reset_interval=3000,
               densify_grad_threshold=0.0002, sh_degree=3, sh_degree_interval=1000,
               position_lr_max_steps=iters, log_interval=500, seed=0)
    scorer = abi.Scorer(ctx, seed=0)

    def run(mode, threshold=0.5, grad_threshold=0.0002, share_prune=0.0, share_prune_from=0,
            prune_threshold=0.5):
        tr = abi.Trainer(ctx, init)
        for c, v in zip(cams, views):
            tr.add_view(c, v)
        if mode:
            tr.attach_scorer(scorer, rec.learn_config(
                mode=mode, samples_per_event=32768, max_events=64, outcome_delay=500,
                eval_views=8, threshold=threshold, prune_threshold=prune_threshold,
                credit_views=credit_views,
                credit_ratio=credit_ratio, share_prune=share_prune,
                share_prune_from=share_prune_from,
                fit=rec.scorer_fit_config(epochs=2, batch=4096, lr=1e-3, seed=1)))
        tot = {"cloned": 0, "split": 0, "pruned": 0}

        def prog(pr):
            tot["cloned"] += pr.cloned
            tot["split"] += pr.split
            tot["pruned"] += pr.pruned
        t = time.perf_counter()
        tr.train(extent, abi.TrainConfig.default(**dict(cfg, densify_grad_threshold=grad_threshold)),
                 on_progress=prog)
        tr.sync()
        sec = time.perf_counter() - t
        psnr = float(np.mean([abi.psnr(tr.render(i, 3), v.astype(np.float32) / np.float32(255))
                              for i, v in enumerate(views)]))
        st = tr.learn_stats() if mode else np.zeros((), rec.LEARN_STATS)
        out = {"psnr_db": round(psnr, 3), "gaussians": int(tr.info()[0]), **tot,
               "it_per_s": round(iters / sec, 2), "seconds": round(sec, 2),
               "samples": int(st["resolved"]), "positives": int(st["positives"]),
               "fit_updates": int(st["updates"]), "fit_loss": round(float(st["last_loss"]), 4)}
        tr.close()
        log(f"mode {mode} threshold {threshold} grad threshold {grad_threshold} share prune "
            f"{share_prune}: {out}")
        return out
    # the reference's decisions; the scorer is trained on their outcomes
    plain = run(1, share_prune=prune_labels[0], share_prune_from=prune_labels[1])
    frontier = [dict(policy="learned", threshold=th,
                     **({"share_prune": sp} if sp else {}),
                     **({"prune_head_from": sf, "prune_threshold": pt} if 0 < sf < iters
                        else {"prune_head": "off"} if sf >= iters else {}),
                     **run(2, th, share_prune=sp, share_prune_from=sf, prune_threshold=pt))
                for th, sp, sf, pt in learned_runs]
    frontier += [dict(policy="plain", grad_threshold=gt, **run(0, grad_threshold=gt))
                 for gt in plain_thresholds]
    learned = frontier[0]
    ctx.close()
    grown = lambda r: r["cloned"] + r["split"]
    # the plain heuristic's PSNR at each learned run's Gaussian count, linear between the
    # plain runs (threshold 2e-4 and the extra ones) that bracket it
    pts = sorted([(plain["gaussians"], plain["psnr_db"])] +
                 [(r["gaussians"], r["psnr_db"]) for r in frontier if r["policy"] == "plain"])
    for r in frontier:
        if r["policy"] != "learned" or "share_prune" in r or "prune_head_from" in r:
            continue  # only runs whose decisions differ from the plain ones in clones alone
        if "prune_head" not in r and prune_labels[0] > 0:
            continue
        for (n0, p0), (n1, p1) in zip(pts, pts[1:]):
            if n0 <= r["gaussians"] <= n1 and n1 > n0:
                at = p0 + (p1 - p0) * (r["gaussians"] - n0) / (n1 - n0)
                r["plain_psnr_at_count_db"] = round(at, 3)
                r["gain_at_matched_count_db"] = round(r["psnr_db"] - at, 3)
    keys = ("policy", "threshold", "share_prune", "prune_head_from", "prune_threshold",
            "prune_head", "grad_threshold",
            "psnr_db", "gaussians",
            "plain_psnr_at_count_db", "gain_at_matched_count_db")
    return {"workload": "300k points on 20 ellipsoids -> densify, 100 cams 1297x840, "
                        f"{iters} iterations",
            "labels": (f"credit over {credit_views} views, bar {credit_ratio} x median share"
                       if credit_views else "survival"),
            "plain": plain, "learned": learned,
            "psnr_delta_db": round(learned["psnr_db"] - plain["psnr_db"], 3),
            "gaussian_ratio": round(learned["gaussians"] / max(plain["gaussians"], 1), 4),
            "densified_ratio": round(grown(learned) / max(grown(plain), 1), 4),
            "frontier": [{k: r[k] for k in keys if k in r} for r in frontier]}


def learned_runs(spec):
    """'threshold[:share_prune[:from[:prune_threshold]]],...' -> [(threshold, share_prune,
    from, prune_threshold), ...]."""
    out = []
    for item in spec.split(","):
        f = item.split(":")
        out.append((float(f[0]), float(f[1]) if len(f) > 1 else 0.0,
                    int(f[2]) if len(f) > 2 else 0, float(f[3]) if len(f) > 3 else 0.5))
    return out


def prune_labels(spec):
    f = spec.split(":")
    return float(f[0]), int(f[1]) if len(f) > 1 else 0


def densify_bench(n, rounds, log=print):
    """Density control at scale: one densify_and_prune (src/densify.rs:125-221) of n
    Gaussians (the 1M-Gaussian scene) with densification statistics drawn so that about
    12% of the rows pass the 2e-4 gradient threshold (split or cloned by their size) and
    about 3% are pruned, restored before each of `rounds` timed rounds.  Reports the mean
    time of the compaction kernels and their algorithmic bytes: densify_flags reads the
    rows' log-scales, opacity and statistics (28 B) and writes 5 flags (20 B); the
    compaction (densify_copy for kept rows and clones, densify_children for split rows,
    reported together as densify_scatter) reads parameters and both moments of every row
    (708 B) and writes them for every output row (708 B x N_out)."""
    import numpy as np
    from splatrs_amd import abi, scenes
    from splatrs_amd import records as rec
    ctx = abi.Context(0)
    g = scenes.large_scene(0, n)
    rng = np.random.default_rng(5)
    denom = rng.integers(1, 60, n).astype(np.uint32)
    mean_grad = np.exp(rng.normal(np.log(1.2e-4), 0.55, n)).astype(np.float32)
    state = {"params": g, "grad_accum": (mean_grad * denom).astype(np.float32), "denom": denom,
             "max_radii": rng.integers(0, 40, n).astype(np.uint32)}
    g[rng.random(n) < 0.03, 10] = -7.0  # opacity below the 0.005 prune limit
    tr = abi.Trainer(ctx, g)
    extent = 6.0
    cfg = rec.densify_config(grad_threshold=0.0002)
    tr.restore(state, 1)
    tr.densify(cfg, extent, 1)  # warm-up: sizes the output planes
    ctx.set_timing(True)
    reports = []
    walls = []
    for r in range(rounds):
        tr.restore(state, 1)
        ctx.reset_timing()
        t0 = time.perf_counter()
        rep = tr.densify(cfg, extent, 2 + r)
        walls.append(time.perf_counter() - t0)
        times = ctx.kernel_times()
        reports.append((rep, times))
    ctx.set_timing(False)
    n_out = int(tr.info()[0])
    rep = reports[-1][0]
    out = {"workload": f"densify_and_prune of {n} Gaussians (1M-Gaussian scene, drawn "
                       "statistics), threshold 2e-4, extent 6",
           "n_in": n, "n_out": n_out, "cloned": int(rep["cloned"]), "split": int(rep["split"]),
           "pruned": int(rep["pruned"]),
           "call_ms": round(float(np.mean(walls)) * 1e3, 3), "kernels": {}}
    alg = {"densify_flags": 48 * n, "densify_scatter": 708 * n + 708 * n_out}
    for _, t in reports:  # the compaction's two kernels as one
        parts = [t[k] for k in ("densify_copy", "densify_children") if k in t]
        t["densify_scatter"] = (1, sum(p[1] / max(p[0], 1) for p in parts))
    for name in sorted(reports[-1][1]):
        ms = np.mean([t[name][1] / max(t[name][0], 1) for _, t in reports if name in t])
        k = {"mean_ms": round(float(ms), 4),
             "launches": int(reports[-1][1][name][0])}
        if name in alg:
            k["algorithmic_bytes"] = alg[name]
            k["gbs"] = round(alg[name] / (ms * 1e-3) / 1e9, 1)
            k["frac_of_hbm_peak"] = round(alg[name] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)
        out["kernels"][name] = k
    tr.close()
    ctx.close()
    log(f"densify: {out}")
    return out


def batched_bench(steps, log=print):
    """Batched configuration: 16 scenes of 50k-800k Gaussians (log-uniform; the small
    configuration's generator) and 64 views (4 per scene, 320x240 .. 1600x1200) trained
    in one graph-captured step, launch shapes autotuned.  Also times the same 64 views as
    single-view trainer steps for comparison."""
    import numpy as np
    from splatrs_amd import abi, scenes
    from splatrs_amd import records as rec
    ctx = abi.Context(0)
    t0 = time.perf_counter()
    sizes = scenes.batch_scene_sizes(0)
    pairs = scenes.batch_views(0)
    models = [scenes.random_gaussians(np.random.default_rng(100 + s), n)
              for s, n in enumerate(sizes)]
    batch = abi.Batch(ctx)
    for g in models:
        batch.add_scene(g)
    for s, n in enumerate(sizes):  # targets from a second seed of each scene
        gt = abi.Trainer(ctx, scenes.random_gaussians(np.random.default_rng(200 + s), n))
        mine = [(k, c) for k, (ps, c) in enumerate(pairs) if ps == s]
        for _, c in mine:
            gt.add_view(c, np.zeros((int(c["height"]), int(c["width"]), 3), np.uint8))
        for j, (k, c) in enumerate(mine):
            batch.add_view(s, c, abi.to_rgb8(gt.render(j, 3)))
        gt.close()
    extents = [scenes.scene_extent([c for ps, c in pairs if ps == s]) for s in range(len(sizes))]
    views = list(range(len(pairs)))
    pixels = sum(int(c["width"]) * int(c["height"]) for _, c in pairs)
    log(f"batched scenes ready in {time.perf_counter() - t0:.1f} s: {sum(sizes)} Gaussians, "
        f"{len(pairs)} views, {pixels / 1e6:.1f} MP")
    lr = rec.learning_rates(0.0)
    p = rec.batch_step_params(3, lr)
    tune = batch.tune(views, p)
    it = 0

    def lr_mean(it):
        return [rec.position_lr(it, e) for e in extents]
    for _ in range(2):
        it += 1
        batch.step(views, p, lr_mean(it), want_losses=False)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        it += 1
        batch.step(views, p, lr_mean(it), want_losses=False)
    ctx.sync()
    sec = time.perf_counter() - t
    info = batch.info()
    batch.close()
    # the same views as single-view trainer steps (one trainer per scene)
    single_sec = 0.0
    for s, n in enumerate(sizes):
        tr = abi.Trainer(ctx, models[s])
        mine = [c for ps, c in pairs if ps == s]
        for c in mine:
            tr.add_view(c, np.zeros((int(c["height"]), int(c["width"]), 3), np.uint8))
        for j in range(len(mine)):
            tr.step(j, rec.step_params(3, rec.learning_rates(1e-4)), result=False)
        tr.sync()
        t = time.perf_counter()
        for _ in range(2):
            for j in range(len(mine)):
                tr.step(j, rec.step_params(3, rec.learning_rates(1e-4)), result=False)
        tr.sync()
        single_sec += (time.perf_counter() - t) / 2
        tr.close()
    ctx.close()
    return {"workload": f"{len(sizes)} scenes ({min(sizes)}-{max(sizes)} Gaussians, "
                        f"{sum(sizes)} total), {len(pairs)} views 320x240..1600x1200 "
                        f"({pixels / 1e6:.1f} MP) per step",
            "steps_per_s": round(steps / sec, 3), "ms_per_step": round(sec / steps * 1e3, 3),
            "views_per_s": round(len(pairs) * steps / sec, 1),
            "single_view_steps_views_per_s": round(len(pairs) / single_sec, 1),
            "tuned_shapes": {"fwd_batch": int(tune["fwd_batch"]),
                             "fwd_persistent": bool(tune["fwd_persistent"]),
                             "fwd_per_cu": int(tune["fwd_per_cu"]),
                             "bwd_batch": int(tune["bwd_batch"]),
                             "bwd_persistent": bool(tune["bwd_persistent"]),
                             "bwd_per_cu": int(tune["bwd_per_cu"]),
                             "bwd_row_lists": bool(tune["bwd_row_lists"]),
                             "fwd_ms": round(float(tune["fwd_ms"]), 3),
                             "bwd_ms": round(float(tune["bwd_ms"]), 3),
                             "default_fwd_ms": round(float(tune["default_fwd_ms"]), 3),
                             "default_bwd_ms": round(float(tune["default_bwd_ms"]), 3)},
            "graph": {"captures": info["captures"], "replays": info["replays"],
                      "aborted_steps": info["aborts"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ordered", action="store_true",
                    help="reference summation order in the backward pass")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="oracle train steps timed for the CPU baseline (0 = skip)")
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (reference kernel of "
                         "known traffic for the HBM byte counters) and exit")
    ap.add_argument("--learned-only", action="store_true",
                    help="only run the plain-vs-learned densification comparison")
    ap.add_argument("--learned-iters", type=int, default=7000)
    ap.add_argument("--learned-runs", default="0.2:0:4500:0.4,0.35:0:100000",
                    help="learned runs, comma separated, each threshold[:share_prune[:from "
                         "iteration[:prune_threshold]]]: the scorer's clone / split "
                         "threshold, the measured image-share pruning bar (fraction of the "
                         "median share; 0: off), the iteration from which pruning (measured "
                         "and by the scorer's prune head) applies, the prune head's threshold")
    ap.add_argument("--prune-labels", default="0.15:2500",
                    help="bar:from -- the learn run labels rows below bar x the median image "
                         "share as prune (from that iteration; bar 0: the reference's prunes)")
    ap.add_argument("--plain-thresholds", default="0.000225,0.00025",
                    help="extra plain runs at these densify gradient thresholds")
    ap.add_argument("--credit-views", type=int, default=8,
                    help="label the scorer's samples by image-share credit over this many "
                         "views (0: by survival)")
    ap.add_argument("--credit-ratio", type=float, default=1.0,
                    help="credit bar, as a fraction of the median Gaussian's share")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the batched many-scene configuration")
    ap.add_argument("--densify-only", action="store_true",
                    help="only time density control (compaction) at 1M Gaussians")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched and plain-vs-learned sections of the JSON line")
    args = ap.parse_args()
    if args.densify_only:
        r = densify_bench(args.gaussians, 5, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(r), flush=True)
        return
    if args.batched_only:
        r = batched_bench(args.batched_steps, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(r), flush=True)
        return
    if args.learned_only:
        r = learned_compare(args.learned_iters,
                            log=lambda m: print(m, file=sys.stderr, flush=True),
                            learned_runs=learned_runs(args.learned_runs),
                            plain_thresholds=[float(x) for x in args.plain_thresholds.split(",") if x],
                            credit_views=args.credit_views, credit_ratio=args.credit_ratio,
                            prune_labels=prune_labels(args.prune_labels))
        print(json.dumps(r), flush=True)
        return
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, device="cuda:0", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 4, "copies": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the training path runs on one GPU; a sharded version is not built)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from splatrs_amd import abi, scenes
    from splatrs_amd import records as rec

    ctx = abi.Context(0)
    g = scenes.large_scene(0, args.gaussians)
    cam = scenes.hd_camera()
    W, H = int(cam["width"]), int(cam["height"])
    gt = scenes.large_scene(1, args.gaussians)
    target = abi.to_rgb8(ctx.render(gt, cam, rec.proj_opts(3)))
    del gt
    trainer = abi.Trainer(ctx, g)
    view = trainer.add_view(cam, target)
    extent = scenes.scene_extent([cam])

    def params(it, want_loss=False):
        lr = rec.learning_rates(rec.position_lr(it, extent))
        return rec.step_params(3, lr, ordered_backward=args.ordered, want_loss=want_loss)

    it = 0
    info = None
    for _ in range(max(args.warmup, 1)):
        it += 1
        r = trainer.step(view, params(it, want_loss=(info is None)))
        if info is None:
            info = dict(loss0=float(r["loss"]), entries=int(r["num_entries"]),
                        visible=int(r["num_visible"]))
    ctx.sync()

    # autotuned compositing launch shapes (results do not depend on them)
    tune = trainer.tune(view, params(it + 1))
    tuned = {"fwd_batch": int(tune["fwd_batch"]), "fwd_persistent": bool(tune["fwd_persistent"]),
             "fwd_per_cu": int(tune["fwd_per_cu"]), "bwd_batch": int(tune["bwd_batch"]),
             "bwd_persistent": bool(tune["bwd_persistent"]), "bwd_per_cu": int(tune["bwd_per_cu"]),
             "bwd_row_lists": bool(tune["bwd_row_lists"]),
             "fwd_ms": round(float(tune["fwd_ms"]), 4), "bwd_ms": round(float(tune["bwd_ms"]), 4),
             "default_fwd_ms": round(float(tune["default_fwd_ms"]), 4),
             "default_bwd_ms": round(float(tune["default_bwd_ms"]), 4)}

    def timed_pass():
        nonlocal it
        trainer.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            it += 1
            trainer.step(view, params(it), result=False)
        trainer.sync()
        return time.perf_counter() - t0

    # graph-captured steps (the default mode): the measured value
    sec = timed_pass()
    value = args.steps / sec
    graph_stats = trainer.graph_stats()
    # the same kernels launched one by one (for comparison)
    trainer.set_graph(False)
    eager_sec = timed_pass()

    # per-kernel attribution (separate eager pass: events add a little launch overhead)
    ctx.set_timing(True)
    ctx.reset_timing()
    for _ in range(args.steps):
        it += 1
        trainer.step(view, params(it), result=False)
    times = ctx.kernel_times()
    ctx.set_timing(False)
    trainer.set_graph(True)
    r = trainer.step(view, params(it + 1, want_loss=True))
    N, V, E, P = args.gaussians, int(r["num_visible"]), int(r["num_entries"]), W * H

    kernels = {}
    for name, (launches, total_ms) in times.items():
        per_step = total_ms / args.steps
        kernels[name] = {"ms_per_step": round(per_step, 4),
                         "launches_per_step": launches / args.steps,
                         "mean_ms": round(total_ms / max(launches, 1), 4)}
        b = algorithmic_bytes(name, N, V, E, P)
        if b is not None and launches:
            kernels[name]["gbs"] = round(b / (total_ms / launches * 1e-3) / 1e9, 1)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    roof_name = dom if algorithmic_bytes(dom, N, V, E, P) is not None else max(
        (k for k in kernels if algorithmic_bytes(k, N, V, E, P) is not None),
        key=lambda k: kernels[k]["ms_per_step"])
    b = algorithmic_bytes(roof_name, N, V, E, P)
    mean_s = kernels[roof_name]["mean_ms"] * 1e-3
    achieved = b / mean_s / 1e9
    # HBM bytes per launch measured by the rocprofv3 FETCH_SIZE / WRITE_SIZE passes on this
    # workload (profiles/), when they cover this kernel.
    traffic = None
    valu = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_1M_1080p.json")
    if args.gaussians == 1_000_000 and os.path.exists(tpath):
        t = json.load(open(tpath))["kernels"].get(roof_name)
        if t:
            traffic = int(t["read_bytes"] + t["write_bytes"])
            if "valu_insts" in t and "valu_dual_quadcycles" in t:
                # The compositing kernels are bound by instruction issue, not by HBM: wave64
                # VALU instructions per launch (rocprofv3 SQ_INSTS_VALU on this workload),
                # of which some issue as a second instruction in a quad-cycle
                # (SQ_ACTIVE_INST_VALU2).  busy_frac = quad-cycles that issued a VALU
                # instruction / (256 CUs x 4 SIMDs x 2.4 GHz x launch time), at most 1.
                rate = t["valu_insts"] / mean_s / 1e9
                single = 1.0 - t["valu_dual_quadcycles"] / t["valu_insts"]
                valu = {"insts_per_launch": int(t["valu_insts"]),
                        "dual_issue_quadcycles_per_launch": int(t["valu_dual_quadcycles"]),
                        "busy_frac": round(rate * single / VALU_PEAK_GINST, 3)}
    # exact (Gaussian, pixel) pair counts of this view (counting kernels, outside the timed
    # region): the compositing kernels' FP32 roofline
    pairs = trainer.pair_counts(view, params(it + 1))
    evaluated = {"raster_fwd": pairs["fwd_evaluated"], "raster_bwd": pairs["bwd_evaluated"]}
    for name, ops in FLOPS_PER_PAIR.items():
        if name in kernels:
            tf = evaluated[name] * ops / (kernels[name]["mean_ms"] * 1e-3) / 1e12
            kernels[name]["fp32"] = {"evaluated_pairs": evaluated[name], "ops_per_pair": ops,
                                     "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS,
                                     "unit": "TFLOP/s", "frac": round(tf / FP32_PEAK_TFLOPS, 4)}
    roofline = {"kernel": roof_name, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": int(b),
                "mean_ms": kernels[roof_name]["mean_ms"],
                "fp32": kernels[roof_name].get("fp32"), "valu_issue": valu}

    cpu = None
    if args.cpu_steps > 0:
        # The oracle (C++ restatement of the reference, compiled like its release build)
        # with its loops parallel along the reference's rayon axes on every host core; a
        # bounded sample of full train steps of the same workload.  One step on one core
        # is reported beside it.
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        cores = O.set_threads(0)
        st = O.TrainState(g)
        lr = rec.learning_rates(rec.position_lr(1, extent))
        st.train_step(cam, target, 3, lr)  # warm-up (page-in, allocations)
        t0 = time.perf_counter()
        for k in range(args.cpu_steps):
            st.train_step(cam, target, 3, lr)
        cs = time.perf_counter() - t0
        O.set_threads(1)
        t1 = time.perf_counter()
        st.train_step(cam, target, 3, lr)
        c1 = time.perf_counter() - t1
        O.set_threads(cores)
        cpu = {"value": round(args.cpu_steps / cs, 4), "unit": "it/s", "cores": cores,
               "kind": "port",
               "sample": f"{args.cpu_steps} full train steps of the same 1M/1080p workload "
                         f"on {cores} threads ({cs:.1f} s); C++ oracle, OpenMP over the "
                         "reference's rayon axes",
               "one_core_it_per_s": round(1.0 / c1, 4)}

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "it/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sec / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1M Gaussians, 1920x1080 single-view train step (plain method)",
                   "gaussians": N, "width": W, "height": H, "tile": 16, "sh_degree": 3,
                   "visible": V, "tile_entries": E,
                   "backward": "ordered" if args.ordered else "tree-reduced",
                   "loss_at_step1": round(info["loss0"], 6),
                   "step_mode": "hipGraph replay",
                   "eager_it_per_s": round(args.steps / eager_sec, 3),
                   "graph_captures_replays_aborts": list(graph_stats),
                   "tuned_shapes": tuned, "pair_counts": pairs},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    if not args.no_extras:
        trainer.close()
        ctx.close()
        err = lambda m: print(m, file=sys.stderr, flush=True)
        for key, fn in (("batched", lambda: batched_bench(args.batched_steps, err)),
                        ("learned", lambda: learned_compare(
                            args.learned_iters, err, learned_runs(args.learned_runs),
                            [float(x) for x in args.plain_thresholds.split(",") if x],
                            args.credit_views, args.credit_ratio,
                            prune_labels(args.prune_labels)))):
            try:
                out[key] = fn()
            except Exception as e:  # the headline line is printed regardless
                out[key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
