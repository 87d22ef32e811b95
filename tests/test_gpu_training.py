"""Training loop and initialisation on the GPU (needs a gfx950 GPU).

- from_points / mean_nn_dist2 against the oracle (bitwise, incl. coincident points);
- ports of the reference's end-to-end training tests (tests/pipeline.rs:17-130): PSNR
  gains without and with densification, report and save counts;
- splat_train against a Python restatement of src/train.rs:196-303 driving single steps,
  densify and reset (bitwise, ordered backward);
- the 10k-Gaussian / 8-view 256x256 configuration against the oracle's train step:
  identical model after the run (ordered backward) and the same PSNR.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

import oracle_lib as O
from helpers import ground_truth, init_points, ring_cameras
from splatrs_amd import abi, scenes
from splatrs_amd import records as rec

pytestmark = pytest.mark.gpu
LIBM = ctypes.CDLL("libm.so.6")
LIBM.logf.restype = ctypes.c_float
LIBM.logf.argtypes = [ctypes.c_float]
MASK = (1 << 64) - 1


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


def test_from_points_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 4, 200, 5000):
        say(f"mean_nn_dist2 n={n}")
        pts = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        if n >= 200:
            pts[10:50] = pts[5]          # coincident points
            pts[60:62] = pts[61]
        got = ctx.mean_nn_dist2(pts)
        want = O.mean_nn_dist2(pts.astype(np.float64))
        assert np.array_equal(got, want)
    rgb = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    g = ctx.from_points(pts, rgb)
    f = np.float32
    for i in range(0, n, 97):
        d2 = max(f(want[i]), f(1e-7))
        assert g[i, 3] == LIBM.logf(f(np.sqrt(d2))) and g[i, 3] == g[i, 5]
        assert g[i, 10] == LIBM.logf(f(0.1) / (f(1.0) - f(0.1)))
        assert list(g[i, 6:10]) == [1, 0, 0, 0] and not g[i, 14:].any()
        for c in range(3):
            assert g[i, 11 + c] == (f(rgb[i, c]) / f(255.0) - f(0.5)) / f(0.28209479177387814)


def render_views(ctx, gt, cams, bg, degree):
    return [abi.to_rgb8(ctx.render(gt, c, rec.proj_opts(degree), bg)) for c in cams]


def evaluate(ctx, model, cams, views, degree, bg):
    """evaluate, src/train.rs:309-321."""
    tot = 0.0
    for c, v in zip(cams, views):
        img = ctx.render(model, c, rec.proj_opts(degree), bg)
        tot += abi.psnr(img, v.astype(np.float32) / np.float32(255.0))
    return tot / len(cams)


def test_synthetic_scene_trains_without_densification(ctx):
    """tests/pipeline.rs:17-60."""
    rng = np.random.default_rng(100)
    gt = ground_truth(rng, 40)
    cams = ring_cameras(10, 64, 48, 2.6)
    views = render_views(ctx, gt, cams, (0, 0, 0), 1)
    extent = scenes.scene_extent(cams)
    init = ctx.from_points(*init_points(gt, rng))
    cfg = abi.TrainConfig.default(iterations=600, densify=0, sh_degree=1, sh_degree_interval=200,
                                  position_lr_max_steps=600, log_interval=0, seed=3)
    before = evaluate(ctx, init, cams, views, 0, (0, 0, 0))
    tr = abi.Trainer(ctx, init)
    for c, v in zip(cams, views):
        tr.add_view(c, v)
    saves = []
    tr.train(extent, cfg, on_save=lambda it, t: saves.append((it, t.info()[0])))
    assert saves == [(600, 40)]
    after = evaluate(ctx, tr.download()["params"], cams, views, 1, (0, 0, 0))
    say(f"no densification: PSNR {before:.2f} -> {after:.2f} dB")
    assert after > before + 6.0 and after > 24.0


def test_synthetic_scene_trains_with_densification(ctx):
    """tests/pipeline.rs:62-130."""
    rng = np.random.default_rng(200)
    gt = ground_truth(rng, 30)
    cams = ring_cameras(8, 48, 36, 2.6)
    views = render_views(ctx, gt, cams, (1, 1, 1), 1)
    extent = scenes.scene_extent(cams)
    xyz, rgb = init_points(gt, rng)
    init = ctx.from_points(xyz[:10], rgb[:10])
    cfg = abi.TrainConfig.default(iterations=500, densify=1, densify_from_iter=40,
                                  densify_until_iter=400, densification_interval=40,
                                  opacity_reset_interval=200, densify_grad_threshold=0.0005,
                                  white_background=1, sh_degree=1, sh_degree_interval=150,
                                  position_lr_max_steps=500, log_interval=50,
                                  save_iterations=[250], seed=5)
    before = evaluate(ctx, init, cams, views, 0, (1, 1, 1))
    tr = abi.Trainer(ctx, init)
    for c, v in zip(cams, views):
        tr.add_view(c, v)
    reports, saves = [], []
    tr.train(extent, cfg, on_progress=lambda p: reports.append((p.iteration, p.loss,
                                                                   p.num_gaussians)),
             on_save=lambda it, t: saves.append((it, t.info()[0])))
    assert len(reports) == 10 and all(np.isfinite(r[1]) for r in reports)
    assert [s[0] for s in saves] == [250, 500]
    trained = tr.download()["params"]
    assert len(trained) > 10
    after = evaluate(ctx, trained, cams, views, 1, (1, 1, 1))
    say(f"with densification: PSNR {before:.2f} -> {after:.2f} dB, 10 -> {len(trained)}")
    assert after > before + 4.0
    assert np.isfinite(trained[:, :10]).all()


def sm64(state):
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def position_lr(it, extent, max_steps):
    """TrainConfig::position_lr with glibc logf/expf."""
    f = np.float32
    LIBM.expf.restype = ctypes.c_float
    LIBM.expf.argtypes = [ctypes.c_float]
    t = min(max(f(f(it) / f(max(max_steps, 1))), f(0)), f(1))
    a = LIBM.logf(f(f(0.00016) * f(extent)))
    b = LIBM.logf(f(f(0.0000016) * f(extent)))
    return LIBM.expf(f(f(a * f(f(1) - t)) + f(b * t)))


def test_train_loop_matches_restated_schedule(ctx):
    rng = np.random.default_rng(7)
    gt = ground_truth(rng, 60)
    cams = ring_cameras(6, 56, 40, 2.6)
    views = render_views(ctx, gt, cams, (0, 0, 0), 1)
    extent = scenes.scene_extent(cams)
    init = ctx.from_points(*init_points(gt, rng))
    kw = dict(iterations=160, densify=1, densify_from_iter=30, densify_until_iter=140,
              densification_interval=20, opacity_reset_interval=60,
              densify_grad_threshold=0.0002, sh_degree=2, sh_degree_interval=50,
              position_lr_max_steps=160, log_interval=0, seed=11, ordered_backward=1)
    cfg = abi.TrainConfig.default(**kw)
    a = abi.Trainer(ctx, init)
    b = abi.Trainer(ctx, init)
    for c, v in zip(cams, views):
        a.add_view(c, v)
        b.add_view(c, v)
    say("splat_train")
    a.train(extent, cfg)
    say("restated schedule over single steps")
    sh = 0
    for it in range(1, kw["iterations"] + 1):
        if it % kw["sh_degree_interval"] == 0 and sh < kw["sh_degree"]:
            sh += 1
        view = abi.lib().splat_train_view_at(kw["seed"], len(cams), it)
        window = it < kw["densify_until_iter"]
        lr = rec.learning_rates(position_lr(it, extent, kw["position_lr_max_steps"]))
        b.step(view, rec.step_params(sh, lr, accumulate_stats=window, ordered_backward=True))
        if window:
            if it > kw["densify_from_iter"] and it % kw["densification_interval"] == 0:
                cfgd = rec.densify_config(grad_threshold=kw["densify_grad_threshold"],
                                          prune_big_world=it > kw["opacity_reset_interval"])
                seed = sm64((kw["seed"] ^ ((0x5851F42D4C957F2D * it) & MASK)) & MASK)
                b.densify(cfgd, extent, seed)
            if it % kw["opacity_reset_interval"] == 0:
                b.reset_opacity(0.01)
    pa, pb = a.download(moments=True), b.download(moments=True)
    assert pa["params"].shape == pb["params"].shape and len(pa["params"]) != 60
    for k in ("params", "m", "v"):
        assert np.array_equal(pa[k].view(np.uint32), pb[k].view(np.uint32)), k


def test_train_resumes_from_a_checkpoint(ctx, tmp_path):
    """splat_train to 160 iterations (SH steps at 50 and 100, densification every 20 from
    30 to 140, opacity resets at 60 and 120) against a run whose state after iteration 90
    is checkpointed from on_save, restored into a new trainer and trained on from 91:
    model and moments equal the uninterrupted run's bitwise."""
    rng = np.random.default_rng(7)
    gt = ground_truth(rng, 60)
    cams = ring_cameras(6, 56, 40, 2.6)
    views = render_views(ctx, gt, cams, (0, 0, 0), 1)
    extent = scenes.scene_extent(cams)
    init = ctx.from_points(*init_points(gt, rng))
    kw = dict(iterations=160, densify=1, densify_from_iter=30, densify_until_iter=140,
              densification_interval=20, opacity_reset_interval=60,
              densify_grad_threshold=0.0002, sh_degree=2, sh_degree_interval=50,
              position_lr_max_steps=160, log_interval=0, seed=11, ordered_backward=1)

    def trainer(model):
        tr = abi.Trainer(ctx, model)
        for c, v in zip(cams, views):
            tr.add_view(c, v)
        return tr
    say("uninterrupted splat_train")
    a = trainer(init)
    a.train(extent, abi.TrainConfig.default(**kw))
    want = a.download(moments=True, stats=True)
    want_step = a.info()[1]
    a.close()
    say("splat_train with a checkpoint written at iteration 90")
    path = tmp_path / "run.ckpt"

    def on_save(it, tr):
        if it == 90:
            abi.checkpoint_write(path, tr.download(moments=True, stats=True), tr.info()[1], 1)
    b = trainer(init)
    b.train(extent, abi.TrainConfig.default(save_iterations=[90], **kw), on_save=on_save)
    b.close()
    saved, step, _ = abi.checkpoint_read(path)
    assert step == 90
    say("restore into a new trainer, splat_train from iteration 91")
    c = trainer(init[:10])
    c.restore(saved, step)
    c.train(extent, abi.TrainConfig.default(start_iteration=91, **kw))
    got = c.download(moments=True, stats=True)
    assert c.info()[1] == want_step
    c.close()
    for k in want:
        assert got[k].shape == want[k].shape, k
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k
    with pytest.raises(Exception):
        d = trainer(init)
        try:
            d.train(extent, abi.TrainConfig.default(start_iteration=161, **kw))
        finally:
            d.close()


def test_small_config_matches_oracle(ctx):
    """The whole small configuration: 10k Gaussians (seed 0), 8 views 256x256 at radius 4,
    targets from seed 1, 500 iterations of L1 + 0.2 D-SSIM with Adam, no densification
    (SH degree 0 throughout: the degree steps up every 1000 iterations)."""
    iters = 500
    g = scenes.small_scene(0)
    cams = scenes.small_scene_cameras()
    views = render_views(ctx, scenes.small_scene(1), cams, (0, 0, 0), 3)
    extent = scenes.scene_extent(cams)
    cfg = abi.TrainConfig.default(iterations=iters, densify=0, position_lr_max_steps=30000,
                                  log_interval=0, seed=0, ordered_backward=1)
    tr = abi.Trainer(ctx, g)
    for c, v in zip(cams, views):
        tr.add_view(c, v)
    say("small configuration on the GPU")
    tr.train(extent, cfg)
    gpu = tr.download(moments=True)
    say("small configuration on the oracle")
    st = O.TrainState(g)
    for it in range(1, iters + 1):
        view = abi.lib().splat_train_view_at(0, len(cams), it)
        lr = rec.learning_rates(position_lr(it, extent, 30000))
        st.train_step(cams[view], views[view], 0, lr)
    same = np.mean(gpu["params"].view(np.uint32) == st.params.view(np.uint32))
    say(f"bitwise-equal parameters after {iters} iterations: {same:.6f}")
    pg = evaluate(ctx, gpu["params"], cams, views, 0, (0, 0, 0))
    po = evaluate(ctx, st.params, cams, views, 0, (0, 0, 0))
    p0 = evaluate(ctx, g, cams, views, 0, (0, 0, 0))
    say(f"PSNR gpu {pg:.3f} dB, oracle {po:.3f} dB (initial {p0:.3f} dB)")
    assert abs(pg - po) <= 0.01
    assert pg > p0 + 1.0
    assert same >= 0.999
