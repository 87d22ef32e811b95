"""The batched and densifying configurations at their own sizes (needs a gfx950 GPU).

Batched (16 scenes of 50k-800k Gaussians, 64 views of 320x240 .. 1600x1200, one step in
the reference's summation order): the smallest and the largest scene equal the oracle
composed view by view (each view's gradient, src/train.rs:211-240, summed per scene in
step order, then Adam, src/adam.rs:44-61) bitwise; on every scene, only Gaussians some
view reaches change, and the per-view losses are within 2e-6 of the binary64 sums of the
oracle's loss terms.

Densifying (300k points sampled on 20 ellipsoids, 100 cameras 1297x840 at radius 6,
targets rendered from a 2M-Gaussian ground truth): splat_train runs iterations 1-599
(statistics accumulating, no density control before iteration 600), iteration 600 is
stepped as splat_train steps it, and the first densification (src/densify.rs:125-221) on
the resident model equals the oracle's on the downloaded state.  At the same event the
learned scorer's 12 features equal the oracle's restatement and its BF16-MFMA logits
match the BF16 oracle MLP.
"""
import math
import sys

import numpy as np
import pytest

import oracle_lib as O
from splatrs_amd import abi, scenes
from splatrs_amd import records as rec

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


def splitmix64(x):
    """SplitMix64 on uint64 arrays (wrapping arithmetic)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return x ^ (x >> np.uint64(31))


def trainer_normals(seed, count):
    """The trainer's split-child normal draws: SplitMix64 counter stream and Box-Muller in
    binary64 (the loop form is test_gpu_trainer.normals)."""
    with np.errstate(over="ignore"):
        i = np.arange(count, dtype=np.uint64)
        s = np.uint64(seed)
        a = splitmix64(s ^ splitmix64(np.uint64(2) * i))
        b = splitmix64(s ^ splitmix64(np.uint64(2) * i + np.uint64(1)))
    u1 = np.maximum((a >> np.uint64(11)).astype(np.float64) * 2.0 ** -53, 1e-12)
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * 3.141592653589793 * u2)).astype(np.float32)


def test_trainer_normals_vectorised_form():
    from test_gpu_trainer import normals
    assert np.array_equal(trainer_normals(1234, 50), normals(1234, 50))


# ---- batched configuration -------------------------------------------------------------

def oracle_view(g, cam, target, degree):
    """One view's parameter gradient in the oracle (project -> bin -> composite -> loss ->
    backward) and its loss as the binary64 sum of the binary32 SSIM and L1 terms: the
    oracle's own scalar is a sequential binary32 sum, which at a million terms is off by
    several percent (the reference leaves the order to rayon's parallel sum)."""
    opts = rec.proj_opts(degree)
    W, H = int(cam["width"]), int(cam["height"])
    projs = O.project_all(g, cam, opts)
    binning = O.bin(projs, W, H)
    settings = rec.raster_settings(W, H, (0.0, 0.0, 0.0))
    image, final_t, n_contrib = O.rasterize_forward(projs, binning, settings)
    tgt = target.astype(np.float32) / np.float32(255.0)
    _, dimage = O.photometric_loss(image, tgt, W, H, 0.2)
    ssim, l1 = O.loss_sums64(image, tgt, W, H)
    n = W * H * 3
    loss = 0.8 * (l1 / n) + 0.2 * (1.0 - ssim / n)
    pgrads, _ = O.rasterize_backward(projs, binning, settings, final_t, n_contrib, dimage)
    return loss, O.project_backward_all(g, cam, opts, pgrads)


def test_batched_full_size_step_matches_oracle(ctx):
    sizes = scenes.batch_scene_sizes(0)
    pairs = scenes.batch_views(0)
    say(f"batched: {len(sizes)} scenes, {sum(sizes)} Gaussians, {len(pairs)} views")
    models = [scenes.random_gaussians(np.random.default_rng(100 + s), n)
              for s, n in enumerate(sizes)]
    targets = []
    for k, (s, c) in enumerate(pairs):
        gt = scenes.random_gaussians(np.random.default_rng(200 + s), sizes[s])
        targets.append(abi.to_rgb8(ctx.render(gt, c, rec.proj_opts(3))))
    b = abi.Batch(ctx)
    for g in models:
        b.add_scene(g)
    for (s, c), t in zip(pairs, targets):
        b.add_view(s, c, t)
    order = list(range(len(pairs)))
    lr = rec.learning_rates(0.0)
    lr_mean = np.array([rec.position_lr(1, scenes.scene_extent(
        [c for ps, c in pairs if ps == s])) for s in range(len(sizes))], np.float32)
    say("one batched step, reference summation order")
    losses = b.step(order, rec.batch_step_params(3, lr, ordered_backward=True), lr_mean)
    assert np.isfinite(losses).all()
    small, large = int(np.argmin(sizes)), int(np.argmax(sizes))
    for s in range(len(sizes)):
        got = b.download(s, moments=True)
        changed = (got["params"] != models[s]).any(axis=1)
        say(f"scene {s}: {sizes[s]} Gaussians, {int(changed.sum())} changed")
        assert np.isfinite(got["params"]).all() and changed.any()
        if s not in (small, large):
            continue
        total = None
        for k in order:
            if pairs[k][0] != s:
                continue
            loss, gv = oracle_view(models[s], pairs[k][1], targets[k], 3)
            assert abs(losses[k] - loss) <= 2e-6 * abs(loss), (k, losses[k], loss)
            total = gv.copy() if total is None else total + gv
        params, m, v = models[s].copy(), np.zeros_like(models[s]), np.zeros_like(models[s])
        lrs = lr.copy()
        lrs["mean"] = lr_mean[s]
        O.adam_step(params, total, m, v, rec.adam_state(), lrs)
        # untouched rows: a zero gradient leaves parameters and moments at their values
        reached = (total != 0).any(axis=1)
        assert not changed[~reached].any()
        for key, want in (("params", params), ("m", m), ("v", v)):
            same = np.mean(got[key].view(np.uint32) == want.view(np.uint32))
            say(f"scene {s} {key}: bitwise-equal fraction {same:.7f}")
            assert same >= 0.9999, (s, key)
    b.close()


# ---- densifying configuration ----------------------------------------------------------

@pytest.fixture(scope="module")
def densify_event(ctx):
    """The densifying run's state at its first density-control event (iteration 600)."""
    say("ellipsoid scene: 2M-Gaussian ground truth, 300k points, 100 cameras 1297x840")
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
    kw = dict(densify=1, densify_from_iter=500, densify_until_iter=5000,
              densification_interval=100, opacity_reset_interval=3000,
              densify_grad_threshold=0.0002, sh_degree=3, sh_degree_interval=1000,
              position_lr_max_steps=7000, log_interval=0, seed=0, ordered_backward=1)
    tr = abi.Trainer(ctx, init)
    for c, v in zip(cams, views):
        tr.add_view(c, v)
    tr.attach_scorer(abi.Scorer(ctx, seed=0), rec.learn_config(mode=1))
    say("splat_train iterations 1-599")
    tr.train(extent, abi.TrainConfig.default(iterations=599, **kw))
    it = 600
    view = abi.lib().splat_train_view_at(0, len(cams), it)
    lr = rec.learning_rates(rec.position_lr(it, extent, max_steps=7000))
    tr.step(view, rec.step_params(0, lr, accumulate_stats=True, ordered_backward=True))
    yield tr, extent, len(init)
    tr.close()


def test_scorer_features_and_logits_at_scale(ctx, densify_event):
    tr, extent, _ = densify_event
    d = tr.download(stats=True)
    lin = tr.lineage()
    it = 600
    say(f"scorer features of {len(d['params'])} Gaussians")
    got = tr.features(extent, 0.0002, 0.01, it)
    want = O.scorer_features(d["params"], d["grad_accum"], d["denom"], d["max_radii"],
                             lin["max_grad"], lin["strong"], lin["birth"], 0.0002,
                             np.float32(math.log(np.float32(0.01) * np.float32(extent))), it)
    err = np.abs(got - want).max()
    say(f"features: max abs difference {err:.3e}")
    assert err <= 1e-6
    scorer = abi.Scorer(ctx, seed=11)
    logits = scorer.forward(got)
    ref = O.mlp_forward(scorer.params(), got, bf16=True)
    scale = np.abs(ref).max()
    e = np.abs(logits - ref)
    say(f"logits: max abs difference {e.max():.3e} (scale {scale:.3e}), mean {e.mean():.3e}")
    assert e.max() <= 1e-2 * scale and e.mean() <= 1e-3 * scale
    scorer.close()


def test_first_densification_at_scale_matches_oracle(densify_event):
    tr, extent, n0 = densify_event
    before = tr.download(moments=True, stats=True)
    n = len(before["params"])
    assert n == n0 and int(before["denom"].max()) > 0
    cfg = rec.densify_config(grad_threshold=0.0002)  # prune_big_world: 600 <= 3000
    seed = 600
    say(f"densify {n} Gaussians on the device and in the oracle")
    rep = tr.densify(cfg, extent, seed)
    after = tr.download(moments=True, stats=True)
    nrm = trainer_normals(seed, 3 * int(cfg["split_children"]) * n)
    po, mo, vo, repo = O.densify_and_prune(before["params"], before["m"], before["v"],
                                           before["grad_accum"], before["denom"],
                                           before["max_radii"], cfg, extent, nrm)
    say(f"report {rep.tolist()}, oracle {repo.tolist()}; {len(po)} Gaussians after")
    assert tuple(rep.tolist()) == tuple(repo.tolist())
    assert int(rep["cloned"]) + int(rep["split"]) > 0
    assert after["params"].shape == po.shape
    # every field but the split children's means is copied or shifted exactly; the means
    # add R (s * normal) (binary64 Box-Muller draws on both sides)
    assert np.array_equal(after["params"][:, 3:].view(np.uint32), po[:, 3:].view(np.uint32))
    assert np.allclose(after["params"][:, :3], po[:, :3], rtol=1e-6, atol=1e-7)
    assert np.array_equal(after["m"].view(np.uint32), mo.view(np.uint32))
    assert np.array_equal(after["v"].view(np.uint32), vo.view(np.uint32))
    assert not after["grad_accum"].any() and not after["denom"].any()
