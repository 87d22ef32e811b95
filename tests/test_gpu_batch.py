"""Batched training of many scenes and views (needs a gfx950 GPU).

- one step over three scenes of different sizes and six views of different resolutions,
  in an interleaved order, equals the oracle composed view by view: each view's
  reference gradient (project -> bin -> composite -> loss -> backward), summed per scene
  in list order, then Adam with the scene's position learning rate (bitwise, ordered
  backward);
- a batch of one view of one scene steps exactly like the single-view trainer;
- the launch shapes chosen by the autotuner give bitwise the same model as the default
  shapes (the shapes change the schedule, not the arithmetic).
"""
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


SIZES = [3000, 5000, 2000]
VIEWS = [(0, 96, 64), (1, 64, 64), (1, 200, 120), (0, 128, 100), (2, 100, 100), (1, 80, 144)]
ORDER = [3, 0, 5, 1, 4, 2]


def setup(ctx):
    gs = [scenes.random_gaussians(np.random.default_rng(40 + s), n) for s, n in enumerate(SIZES)]
    gts = [scenes.random_gaussians(np.random.default_rng(80 + s), n) for s, n in enumerate(SIZES)]
    dirs = scenes.fibonacci_sphere(len(VIEWS))
    cams, targets = [], []
    for k, (s, w, h) in enumerate(VIEWS):
        d = dirs[k]
        eye = (4.0 * d[0], 4.0 * d[1], 4.0 * d[2])
        up = (0.0, -1.0, 0.0) if abs(d[1]) < 0.99 else (0.0, 0.0, 1.0)
        cam = scenes.look_at(w, h, 1.1 * w, 1.1 * w, eye, (0.0, 0.0, 0.0), up)
        cams.append(cam)
        targets.append(abi.to_rgb8(ctx.render(gts[s], cam, rec.proj_opts(3))))
    return gs, cams, targets


def make_batch(ctx, gs, cams, targets):
    b = abi.Batch(ctx)
    for g in gs:
        b.add_scene(g)
    for (s, _, _), c, t in zip(VIEWS, cams, targets):
        b.add_view(s, c, t)
    return b


def oracle_view_grad(g, cam, target, degree):
    opts = rec.proj_opts(degree)
    W, H = int(cam["width"]), int(cam["height"])
    projs = O.project_all(g, cam, opts)
    binning = O.bin(projs, W, H)
    settings = rec.raster_settings(W, H, (0.0, 0.0, 0.0))
    image, final_t, n_contrib = O.rasterize_forward(projs, binning, settings)
    tgt = target.astype(np.float32) / np.float32(255.0)
    loss, dimage = O.photometric_loss(image, tgt, W, H, 0.2)
    pgrads, _ = O.rasterize_backward(projs, binning, settings, final_t, n_contrib, dimage)
    return loss, O.project_backward_all(g, cam, opts, pgrads)


def test_batched_step_matches_oracle(ctx):
    gs, cams, targets = setup(ctx)
    b = make_batch(ctx, gs, cams, targets)
    lr = rec.learning_rates(0.0)
    lr_mean = np.array([1.6e-4, 2.4e-4, 0.8e-4], np.float32)
    p = rec.batch_step_params(3, lr, ordered_backward=True)
    say("batched step over 3 scenes / 6 views")
    losses = b.step(ORDER, p, lr_mean)
    for s, g0 in enumerate(gs):
        total = None
        for k in ORDER:
            if VIEWS[k][0] != s:
                continue
            loss, gv = oracle_view_grad(g0, cams[k], targets[k], 3)
            assert abs(losses[ORDER.index(k)] - loss) <= 1e-5 * abs(loss), (k, losses, loss)
            total = gv.copy() if total is None else total + gv
        params, m, v = g0.copy(), np.zeros_like(g0), np.zeros_like(g0)
        lrs = lr.copy()
        lrs["mean"] = lr_mean[s]
        O.adam_step(params, total, m, v, rec.adam_state(), lrs)  # increments to step 1
        got = b.download(s, moments=True)
        for key, want in (("params", params), ("m", m), ("v", v)):
            same = np.mean(got[key].view(np.uint32) == want.view(np.uint32))
            say(f"scene {s} {key}: bitwise-equal fraction {same:.6f}")
            assert same == 1.0, (s, key)


def test_single_view_batch_equals_trainer(ctx):
    g = scenes.small_scene(0, 4000)
    cams = scenes.small_scene_cameras()
    target = abi.to_rgb8(ctx.render(scenes.small_scene(1, 4000), cams[2], rec.proj_opts(3)))
    b = abi.Batch(ctx)
    b.add_scene(g)
    b.add_view(0, cams[2], target)
    tr = abi.Trainer(ctx, g)
    tr.add_view(cams[2], target)
    extent = scenes.scene_extent(cams)
    for it in range(1, 4):
        lr = rec.learning_rates(rec.position_lr(it, extent))
        tr.step(0, rec.step_params(3, lr))
        b.step([0], rec.batch_step_params(3, lr), [lr["mean"]])
    a, c = tr.download(moments=True), b.download(0, moments=True)
    for k in ("params", "m", "v"):
        assert np.array_equal(a[k].view(np.uint32), c[k].view(np.uint32)), k


def test_tuned_shapes_give_the_same_model(ctx):
    gs, cams, targets = setup(ctx)
    a = make_batch(ctx, gs, cams, targets)
    b = make_batch(ctx, gs, cams, targets)
    lr = rec.learning_rates(0.0)
    p = rec.batch_step_params(3, lr)
    r = b.tune(ORDER, p)
    say(f"tuned: fwd {int(r['fwd_batch'])}/{int(r['fwd_persistent'])}/{int(r['fwd_per_cu'])} "
        f"{float(r['fwd_ms']):.4f} ms (default {float(r['default_fwd_ms']):.4f}), bwd "
        f"{int(r['bwd_batch'])}/{int(r['bwd_persistent'])}/{int(r['bwd_per_cu'])} "
        f"{float(r['bwd_ms']):.4f} ms (default {float(r['default_bwd_ms']):.4f}), "
        f"{int(r['candidates'])} candidates")
    assert r["candidates"] == 32 and r["fwd_ms"] <= r["default_fwd_ms"]
    # force persistent shapes on a third batch as well
    c = make_batch(ctx, gs, cams, targets)
    forced = r.copy()
    forced["fwd_batch"], forced["fwd_persistent"], forced["fwd_per_cu"] = 128, 1, 2
    forced["bwd_batch"], forced["bwd_persistent"], forced["bwd_per_cu"] = 16, 1, 8
    forced["bwd_row_lists"] = 1 - int(r["bwd_row_lists"])  # the other entry walk
    c.set_shapes(forced)
    lr_mean = [1e-4, 2e-4, 3e-4]
    for _ in range(3):
        la = a.step(ORDER, p, lr_mean)
        lb = b.step(ORDER, p, lr_mean)
        lc = c.step(ORDER, p, lr_mean)
        assert np.array_equal(la, lb) and np.array_equal(la, lc)
    for s in range(len(SIZES)):
        x, y, z = a.download(s, True), b.download(s, True), c.download(s, True)
        for k in ("params", "m", "v"):
            assert np.array_equal(x[k].view(np.uint32), y[k].view(np.uint32)), (s, k)
            assert np.array_equal(x[k].view(np.uint32), z[k].view(np.uint32)), (s, k)
    info = a.info()
    assert info["adam_step"] == 3 and info["captures"] >= 1 and info["replays"] >= 3
