"""The benchmarked train step, at its own size, against the oracle (needs a gfx950 GPU).

Workload: 1M Gaussians (means U[-5,5]^3, log-scales N(-5,0.7)), one 1920x1080 camera
(fx = fy = 1600) at distance 8, an RGB8 target rendered from a second scene of the same
generator, SH degree 3.  The trainer runs exactly as the benchmark times it: graph-replayed
steps, compositing launch shapes chosen by the autotuner, the tree-reduced backward and the
fused slot-reduce / statistics / projection-backward / Adam kernel.  Each step's parameter
gradient is read back through the first Adam moment: from zero moments, Adam's
m = beta1 m + (1 - beta1) g is the single rounding of (1 - beta1) g, the same map on both
sides (src/adam.rs:44-61).

Bars:
  * tree-reduced steps: at every step the gradient is no less accurate than the
    reference's binary32 order (the oracle) against the binary64 gradient of the same
    binary32 forward, per parameter group, at the median, p99 and p99.9; three continuous
    steps keep params within Adam's step bound of the oracle's and the visibility
    statistics bitwise;
  * the ordered backward: three continuous steps give params, both moments and the
    statistics bitwise equal to the oracle's (up to the rare ulp of ln(255 o), where the
    device's correctly rounded log and glibc's logf differ);
  * the loss at 1920x1080: the photometric-loss and SSIM gradients bitwise; the scalars
    within 2e-6 (loss) and 1e-5 (SSIM) of the binary64 sums of the same binary32 terms
    and no less accurate than the reference's sequential sum.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from helpers import no_less_accurate
from splatrs_amd import abi, scenes
from splatrs_amd import records as rec

pytestmark = pytest.mark.gpu

GROUPS = [("mean", 0, 3), ("log_scale", 3, 6), ("quat", 6, 10), ("opacity", 10, 11),
          ("sh_dc", 11, 14), ("sh_rest", 14, 59)]
C1 = np.float32(np.float32(1.0) - np.float32(0.9))  # 1 - beta1 in binary32


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def problem(ctx):
    say("building the 1M / 1920x1080 workload")
    g = scenes.large_scene(0)
    cam = scenes.hd_camera()
    rgb = abi.to_rgb8(ctx.render(scenes.large_scene(1), cam, rec.proj_opts(3)))
    return g, cam, rgb, scenes.scene_extent([cam])


def bench_trainer(ctx, g, cam, rgb, extent, ordered):
    """The trainer as the benchmark builds it: graph mode (the default), tuned shapes."""
    tr = abi.Trainer(ctx, g)
    tr.add_view(cam, rgb)
    lr = rec.learning_rates(rec.position_lr(1, extent))
    shapes = tr.tune(0, rec.step_params(3, lr, ordered_backward=ordered))
    say(f"tuned shapes: fwd batch {int(shapes['fwd_batch'])} persistent "
        f"{int(shapes['fwd_persistent'])}, bwd batch {int(shapes['bwd_batch'])} persistent "
        f"{int(shapes['bwd_persistent'])}")
    return tr


def step_params(it, extent, ordered):
    lr = rec.learning_rates(rec.position_lr(it, extent))
    return rec.step_params(3, lr, accumulate_stats=True, ordered_backward=ordered), lr


def test_timed_step_gradients_against_binary64_truth(ctx, problem):
    g, cam, rgb, extent = problem
    tr = bench_trainer(ctx, g, cam, rgb, extent, ordered=False)
    ts = O.TrainState(g)
    for it in range(1, 4):
        say(f"tree-reduced step {it}: restore the oracle's parameters, zero moments")
        tr.restore({"params": ts.params}, adam_step=it - 1)
        p, lr = step_params(it, extent, False)
        tr.step(0, p)
        m = tr.download(moments=True)["m"]
        grads = O.step_gradients(ts.params, cam, rgb, 3)
        m_ref = C1 * grads["ref"]  # binary32 products, as the kernel rounds them
        m_true = np.float64(C1) * grads["true"]
        for name, a, b in GROUPS:
            no_less_accurate(f"step {it} {name}", m[:, a:b], m_ref[:, a:b], m_true[:, a:b], say)
        ts.train_step(cam, rgb, 3, lr, accumulate_stats=True)
    captures, replays, aborts = tr.graph_stats()
    say(f"graph captures {captures}, replays {replays}, aborted steps {aborts}")
    assert replays >= 3
    tr.close()


def test_timed_steps_follow_oracle(ctx, problem):
    g, cam, rgb, extent = problem
    tr = bench_trainer(ctx, g, cam, rgb, extent, ordered=False)
    ts = O.TrainState(g)
    for it in range(1, 4):
        say(f"tree-reduced continuous step {it}")
        p, lr = step_params(it, extent, False)
        r = tr.step(0, p)
        _, e = ts.train_step(cam, rgb, 3, lr, accumulate_stats=True)
        assert int(r["num_entries"]) == e
    out = tr.download(moments=True, stats=True)
    # Adam (eps 1e-15) turns rounding-level differences of near-zero gradients into steps
    # of up to the learning rate: each parameter stays within 3 steps x 2 rates.
    lr = rec.learning_rates(rec.position_lr(1, extent))
    rates = np.array([lr["mean"]] * 3 + [lr["log_scale"]] * 3 + [lr["quat"]] * 4 +
                     [lr["opacity"]] + [lr["sh_dc"]] * 3 + [lr["sh_rest"]] * 45)
    d = np.abs(out["params"].astype(np.float64) - ts.params)
    assert (d <= 6 * rates + 1e-6).all()
    same = np.mean(out["params"] == ts.params)
    say(f"params bitwise-equal fraction after 3 tree-reduced steps: {same:.6f}")
    assert np.array_equal(out["denom"], ts.denom)
    assert np.array_equal(out["max_radii"], ts.max_radii)
    tr.close()


def test_timed_steps_ordered_bitwise(ctx, problem):
    g, cam, rgb, extent = problem
    tr = bench_trainer(ctx, g, cam, rgb, extent, ordered=True)
    ts = O.TrainState(g)
    for it in range(1, 4):
        say(f"ordered step {it}")
        p, lr = step_params(it, extent, True)
        tr.step(0, p)
        ts.train_step(cam, rgb, 3, lr, accumulate_stats=True)
    out = tr.download(moments=True, stats=True)
    for name, a, b in (("params", out["params"], ts.params), ("m", out["m"], ts.m),
                       ("v", out["v"], ts.v), ("grad_accum", out["grad_accum"], ts.grad_accum)):
        same = np.mean(a.view(np.uint32) == b.view(np.uint32))
        rel = np.abs(a.astype(np.float64) - b) / (np.abs(b).astype(np.float64) + 1e-30)
        say(f"{name}: bitwise-equal fraction {same:.7f}, max relative difference {rel.max():.2e}")
        assert same >= 0.9999, name
    assert np.array_equal(out["denom"], ts.denom)
    assert np.array_equal(out["max_radii"], ts.max_radii)
    tr.close()


def test_loss_gradient_1080p_bitwise(ctx, problem):
    g, cam, rgb, _ = problem
    w, h = int(cam["width"]), int(cam["height"])
    say("photometric loss and SSIM at 1920x1080")
    pred = ctx.render(g, cam, rec.proj_opts(3))
    target = rgb.astype(np.float32) / np.float32(255.0)
    loss, grad = ctx.photometric_loss(pred, target, w, h, 0.2)
    lo, go = O.photometric_loss(pred, target, w, h, 0.2)
    assert np.array_equal(grad.view(np.uint32), go.view(np.uint32))
    # The scalars are sums of 6.2M terms whose order the reference leaves to rayon's
    # parallel sum (src/loss.rs:96-121, 146-151); the oracle adds them sequentially in
    # binary32.  Against the binary64 sums of the same binary32 terms, the device's sums
    # (binary32 over each workgroup's 8192 terms, by shuffles: error about 13 ulp of the
    # block's magnitude; binary64 over the blocks) must be within 2e-6 relative for the
    # loss and 1e-5 for the mean SSIM (its terms change sign), and no less accurate than
    # the sequential order.
    ssim64, l164 = O.loss_sums64(pred, target, w, h)
    n = w * h * 3
    s64 = ssim64 / n
    l64 = 0.8 * (l164 / n) + 0.2 * (1.0 - s64)
    say(f"loss {loss!r}, oracle {lo!r}, binary64 sum {l64!r}")
    assert abs(loss - l64) <= 2e-6 * abs(l64)
    assert abs(loss - l64) <= abs(lo - l64) + 1e-7 * abs(l64)
    s, sg = ctx.ssim_with_grad(pred, target, w, h)
    so, sgo = O.ssim_with_grad(pred, target, w, h)
    assert np.array_equal(sg.view(np.uint32), sgo.view(np.uint32))
    say(f"ssim {s!r}, oracle {so!r}, binary64 sum {s64!r}")
    assert abs(s - s64) <= 1e-5 * abs(s64)
    assert abs(s - s64) <= abs(so - s64) + 1e-7 * abs(s64)


def test_pair_counts_1080p_equal_the_oracle(ctx, problem):
    """The counting compositing kernels at 1M / 1080p: forward evaluated, passing and
    composited pairs and backward evaluated and passing pairs equal the oracle's walk; the
    backward passes exactly the composited pairs."""
    g, cam, rgb, extent = problem
    w, h = int(cam["width"]), int(cam["height"])
    tr = abi.Trainer(ctx, g)
    tr.add_view(cam, rgb)
    p, _ = step_params(1, extent, False)
    got = tr.pair_counts(0, p)
    tr.close()
    say(f"device pair counts: {got}")
    opts = rec.proj_opts(3)
    po = O.project_all(g, cam, opts)
    bo = O.bin(po, w, h)
    st = rec.raster_settings(w, h)
    _, _, no = O.rasterize_forward(po, bo, st)
    want = O.raster_pair_counts(po, bo, st, no)
    say(f"oracle pair counts: {want}")
    for k in ("fwd_passing", "fwd_composited", "bwd_passing"):
        assert got[k] == want[k], k
    # sigma_max = ln(255 o): where the device's correctly rounded log and glibc's logf
    # differ by an ulp, a pair at sigma = sigma_max is evaluated on one side only (and fails
    # alpha_min on both)
    for k in ("fwd_evaluated", "bwd_evaluated"):
        assert abs(got[k] - want[k]) <= 1e-6 * want[k], k
    assert got["bwd_passing"] == got["fwd_composited"]
    for d in ("fwd", "bwd"):
        assert got[f"{d}_active_visits"] <= got[f"{d}_visits"]
        assert got[f"{d}_active_visits"] <= got[f"{d}_evaluated"] <= 64 * got[f"{d}_active_visits"]
