"""Pins of the oracle's binning, compositing and loss (src/rasterize.rs, src/loss.rs).

- bin: sorted list equals a brute-force sort by (tile, depth bits, entry), offsets and
  ranges consistent (src/rasterize.rs:60-128);
- forward: tiled compositor vs a per-pixel compositor (port of tests/pipeline.rs:134-224);
- backward: f64 central differences over 9 screen-space Gaussians on 40x28 with the AbsGS
  bound (port of tests/gradcheck_f64.rs:63-104);
- loss: SSIM(x, x) = 1 with vanishing gradient (src/loss.rs:196-207), f64 central
  differences of SSIM and the photometric loss (tests/gradcheck_f64.rs:106-140), and torch
  conv2d + autograd as an independent library routine for blur, SSIM and its gradient;
- chain: parameters -> loss in f64 by central differences (tests/gradcheck_f64.rs:146-217).
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import (F64, central, check, ground_truth, naive_render, random_gaussian,
                     ref_test_camera, ring_cameras, rnd)
from splatrs_amd import records as rec


def _tiles_of(p):
    return (int(p["tile_max"][0]) - int(p["tile_min"][0])) * \
        (int(p["tile_max"][1]) - int(p["tile_min"][1])) if p["radius"] > 0 else 0


def brute_force_bin(projs, width, height):
    tx = (width + 15) // 16
    entries = []
    for gi, p in enumerate(projs):
        if p["radius"] == 0:
            continue
        db = int(np.float32(p["depth"]).view(np.uint32))
        for ty in range(p["tile_min"][1], p["tile_max"][1]):
            for x in range(p["tile_min"][0], p["tile_max"][0]):
                entries.append((((ty * tx + x) << 32) | db, gi))
    keyed = sorted((k, e) for e, (k, _) in enumerate(entries))
    return entries, keyed


def test_bin_matches_brute_force_sort():
    rng = np.random.default_rng(500)
    gt = ground_truth(rng, 120)
    gt[::10, 3:6] = [-0.7, -0.9, -0.8]
    cam = ring_cameras(1, 70, 50, 2.4)[0]
    projs = O.project_all(gt, cam, rec.proj_opts(1, True))
    b = O.bin(projs, 70, 50)
    entries, keyed = brute_force_bin(projs, 70, 50)
    assert b.num_entries == len(entries) > 100
    assert [e[1] for e in entries] == list(b.entry_gaussian)
    assert [k for k, _ in keyed] == list(b.sorted_keys)
    assert [e for _, e in keyed] == list(b.sorted_entries)
    off = np.concatenate([[0], np.cumsum([_tiles_of(p) for p in projs])])
    assert np.array_equal(off, b.entry_offset)
    for t in range(b.tiles_x * b.tiles_y):
        s, e = b.tile_range[t]
        assert np.all((b.sorted_keys[s:e] >> np.uint64(32)) == t)


def test_bin_ties_break_by_entry():
    # Identical Gaussians (as after a clone) share tile and depth: ties go by entry index.
    rng = np.random.default_rng(1)
    g = ground_truth(rng, 5)
    g = np.concatenate([g, g[:3]])
    cam = ring_cameras(1, 48, 40, 2.4)[0]
    projs = O.project_all(g, cam, rec.proj_opts(1))
    b = O.bin(projs, 48, 40)
    pairs = list(zip(b.sorted_keys.tolist(), b.sorted_entries.tolist()))
    assert pairs == sorted(pairs)
    assert len(set(b.sorted_keys.tolist())) < len(pairs)


def test_tiled_rasterizer_matches_naive_compositing():
    rng = np.random.default_rng(500)
    model = ground_truth(rng, 120)
    for i in range(len(model)):
        if i % 10 == 0:
            model[i, 3:6] = [-0.7, -0.9, -0.8]
        if i % 7 == 0:
            model[i, 10] = -5.0
    cam = ring_cameras(1, 70, 50, 2.4)[0]
    projs = O.project_all(model, cam, rec.proj_opts(1, True))
    bg = (0.3, 0.6, 0.1)
    b = O.bin(projs, 70, 50)
    tiled, _, _ = O.rasterize_forward(projs, b, rec.raster_settings(70, 50, bg))
    naive = naive_render(projs, 70, 50, bg)
    assert np.abs(tiled - naive).max() < 1e-5
    assert int((projs["radius"] > 0).sum()) > 100


def random_projs(rng, width, height, n):
    """tests/gradcheck_f64.rs:28-55: screen-space Gaussians over the whole image."""
    tx, ty = (width + 15) // 16, (height + 15) // 16
    out = np.zeros((n, 15))
    for i in range(n):
        l00, l10, l11 = rnd(rng, 1.5, 4.0), rnd(rng, -2.0, 2.0), rnd(rng, 1.5, 4.0)
        a, b, c = l00 * l00, l00 * l10, l10 * l10 + l11 * l11
        det = a * c - b * b
        out[i, 0:2] = [rnd(rng, 4.0, width - 4.0), rnd(rng, 4.0, height - 4.0)]
        out[i, 2:5] = [c / det, -b / det, a / det]
        out[i, 5] = 1.0 + i * 0.37
        out[i, 6] = 1000
        out[i, 7] = rnd(rng, 0.3, 0.9)
        out[i, 8:11] = [rnd(rng, 0, 1) for _ in range(3)]
        out[i, 11:15] = [0, 0, tx, ty]
    return out


def test_rasterizer_gradient_f64():
    rng = np.random.default_rng(21)
    w, h = 40, 28
    bg = [0.2, 0.5, 0.8]
    projs = random_projs(rng, w, h, 9)
    wt = rng.uniform(-1, 1, (h, w, 3))
    grads, absg = O.raster_backward_projs_f64(projs, w, h, bg, wt)
    fields = [0, 1, 2, 3, 4, 7, 8, 9, 10]  # mean2d, conic, opacity, color in Proj layout
    for i in range(len(projs)):
        for f, col in enumerate(fields):
            hstep = 1e-6 if 2 <= f <= 4 else F64["H"]

            def loss(e):
                p = projs.copy()
                p[i, col] += e
                img, _, _ = O.render_projs_f64(p, w, h, bg)
                return float((img * wt).sum())
            num = central(0.0, hstep, loss)
            check(f"gaussian {i} field {f}", float(grads[i, f]), num, 1.0, F64)
        assert absg[i, 0] + 1e-9 >= abs(grads[i, 0])
        assert absg[i, 1] + 1e-9 >= abs(grads[i, 1])


# ---- loss -------------------------------------------------------------------------------

def test_identical_images_have_unit_ssim():
    w, h = 20, 14
    img = (np.arange(w * h * 3) * 7919 % 255).astype(np.float32) / np.float32(255.0)
    s, g = O.ssim_with_grad(img, img, w, h)
    assert abs(s - 1.0) < 1e-4
    assert np.abs(g).max() < 1e-3


def test_ssim_gradient_f64():
    rng = np.random.default_rng(31)
    w, h = 18, 13
    x = rng.uniform(0, 1, w * h * 3)
    y = rng.uniform(0, 1, w * h * 3)
    _, grad = O.ssim_with_grad_f64(x, y, w, h)
    for _ in range(40):
        i = int(rng.integers(0, x.size))

        def f(e):
            xp = x.copy()
            xp[i] += e
            return O.ssim_with_grad_f64(xp, y, w, h)[0]
        check(f"ssim d/dx[{i}]", float(grad[i]), central(0.0, F64["H"], f), 1.0, F64)


def test_photometric_loss_gradient_f64():
    rng = np.random.default_rng(32)
    w, h = 16, 12
    x = rng.uniform(0, 1, w * h * 3)
    y = rng.uniform(0, 1, w * h * 3)
    _, grad = O.photometric_loss_f64(x, y, w, h, 0.2)
    for _ in range(40):
        i = int(rng.integers(0, x.size))

        def f(e):
            xp = x.copy()
            xp[i] += e
            return O.photometric_loss_f64(xp, y, w, h, 0.2)[0]
        check(f"loss d/dx[{i}]", float(grad[i]), central(0.0, F64["H"], f), 1.0, F64)


def _torch_ssim(x, y, w, h):
    import torch
    k = torch.exp(-(torch.arange(11, dtype=torch.float64) - 5) ** 2 / (2 * 1.5 ** 2))
    k = k / k.sum()
    win = (k[:, None] * k[None, :])[None, None].repeat(3, 1, 1, 1)

    def conv(t):
        return torch.nn.functional.conv2d(t, win, padding=5, groups=3)
    X = x.reshape(h, w, 3).permute(2, 0, 1)[None]
    Y = y.reshape(h, w, 3).permute(2, 0, 1)[None]
    mx, my = conv(X), conv(Y)
    sx, sy, sxy = conv(X * X) - mx * mx, conv(Y * Y) - my * my, conv(X * Y) - mx * my
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    s = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sx + sy + c2))
    return s.mean(), conv


def test_loss_matches_torch_conv2d_autograd():
    import torch
    rng = np.random.default_rng(33)
    w, h = 23, 17
    x = rng.uniform(0, 1, w * h * 3)
    y = rng.uniform(0, 1, w * h * 3)
    xt = torch.tensor(x, requires_grad=True)
    s_t, conv = _torch_ssim(xt, torch.tensor(y), w, h)
    s_t.backward()
    s_o, g_o = O.ssim_with_grad_f64(x, y, w, h)
    assert abs(s_o - s_t.item()) < 1e-12
    assert np.allclose(g_o, xt.grad.numpy(), rtol=1e-9, atol=1e-15)
    # the blur itself
    xb = O.blur(x.astype(np.float32), w, h, 3)
    tb = conv(torch.tensor(x.astype(np.float32), dtype=torch.float64).reshape(h, w, 3)
              .permute(2, 0, 1)[None])[0].permute(1, 2, 0).reshape(-1).numpy()
    assert np.abs(xb - tb).max() < 1e-6
    # f32 loss gradient vs the f64 one
    l32, g32 = O.photometric_loss(x.astype(np.float32), y.astype(np.float32), w, h, 0.2)
    l64, g64 = O.photometric_loss_f64(x.astype(np.float32).astype(np.float64),
                                      y.astype(np.float32).astype(np.float64), w, h, 0.2)
    assert abs(l32 - l64) < 1e-5
    assert np.abs(g32 - g64).max() < 1e-8


def test_gaussian_kernel_and_psnr():
    k = O.gaussian_kernel()
    ref = np.exp(-(np.arange(11) - 5.0) ** 2 / 4.5)
    assert np.allclose(k, ref / ref.sum(), rtol=1e-6)
    a = np.array([0.0, 0.5, 1.2, -0.1], np.float32)
    b = np.array([0.1, 0.5, 1.0, 0.0], np.float32)
    mse = np.mean((np.clip(a, 0, 1).astype(np.float64) - np.clip(b, 0, 1)) ** 2)
    assert abs(O.psnr(a, b) - (-10 * np.log10(mse))) < 1e-5
    assert O.psnr(a, a) == np.inf


@pytest.mark.slow
def test_end_to_end_gradient_f64():
    rng = np.random.default_rng(41)
    cam = ref_test_camera(48, 32)
    w, h = 48, 32
    gs = []
    for _ in range(6):
        g = random_gaussian(rng)
        g[3:6] = [rnd(rng, -1.6, -0.9) for _ in range(3)]
        gs.append(g)
    model = np.array(gs)
    target = rng.uniform(0, 1, (h, w, 3))
    c64, o64 = O.cam64(cam), O.opts64(3, True)
    bg = [0.1, 0.1, 0.1]
    p = O.project_all_f64(model, c64, o64)
    assert np.all(p[:, 6] > 0)
    _, grads = O.chain_loss_f64(model, c64, o64, bg, target, 1.0, want_grad=True)
    hstep = 1e-5
    for gi in range(len(model)):
        for i in range(59):
            def f(e):
                m = model.copy()
                m[gi, i] += e
                return O.chain_loss_f64(m, c64, o64, bg, target, 1.0)
            check(f"gaussian {gi} param {i}", float(grads[gi, i]), central(0.0, hstep, f),
                  1.0, F64)


def _projs32(p64):
    """Binary32 Proj records of f64 screen-space Gaussians (the layout of random_projs)."""
    out = np.zeros(len(p64), rec.PROJ)
    out["mean2d"] = p64[:, 0:2]
    out["conic"] = p64[:, 2:5]
    out["depth"] = p64[:, 5]
    out["radius"] = p64[:, 6].astype(np.uint32)
    out["opacity"] = p64[:, 7]
    out["color"] = p64[:, 8:11]
    out["tile_min"] = p64[:, 11:13].astype(np.uint32)
    out["tile_max"] = p64[:, 13:15].astype(np.uint32)
    return out


def test_truth_backward_equals_f64_backward_of_binary32_inputs():
    # The binary64 accuracy reference (binary32 decisions, binary64 values) against the
    # finite-difference-pinned f64 backward (test_rasterizer_gradient_f64) run on the same
    # binary32 values promoted to f64: with no decision near a threshold they are the same
    # function, so they agree to binary64 rounding; and the binary32 reference order lies
    # within binary32 rounding of both.
    rng = np.random.default_rng(33)
    w, h = 40, 28
    bg = np.float32([0.2, 0.5, 0.8]).astype(np.float64)
    p32 = _projs32(random_projs(rng, w, h, 9))
    p64 = np.zeros((9, 15))
    p64[:, 0:2], p64[:, 2:5], p64[:, 5] = p32["mean2d"], p32["conic"], p32["depth"]
    p64[:, 6], p64[:, 7], p64[:, 8:11] = p32["radius"], p32["opacity"], p32["color"]
    p64[:, 11:13], p64[:, 13:15] = p32["tile_min"], p32["tile_max"]
    wt = rng.uniform(-1, 1, (h, w, 3)).astype(np.float32)
    g64, a64 = O.raster_backward_projs_f64(p64, w, h, bg, wt.astype(np.float64))
    st = rec.raster_settings(w, h, bg)
    b = O.bin(p32, w, h)
    _, ft, nc = O.rasterize_forward(p32, b, st)
    gt, at = O.rasterize_backward_truth(p32, b, st, nc, wt)
    scale = np.abs(g64).max(axis=0)
    assert np.all(np.abs(gt - g64) <= 1e-10 * scale)
    assert np.all(np.abs(at - a64) <= 1e-10 * np.abs(a64).max(axis=0))
    g32, _ = O.rasterize_backward(p32, b, st, ft, nc, wt)
    g32 = g32.view(np.float32).reshape(-1, 9)
    assert np.all(np.abs(g32 - gt) <= 1e-5 * scale)
    assert np.abs(g32 - gt).max() > 0  # binary32 rounding is visible against the truth


def test_step_gradient_truth_is_close_to_reference_arithmetic():
    # Parameter gradients of a small training step: the binary32 reference arithmetic is
    # within binary32 rounding of the binary64 truth, field by field.
    from splatrs_amd import scenes
    g = scenes.small_scene(0, 600)
    cam = scenes.small_scene_cameras()[0]
    gt_img = np.clip(np.random.default_rng(3).uniform(0, 1, (256, 256, 3)), 0, 1)
    rgb = np.round(gt_img * 255).astype(np.uint8)
    out = O.step_gradients(g, cam, rgb, 3, screen=True)
    ref, true = out["ref"].astype(np.float64), out["true"]
    scale = np.abs(true).max(axis=0) + 1e-30
    assert np.abs(true).max() > 0
    assert np.all(np.abs(ref - true) <= 1e-4 * scale)
    # the screen-space part equals rasterize_backward_truth's output
    sr = out["screen_ref"].view(np.float32).reshape(-1, 9).astype(np.float64)
    st = out["screen_true"]
    assert np.all(np.abs(sr - st) <= 1e-5 * (np.abs(st).max(axis=0) + 1e-30))


def test_accuracy_bar_rejects_a_dropped_row_partial():
    # The bar the tree-reduced backward is held to (no less accurate than the reference
    # order against the binary64 truth, at p50 / p99 / p99.9) must fail for a plausible
    # reduction mistake: one of a tile's 16 row partials dropped for 1% of the Gaussians.
    from helpers import no_less_accurate
    from splatrs_amd import scenes
    g = scenes.small_scene(0)  # 10k Gaussians: every quantile has its tail samples
    cam = scenes.small_scene_cameras()[0]
    rgb = np.round(np.random.default_rng(3).uniform(0, 1, (256, 256, 3)) * 255).astype(np.uint8)
    out = O.step_gradients(g, cam, rgb, 3, screen=True)
    ref = out["screen_ref"].view(np.float32).reshape(-1, 9)
    truth = out["screen_true"]
    no_less_accurate("reference against itself", ref, ref, truth, say=lambda m: None)
    rng = np.random.default_rng(5)
    bad = ref.copy()
    rows = rng.choice(len(bad), max(1, len(bad) // 100), replace=False)
    bad[rows] *= np.float32(15.0 / 16.0)
    with pytest.raises(AssertionError):
        no_less_accurate("dropped partial", bad, ref, truth, say=lambda m: None)


def test_pair_counts_match_a_naive_compositor():
    # The oracle's pair counts against a per-pixel walk over every Gaussian whose tile
    # rectangle covers the pixel's tile (tests/pipeline.rs:134-186 with counters): passing
    # pairs (alpha >= 1/255 before the pixel is done) and composited pairs.  The backward
    # walk (k < n_contrib) passes exactly the composited pairs.
    rng = np.random.default_rng(21)
    w, h = 40, 28
    gt = ground_truth(rng, 60)
    from helpers import ref_test_camera
    cam = ref_test_camera(w, h)
    opts = rec.proj_opts(1)
    projs = O.project_all(gt, cam, opts)
    b = O.bin(projs, w, h)
    st = rec.raster_settings(w, h)
    _, _, nc = O.rasterize_forward(projs, b, st)
    got = O.raster_pair_counts(projs, b, st, nc)
    f = np.float32
    vis = [i for i in range(len(projs)) if projs[i]["radius"] > 0]
    bits = projs["depth"].astype(np.float32).view(np.uint32)
    order = sorted(vis, key=lambda i: (int(bits[i]), i))
    passing = composited = 0
    for y in range(h):
        for x in range(w):
            tx, ty = x // 16, y // 16
            px, py = f(x) + f(0.5), f(y) + f(0.5)
            t = f(1.0)
            for i in order:
                p = projs[i]
                if not (p["tile_min"][0] <= tx < p["tile_max"][0] and
                        p["tile_min"][1] <= ty < p["tile_max"][1]):
                    continue
                m, cn, o = p["mean2d"].astype(f), p["conic"].astype(f), f(p["opacity"])
                dx, dy = f(px - m[0]), f(py - m[1])
                sigma = f(f(f(0.5) * f(f(f(cn[0] * dx) * dx) + f(f(cn[2] * dy) * dy))) +
                          f(f(cn[1] * dx) * dy))
                if sigma < 0:
                    continue
                og = f(o * np.exp(f(-sigma), dtype=np.float32))
                if og < f(f(1.0) / f(255.0)):
                    continue
                passing += 1
                test_t = f(t * f(f(1.0) - min(og, f(0.99))))
                if test_t < f(1e-4):
                    break
                composited += 1
                t = test_t
    assert composited > 100
    assert got["fwd_passing"] == passing
    assert got["fwd_composited"] == composited
    assert got["bwd_passing"] == composited
    assert got["fwd_evaluated"] >= passing and got["bwd_evaluated"] >= got["bwd_passing"]
