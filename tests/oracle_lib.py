"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker; nothing in the library imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from splatrs_amd import records as rec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u64, u32, i32, f32, f64 = C.c_uint64, C.c_uint32, C.c_int, C.c_float, C.c_double
        sig = {
            "oracle_sh_basis_f32": (None, [i32, P, P]),
            "oracle_sh_basis_grad_f32": (None, [i32, P, P]),
            "oracle_sh_basis_f64": (None, [i32, P, P]),
            "oracle_sh_basis_grad_f64": (None, [i32, P, P]),
            "oracle_project_all_f32": (i32, [P, u64, P, P, P]),
            "oracle_project_backward_all_f32": (i32, [P, u64, P, P, P, P]),
            "oracle_bin_f32": (i32, [P, u64, u32, u32, P]),
            "oracle_rasterize_forward_f32": (i32, [P, u64, P, P, P, P, P]),
            "oracle_rasterize_backward_f32": (i32, [P, u64, P, P, P, P, P, P, P]),
            "oracle_blur_f32": (i32, [P, u32, u32, u32, P]),
            "oracle_gaussian_kernel_f32": (None, [P]),
            "oracle_ssim_with_grad_f32": (i32, [P, P, u32, u32, P, P]),
            "oracle_photometric_loss_f32": (i32, [P, P, u32, u32, f32, P, P]),
            "oracle_raster_pair_counts": (i32, [P, u64, P, P, P, P]),
            "oracle_resize_rgb8": (i32, [P, u32, u32, u32, u32, P]),
            "oracle_loss_sums_f32": (i32, [P, P, u32, u32, P]),
            "oracle_psnr_f32": (f32, [P, P, u64]),
            "oracle_adam_step_f32": (i32, [P, P, P, P, u64, P, P]),
            "oracle_stats_accumulate_f32": (i32, [P, P, P, u64, u32, u32, P, P, P]),
            "oracle_densify_and_prune_f32": (i32, [P, P, P, P, P, P, u64, P, f32, P, u64, P, P,
                                                   P, u64, P, P]),
            "oracle_reset_opacity_f32": (i32, [P, P, P, u64, f32]),
            "oracle_sigmoid_f32": (f32, [f32]),
            "oracle_inverse_sigmoid_f32": (f32, [f32]),
            "oracle_powi_f32": (f32, [f32, i32]),
            "oracle_mean_nn_dist2": (None, [P, u64, P]),
            "oracle_project_all_f64": (None, [P, u64, P, P, P]),
            "oracle_project_backward_all_f64": (None, [P, u64, P, P, P, P]),
            "oracle_render_projs_f64": (i32, [P, u64, u32, u32, P, P, P, P]),
            "oracle_raster_backward_projs_f64": (i32, [P, u64, u32, u32, P, P, P, P]),
            "oracle_ssim_with_grad_f64": (f64, [P, P, u32, u32, P]),
            "oracle_photometric_loss_f64": (f64, [P, P, u32, u32, f64, P]),
            "oracle_chain_loss_f64": (f64, [P, u64, P, P, P, P, f64, P]),
            "oracle_train_step_f32": (f32, [P, P, P, u64, P, P, P, i32, i32, f32, P, P, i32,
                                            i32, P, P, P, P]),
            "oracle_mlp_forward_f32": (None, [P, P, C.c_int64, P, i32]),
            "oracle_mlp_loss_grad_f32": (f32, [P, P, P, P, C.c_int64, P, i32]),
            "oracle_mlp_forward_f64": (None, [P, P, C.c_int64, P]),
            "oracle_mlp_loss_grad_f64": (f64, [P, P, P, P, C.c_int64, P]),
            "oracle_mlp_adam": (None, [P, P, P, P, i32, f32, f32, f32, f32, f32, f32]),
            "oracle_bf16_round": (None, [P, P, C.c_int64]),
            "oracle_set_threads": (i32, [i32]),
            "oracle_rasterize_backward_truth": (i32, [P, u64, P, P, P, P, P, P]),
            "oracle_step_gradients": (i32, [P, u64, P, P, i32, f32, P, P, P, P, P]),
            "oracle_scorer_features_f32": (None, [P, C.c_int64, P, P, P, P, P, P, f32, f32, u32,
                                                  P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """Data pointer of a numpy array or 0-d structured record."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def check(status, what=""):
    if status != 0:
        raise RuntimeError(f"oracle {what} failed with status {status}")


# ---- f32 wrappers with the reference's signatures ------------------------------------

def project_all(g, cam, opts):
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros(len(g), rec.PROJ)
    check(lib().oracle_project_all_f32(ptr(g), len(g), ptr(cam), ptr(opts), ptr(out)))
    return out


def project_backward_all(g, cam, opts, grads):
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros((len(g), rec.NUM_PARAMS), np.float32)
    check(lib().oracle_project_backward_all_f32(ptr(g), len(g), ptr(cam), ptr(opts),
                                                ptr(grads), ptr(out)))
    return out


def bin(projs, width, height):
    from splatrs_amd import abi
    n = len(projs)
    tx, ty = (width + 15) // 16, (height + 15) // 16
    cap = 1024
    while True:
        b = abi.alloc_binning(n, tx, ty, cap)
        s = b.as_struct()
        st = lib().oracle_bin_f32(ptr(projs), n, width, height, C.byref(s))
        if st == 4:  # capacity
            cap = int(s.num_entries)
            continue
        check(st, "bin")
        return abi.trim_binning(b, int(s.num_entries))


def rasterize_forward(projs, binning, settings):
    w, h = int(settings["width"]), int(settings["height"])
    image = np.zeros((h, w, 3), np.float32)
    final_t = np.zeros((h, w), np.float32)
    n_contrib = np.zeros((h, w), np.uint32)
    s = binning.as_struct()
    check(lib().oracle_rasterize_forward_f32(ptr(projs), len(projs), C.byref(s), ptr(settings),
                                             ptr(image), ptr(final_t), ptr(n_contrib)))
    return image, final_t, n_contrib


def rasterize_backward(projs, binning, settings, final_t, n_contrib, dimage):
    n = len(projs)
    grads = np.zeros(n, rec.PROJ_GRAD)
    absg = np.zeros((n, 2), np.float32)
    s = binning.as_struct()
    dimage = np.ascontiguousarray(dimage, np.float32)
    check(lib().oracle_rasterize_backward_f32(ptr(projs), n, C.byref(s), ptr(settings),
                                              ptr(final_t), ptr(n_contrib), ptr(dimage),
                                              ptr(grads), ptr(absg)))
    return grads, absg


def resize_rgb8(image, out_width, out_height):
    """image::imageops::resize(.., Lanczos3) of an RGB8 image (dataset::load's downscale)."""
    image = np.ascontiguousarray(image, np.uint8)
    h, w = image.shape[:2]
    out = np.zeros((out_height, out_width, 3), np.uint8)
    check(lib().oracle_resize_rgb8(ptr(image), w, h, out_width, out_height, ptr(out)))
    return out


def raster_pair_counts(projs, binning, settings, n_contrib):
    """Pair counts of the compositing walks: forward evaluated, passing and composited
    pairs; backward evaluated and passing pairs (a dict)."""
    c = np.zeros(5, np.uint64)
    s = binning.as_struct()
    check(lib().oracle_raster_pair_counts(ptr(projs), len(projs), C.byref(s), ptr(settings),
                                          ptr(np.ascontiguousarray(n_contrib, np.uint32)),
                                          ptr(c)))
    keys = ("fwd_evaluated", "fwd_passing", "fwd_composited", "bwd_evaluated", "bwd_passing")
    return {k: int(v) for k, v in zip(keys, c)}


def rasterize_backward_truth(projs, binning, settings, n_contrib, dimage):
    """Binary64 accuracy reference of rasterize_backward: the binary32 forward's decisions
    (n_contrib and the per-pair tests) with every value in binary64.  Returns (grads n x 9,
    abs n x 2) as float64."""
    n = len(projs)
    grads = np.zeros((n, 9), np.float64)
    absg = np.zeros((n, 2), np.float64)
    s = binning.as_struct()
    check(lib().oracle_rasterize_backward_truth(
        ptr(projs), n, C.byref(s), ptr(settings), ptr(np.ascontiguousarray(n_contrib, np.uint32)),
        ptr(np.ascontiguousarray(dimage, np.float32)), ptr(grads), ptr(absg)), "truth")
    return grads, absg


def step_gradients(params, cam, rgb, sh_degree, lam=0.2, bg=(0, 0, 0), ref=True, truth=True,
                   screen=False):
    """Parameter gradients of one training step's loss at `params` (no Adam): the
    reference's binary32 arithmetic (n x 59 f32) and the binary64 truth (n x 59 f64) of the
    binary32 forward's branch.  With screen=True also the screen-space gradients
    (ProjGrad records, n x 9 f64)."""
    params = np.ascontiguousarray(params, np.float32)
    n = len(params)
    gr = np.zeros((n, 59), np.float32) if ref else None
    gt = np.zeros((n, 59), np.float64) if truth else None
    pr = np.zeros(n, rec.PROJ_GRAD) if screen else None
    pt = np.zeros((n, 9), np.float64) if screen else None
    bgv = np.asarray(bg, np.float32)
    check(lib().oracle_step_gradients(ptr(params), n, ptr(cam),
                                      ptr(np.ascontiguousarray(rgb, np.uint8)), sh_degree, lam,
                                      ptr(bgv), ptr(gr), ptr(gt), ptr(pr), ptr(pt)), "grads")
    out = {"ref": gr, "true": gt}
    if screen:
        out.update(screen_ref=pr, screen_true=pt)
    return out


def blur(src, width, height, channels):
    src = np.ascontiguousarray(src, np.float32)
    dst = np.zeros_like(src)
    check(lib().oracle_blur_f32(ptr(src), width, height, channels, ptr(dst)))
    return dst


def gaussian_kernel():
    k = np.zeros(11, np.float32)
    lib().oracle_gaussian_kernel_f32(ptr(k))
    return k


def ssim_with_grad(x, y, width, height):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = np.zeros_like(x)
    s = np.zeros(1, np.float32)
    check(lib().oracle_ssim_with_grad_f32(ptr(x), ptr(y), width, height, ptr(s), ptr(g)))
    return float(s[0]), g


def photometric_loss(pred, target, width, height, lam):
    pred = np.ascontiguousarray(pred, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    g = np.zeros_like(pred)
    loss = np.zeros(1, np.float32)
    check(lib().oracle_photometric_loss_f32(ptr(pred), ptr(target), width, height, lam,
                                            ptr(loss), ptr(g)))
    return float(loss[0]), g


def loss_sums64(pred, target, width, height):
    """Binary64 sums of the binary32 SSIM map and |pred - target| terms of the loss."""
    pred = np.ascontiguousarray(pred, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    sums = np.zeros(2, np.float64)
    check(lib().oracle_loss_sums_f32(ptr(pred), ptr(target), width, height, ptr(sums)))
    return float(sums[0]), float(sums[1])


def psnr(pred, target):
    pred = np.ascontiguousarray(pred, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    return float(lib().oracle_psnr_f32(ptr(pred), ptr(target), pred.size))


def adam_step(params, grads, m, v, state, lr):
    """In place on params, m, v and state (a 0-d ADAM_STATE record)."""
    check(lib().oracle_adam_step_f32(ptr(params), ptr(grads), ptr(m), ptr(v), len(params),
                                     ptr(state), ptr(lr)))


def stats_accumulate(projs, grads, abs_grads, width, height, grad_accum, denom, max_radii):
    check(lib().oracle_stats_accumulate_f32(ptr(projs), ptr(grads), ptr(abs_grads), len(projs),
                                            width, height, ptr(grad_accum), ptr(denom),
                                            ptr(max_radii)))


def densify_and_prune(params, m, v, grad_accum, denom, max_radii, cfg, extent, normals):
    n = len(params)
    normals = np.ascontiguousarray(normals, np.float32)
    cap = 3 * n + 16
    po = np.zeros((cap, rec.NUM_PARAMS), np.float32)
    mo = np.zeros_like(po)
    vo = np.zeros_like(po)
    n_out = np.zeros(1, np.uint64)
    report = np.zeros((), rec.DENSIFY_REPORT)
    check(lib().oracle_densify_and_prune_f32(
        ptr(np.ascontiguousarray(params, np.float32)), ptr(np.ascontiguousarray(m, np.float32)),
        ptr(np.ascontiguousarray(v, np.float32)), ptr(np.ascontiguousarray(grad_accum, np.float32)),
        ptr(np.ascontiguousarray(denom, np.uint32)), ptr(np.ascontiguousarray(max_radii, np.uint32)),
        n, ptr(cfg), extent, ptr(normals), normals.size, ptr(po), ptr(mo), ptr(vo), cap,
        ptr(n_out), ptr(report)), "densify")
    k = int(n_out[0])
    return po[:k].copy(), mo[:k].copy(), vo[:k].copy(), report


class TrainState:
    """Model, Adam moments and statistics of the oracle trainer (AoS, host)."""

    def __init__(self, gaussians):
        self.params = np.ascontiguousarray(gaussians, np.float32).copy()
        self.m = np.zeros_like(self.params)
        self.v = np.zeros_like(self.params)
        n = len(self.params)
        self.grad_accum = np.zeros(n, np.float32)
        self.denom = np.zeros(n, np.uint32)
        self.max_radii = np.zeros(n, np.uint32)
        self.step = np.zeros(1, np.uint64)

    def train_step(self, cam, rgb, sh_degree, lr, antialias=False, lam=0.2, bg=(0, 0, 0),
                   accumulate_stats=False, absgrad=False):
        """One iteration of src/train.rs:211-247; returns (loss, num_entries)."""
        rgb = np.ascontiguousarray(rgb, np.uint8)
        bgv = np.asarray(bg, np.float32)
        ne = np.zeros(1, np.uint64)
        loss = lib().oracle_train_step_f32(
            ptr(self.params), ptr(self.m), ptr(self.v), len(self.params), ptr(self.step),
            ptr(cam), ptr(rgb), sh_degree, int(antialias), lam, ptr(bgv), ptr(lr),
            int(accumulate_stats), int(absgrad), ptr(self.grad_accum), ptr(self.denom),
            ptr(self.max_radii), ptr(ne))
        return float(loss), int(ne[0])


def reset_opacity(params, m, v, max_alpha):
    check(lib().oracle_reset_opacity_f32(ptr(params), ptr(m), ptr(v), len(params), max_alpha))


def mean_nn_dist2(points):
    pts = np.ascontiguousarray(points, np.float64)
    out = np.zeros(len(pts), np.float64)
    lib().oracle_mean_nn_dist2(ptr(pts), len(pts), ptr(out))
    return out


# ---- f64 ---------------------------------------------------------------------------

def cam64(cam):
    c = np.zeros(18, np.float64)
    c[0], c[1] = cam["width"], cam["height"]
    c[2], c[3], c[4], c[5] = cam["fx"], cam["fy"], cam["cx"], cam["cy"]
    c[6:15] = cam["rot"]
    c[15:18] = cam["trans"]
    return c


def opts64(sh_degree=3, antialias=False, near_plane=0.2, eps2d=0.3):
    return np.array([near_plane, eps2d, 1.0 if antialias else 0.0, sh_degree], np.float64)


def project_all_f64(g, cam64_, opts64_):
    g = np.ascontiguousarray(g, np.float64)
    out = np.zeros((len(g), 15), np.float64)
    lib().oracle_project_all_f64(ptr(g), len(g), ptr(cam64_), ptr(opts64_), ptr(out))
    return out


def project_backward_all_f64(g, cam64_, opts64_, grads9):
    g = np.ascontiguousarray(g, np.float64)
    grads9 = np.ascontiguousarray(grads9, np.float64)
    out = np.zeros((len(g), 59), np.float64)
    lib().oracle_project_backward_all_f64(ptr(g), len(g), ptr(cam64_), ptr(opts64_),
                                          ptr(grads9), ptr(out))
    return out


def render_projs_f64(projs15, width, height, bg):
    projs15 = np.ascontiguousarray(projs15, np.float64)
    bg = np.asarray(bg, np.float64)
    image = np.zeros((height, width, 3))
    ft = np.zeros((height, width))
    nc = np.zeros((height, width), np.uint32)
    check(lib().oracle_render_projs_f64(ptr(projs15), len(projs15), width, height, ptr(bg),
                                        ptr(image), ptr(ft), ptr(nc)))
    return image, ft, nc


def raster_backward_projs_f64(projs15, width, height, bg, dimage):
    projs15 = np.ascontiguousarray(projs15, np.float64)
    bg = np.asarray(bg, np.float64)
    dimage = np.ascontiguousarray(dimage, np.float64)
    grads = np.zeros((len(projs15), 9))
    absg = np.zeros((len(projs15), 2))
    check(lib().oracle_raster_backward_projs_f64(ptr(projs15), len(projs15), width, height,
                                                 ptr(bg), ptr(dimage), ptr(grads), ptr(absg)))
    return grads, absg


def ssim_with_grad_f64(x, y, width, height):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    g = np.zeros_like(x)
    s = lib().oracle_ssim_with_grad_f64(ptr(x), ptr(y), width, height, ptr(g))
    return s, g


def photometric_loss_f64(pred, target, width, height, lam):
    pred = np.ascontiguousarray(pred, np.float64)
    target = np.ascontiguousarray(target, np.float64)
    g = np.zeros_like(pred)
    loss = lib().oracle_photometric_loss_f64(ptr(pred), ptr(target), width, height, lam, ptr(g))
    return loss, g


def chain_loss_f64(g, cam64_, opts64_, bg, target, lam, want_grad=False):
    g = np.ascontiguousarray(g, np.float64)
    bg = np.asarray(bg, np.float64)
    target = np.ascontiguousarray(target, np.float64)
    grads = np.zeros((len(g), 59)) if want_grad else None
    loss = lib().oracle_chain_loss_f64(ptr(g), len(g), ptr(cam64_), ptr(opts64_), ptr(bg),
                                       ptr(target), lam, ptr(grads))
    return (loss, grads) if want_grad else loss


# ---- learned densification scorer (12 -> 64 -> 64 -> 64 -> 3, tanh GELU) ----------------

MLP_PARAMS = 9347


def mlp_forward(params, feats, bf16=True):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(feats, np.float32)
    out = np.zeros((len(x), 3), np.float32)
    lib().oracle_mlp_forward_f32(ptr(p), ptr(x), len(x), ptr(out), int(bf16))
    return out


def mlp_loss_grad(params, feats, labels, weights, bf16=True):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(feats, np.float32)
    y = np.ascontiguousarray(labels, np.uint8)
    w = np.ascontiguousarray(weights, np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = lib().oracle_mlp_loss_grad_f32(ptr(p), ptr(x), ptr(y), ptr(w), len(x), ptr(g), int(bf16))
    return loss, g


def mlp_forward_f64(params, feats):
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(feats, np.float64)
    out = np.zeros((len(x), 3), np.float64)
    lib().oracle_mlp_forward_f64(ptr(p), ptr(x), len(x), ptr(out))
    return out


def mlp_loss_grad_f64(params, feats, labels, weights):
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(feats, np.float64)
    y = np.ascontiguousarray(labels, np.uint8)
    w = np.ascontiguousarray(weights, np.float64)
    g = np.zeros(MLP_PARAMS, np.float64)
    loss = lib().oracle_mlp_loss_grad_f64(ptr(p), ptr(x), ptr(y), ptr(w), len(x), ptr(g))
    return loss, g


def mlp_adam(p, g, m, v, lr, c1, c2, b1=0.9, b2=0.999, eps=1e-8):
    lib().oracle_mlp_adam(ptr(p), ptr(g), ptr(m), ptr(v), len(p), lr, b1, b2, eps, c1, c2)


def bf16_round(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().oracle_bf16_round(ptr(x), ptr(out), x.size)
    return out


def scorer_features(g, grad_accum, denom, max_radii, max_grad, strong, birth, thr, log_size,
                    iteration):
    g = np.ascontiguousarray(g, np.float32)
    out = np.zeros((len(g), 12), np.float32)
    arrs = [np.ascontiguousarray(a, t) for a, t in
            [(grad_accum, np.float32), (denom, np.uint32), (max_radii, np.uint32),
             (max_grad, np.float32), (strong, np.uint32), (birth, np.uint32)]]
    lib().oracle_scorer_features_f32(ptr(g), len(g), *[ptr(a) for a in arrs], thr, log_size,
                                     iteration, ptr(out))
    return out


def set_threads(n):
    """Threads of the oracle's parallel loops (1 = scalar); returns the count in use."""
    return lib().oracle_set_threads(n)
