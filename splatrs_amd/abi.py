"""ctypes binding of libsplat.so (include/splat.h) and a Python mirror of the reference's
host API (src/project.rs, src/rasterize.rs, src/render.rs, src/loss.rs, src/adam.rs,
src/densify.rs), for tests and tools.

Every function here calls the HIP library; there is no CPU fallback.  Loading fails
loudly when libsplat.so is missing or no gfx950 device is present.
"""
import ctypes as C
import os

import numpy as np

from . import records as rec

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplat.so")

P = C.c_void_p
u64, u32, i32, f32, f64 = C.c_uint64, C.c_uint32, C.c_int, C.c_float, C.c_double

OK, ERR_PRECONDITION, ERR_HIP, ERR_OOM, ERR_CAPACITY, ERR_INVALID, ERR_IO, ERR_FORMAT, \
    ERR_NO_DEVICE = range(9)


class SplatError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"libsplat status {status}: {msg}")
        self.status = status


class PreconditionError(SplatError):
    """The reference panics (assert!) on this input."""


class BinningStruct(C.Structure):
    _fields_ = [("tiles_x", u32), ("tiles_y", u32), ("num_entries", u64), ("capacity", u64),
                ("entry_gaussian", P), ("entry_offset", P), ("sorted_keys", P),
                ("sorted_entries", P), ("tile_range", P)]


class Binning:
    """Binning, src/rasterize.rs:26-39, with host numpy arrays."""

    def __init__(self, tiles_x, tiles_y, entry_gaussian, entry_offset, sorted_keys,
                 sorted_entries, tile_range):
        self.tiles_x, self.tiles_y = int(tiles_x), int(tiles_y)
        self.entry_gaussian = entry_gaussian
        self.entry_offset = entry_offset
        self.sorted_keys = sorted_keys
        self.sorted_entries = sorted_entries
        self.tile_range = tile_range

    @property
    def num_entries(self):
        return len(self.sorted_keys)

    def as_struct(self):
        return binning_struct(self)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def alloc_binning(n, tiles_x, tiles_y, capacity):
    return Binning(tiles_x, tiles_y, np.zeros(capacity, np.uint32), np.zeros(n + 1, np.uint32),
                   np.zeros(capacity, np.uint64), np.zeros(capacity, np.uint32),
                   np.zeros((tiles_x * tiles_y, 2), np.uint32))


def trim_binning(b, e):
    return Binning(b.tiles_x, b.tiles_y, b.entry_gaussian[:e].copy(), b.entry_offset,
                   b.sorted_keys[:e].copy(), b.sorted_entries[:e].copy(), b.tile_range)


def binning_struct(b):
    s = BinningStruct()
    s.tiles_x, s.tiles_y = b.tiles_x, b.tiles_y
    s.num_entries = len(b.sorted_keys)
    s.capacity = len(b.sorted_keys)
    s.entry_gaussian = _p(b.entry_gaussian)
    s.entry_offset = _p(b.entry_offset)
    s.sorted_keys = _p(b.sorted_keys)
    s.sorted_entries = _p(b.sorted_entries)
    s.tile_range = _p(b.tile_range)
    return s


class TrainConfig(C.Structure):
    """splat_train_config (TrainConfig, src/train.rs:20-109)."""
    _fields_ = [("iterations", u32), ("position_lr_init", f32), ("position_lr_final", f32),
                ("position_lr_max_steps", u32), ("feature_lr", f32), ("opacity_lr", f32),
                ("scaling_lr", f32), ("rotation_lr", f32), ("lambda_dssim", f32),
                ("densify", i32), ("densify_from_iter", u32), ("densify_until_iter", u32),
                ("densification_interval", u32), ("opacity_reset_interval", u32),
                ("densify_grad_threshold", f32), ("percent_dense", f32), ("min_opacity", f32),
                ("has_max_screen_size", i32), ("max_screen_size", u32), ("absgrad", i32),
                ("antialias", i32), ("white_background", i32), ("sh_degree", i32),
                ("sh_degree_interval", u32), ("near_plane", f32), ("seed", u64),
                ("save_iterations", P), ("num_save_iterations", u32), ("log_interval", u32),
                ("ordered_backward", i32), ("start_iteration", u32)]

    @classmethod
    def default(cls, **overrides):
        c = cls()
        lib().splat_train_config_default(C.byref(c))
        for k, v in overrides.items():
            if k == "save_iterations":
                arr = (u32 * len(v))(*v)
                c._save = arr  # keep alive
                c.save_iterations = C.cast(arr, P)
                c.num_save_iterations = len(v)
            else:
                setattr(c, k, v)
        return c


class Progress(C.Structure):
    """splat_progress (Progress, src/train.rs:140-157)."""
    _fields_ = [("iteration", u32), ("loss", f32), ("ema_loss", f32), ("psnr", f32),
                ("num_gaussians", u64), ("ms_per_iter", f64), ("cloned", u64), ("split", u64),
                ("pruned", u64)]


PROGRESS_FN = C.CFUNCTYPE(None, C.POINTER(Progress), P)
SAVE_FN = C.CFUNCTYPE(C.c_int, u32, P, P)


# ---- library loading -----------------------------------------------------------------

_SIGS = {
    "splat_ctx_create": (i32, [i32, P]),
    "splat_ctx_destroy": (i32, [P]),
    "splat_last_error": (C.c_char_p, [P]),
    "splat_sync": (i32, [P]),
    "splat_ctx_stream": (P, [P]),
    "splat_set_timing": (i32, [P, i32]),
    "splat_set_bin_method": (i32, [P, i32]),
    "splat_reset_timing": (i32, [P]),
    "splat_kernel_times": (i32, [P, P, C.c_size_t, P, P, u32, P]),
    "splat_project_all": (i32, [P, P, u64, P, P, P]),
    "splat_project_backward_all": (i32, [P, P, u64, P, P, P, P]),
    "splat_bin": (i32, [P, P, u64, u32, u32, P]),
    "splat_rasterize_forward": (i32, [P, P, u64, P, P, P, P, P]),
    "splat_rasterize_backward": (i32, [P, P, u64, P, P, P, P, P, i32, P, P]),
    "splat_render": (i32, [P, P, u64, P, P, P, P]),
    "splat_to_rgb8": (i32, [P, u64, P]),
    "splat_photometric_loss": (i32, [P, P, P, u32, u32, f32, P, P]),
    "splat_photometric_loss_device_noalloc": (i32, [P, P, P, u32, u32, f32, P, P]),
    "splat_ssim_with_grad": (i32, [P, P, P, u32, u32, P, P]),
    "splat_blur": (i32, [P, P, u32, u32, u32, P]),
    "splat_psnr": (i32, [P, P, u64, P]),
    "splat_adam_step": (i32, [P, P, P, P, P, u64, P, P]),
    "splat_stats_accumulate": (i32, [P, P, P, P, u64, u32, u32, P, P, P]),
    "splat_densify_and_prune": (i32, [P, P, P, P, P, P, P, u64, P, f32, P, u64, P, P, P, u64,
                                      P, P]),
    "splat_reset_opacity": (i32, [P, P, P, P, u64, f32]),
    "splat_mean_nn_dist2": (i32, [P, P, u64, P]),
    "splat_from_points": (i32, [P, P, P, u64, P]),
    "splat_train_config_default": (None, [P]),
    "splat_train": (i32, [P, f32, P, P, P, P]),
    "splat_train_view_at": (u32, [u64, u32, u32]),
    "splat_ply_write": (i32, [C.c_char_p, P, u64, i32, C.c_char_p, C.c_size_t]),
    "splat_ply_read": (i32, [C.c_char_p, P, u64, P, P, C.c_char_p, C.c_size_t]),
    "splat_checkpoint_write": (i32, [C.c_char_p, P, P, P, P, P, P, u64, u64, i32, C.c_char_p,
                                     C.c_size_t]),
    "splat_checkpoint_read": (i32, [C.c_char_p, P, P, P, P, P, P, u64, P, P, P, C.c_char_p,
                                    C.c_size_t]),
    "splat_trainer_restore": (i32, [P, P, P, P, P, P, P, u64, u64]),
    "splat_colmap_read": (i32, [C.c_char_p, P, u64, P, P, u64, P, P, u64, P, C.c_char_p,
                                C.c_size_t]),
    "splat_colmap_write": (i32, [C.c_char_p, i32, P, u64, P, u64, P, u64, C.c_char_p,
                                 C.c_size_t]),
    "splat_camera_from_colmap": (i32, [P, u64, P, P, C.c_char_p, C.c_size_t]),
    "splat_camera_match_image_size": (i32, [P, u32, u32, P]),
    "splat_resize_rgb8": (i32, [P, P, u32, u32, u32, u32, P]),
    "splat_downscaled_size": (i32, [u32, u32, u32, P, P]),
    "splat_trainer_add_view_downscaled": (i32, [P, P, P, u32, u32, u32, P, P]),
    "splat_split_eval": (i32, [u64, u32, P]),
    "splat_trainer_view": (i32, [P, u32, P, P, u64]),
    "splat_colmap_sort_images": (i32, [P, u64]),
    "splat_trainer_evaluate": (i32, [P, P, u32, i32, i32, P, P, P]),
    "splat_trainer_create": (i32, [P, P, u64, u64, P]),
    "splat_trainer_destroy": (i32, [P]),
    "splat_trainer_add_view": (i32, [P, P, P, P]),
    "splat_trainer_step": (i32, [P, u32, P, P]),
    "splat_trainer_densify": (i32, [P, P, f32, u64, P]),
    "splat_trainer_reset_opacity": (i32, [P, f32]),
    "splat_trainer_render": (i32, [P, u32, i32, i32, P, P]),
    "splat_trainer_info": (i32, [P, P, P]),
    "splat_trainer_set_graph": (i32, [P, i32]),
    "splat_trainer_graph_stats": (i32, [P, P, P, P]),
    "splat_trainer_reserve_entries": (i32, [P, u64]),
    "splat_trainer_sync": (i32, [P]),
    "splat_scorer_create": (i32, [P, P, u64, P]),
    "splat_scorer_destroy": (i32, [P]),
    "splat_scorer_get_params": (i32, [P, P]),
    "splat_scorer_set_params": (i32, [P, P]),
    "splat_scorer_forward": (i32, [P, P, u64, P]),
    "splat_scorer_loss_grad": (i32, [P, P, P, P, u64, P, P]),
    "splat_scorer_fit_config_default": (None, [P]),
    "splat_scorer_fit": (i32, [P, P, P, P, u64, P, P]),
    "splat_learn_config_default": (None, [P]),
    "splat_trainer_tune": (i32, [P, u32, P, P]),
    "splat_trainer_pair_counts": (i32, [P, u32, P, P]),
    "splat_trainer_set_shapes": (i32, [P, P]),
    "splat_batch_create": (i32, [P, P]),
    "splat_batch_destroy": (i32, [P]),
    "splat_batch_add_scene": (i32, [P, P, u64, P]),
    "splat_batch_add_view": (i32, [P, u32, P, P, P]),
    "splat_batch_step": (i32, [P, P, u32, P, P, P]),
    "splat_batch_tune": (i32, [P, P, u32, P, P]),
    "splat_batch_set_shapes": (i32, [P, P]),
    "splat_batch_download": (i32, [P, u32, P, P, P, u64]),
    "splat_batch_info": (i32, [P, P, P, P, P, P]),
    "splat_trainer_attach_scorer": (i32, [P, P, P]),
    "splat_trainer_learn_stats": (i32, [P, P]),
    "splat_trainer_features": (i32, [P, f32, f32, f32, u32, P]),
    "splat_trainer_download_lineage": (i32, [P, P, P, P, P, u64]),
    "splat_trainer_learn_samples": (i32, [P, P, P, P, u64, P]),
    "splat_trainer_download": (i32, [P, P, P, P, P, P, P, u64]),
}

_lib = None


def lib():
    """The loaded libsplat.so.  Raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Function names declared in include/splat.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "splat.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(splat_\w+)\s*\(", text, re.M)))


class Context:
    """A libsplat context (one HIP stream on one device)."""

    def __init__(self, device=0):
        import weakref
        self._h = C.c_void_p()
        self._trainers = weakref.WeakSet()
        st = lib().splat_ctx_create(device, C.byref(self._h))
        if st != OK:
            raise SplatError(st, f"splat_ctx_create(device={device}) failed")

    @property
    def handle(self):
        return self._h

    def close(self):
        for t in list(self._trainers):  # trainers and scorers use this context's stream
            t.close()
        if self._h:
            lib().splat_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, status, what):
        if status == OK:
            return
        msg = lib().splat_last_error(self._h)
        msg = msg.decode() if msg else ""
        cls = PreconditionError if status == ERR_PRECONDITION else SplatError
        raise cls(status, f"{what}: {msg}")

    def sync(self):
        self.check(lib().splat_sync(self._h), "sync")

    # ---- timing ------------------------------------------------------------------------
    def set_timing(self, on):
        self.check(lib().splat_set_timing(self._h, 1 if on else 0), "set_timing")

    def reset_timing(self):
        self.check(lib().splat_reset_timing(self._h), "reset_timing")

    def set_bin_method(self, method):
        """0: per-tile counting and sorts (default); 1: radix sorts (same result)."""
        self.check(lib().splat_set_bin_method(self._h, int(method)), "set_bin_method")

    def kernel_times(self):
        cap = 256
        names = C.create_string_buffer(64 * cap)
        launches = np.zeros(cap, np.uint64)
        total = np.zeros(cap, np.float64)
        count = C.c_uint32()
        self.check(lib().splat_kernel_times(self._h, names, len(names), _p(launches), _p(total),
                                            cap, C.byref(count)), "kernel_times")
        raw = names.raw.split(b"\0")
        out = {}
        for i in range(count.value):
            out[raw[i].decode()] = (int(launches[i]), float(total[i]))
        return out

    # ---- reference API -----------------------------------------------------------------
    def project_all(self, gaussians, cam, opts):
        """project_all, src/project.rs:320-330."""
        g = np.ascontiguousarray(gaussians, np.float32)
        out = np.zeros(len(g), rec.PROJ)
        self.check(lib().splat_project_all(self._h, _p(g), len(g), _p(cam), _p(opts), _p(out)),
                   "project_all")
        return out

    def project_backward_all(self, gaussians, cam, opts, grads):
        """project_backward_all, src/project.rs:477-493."""
        g = np.ascontiguousarray(gaussians, np.float32)
        out = np.zeros((len(g), rec.NUM_PARAMS), np.float32)
        self.check(lib().splat_project_backward_all(self._h, _p(g), len(g), _p(cam), _p(opts),
                                                    _p(grads), _p(out)), "project_backward_all")
        return out

    def bin(self, projs, width, height):
        """bin, src/rasterize.rs:60-128."""
        n = len(projs)
        tx, ty = (width + 15) // 16, (height + 15) // 16
        cap = max(1024, 4 * n)
        while True:
            b = alloc_binning(n, tx, ty, cap)
            s = b.as_struct()
            st = lib().splat_bin(self._h, _p(projs), n, width, height, C.byref(s))
            if st == ERR_CAPACITY:
                cap = int(s.num_entries)
                continue
            self.check(st, "bin")
            return trim_binning(b, int(s.num_entries))

    def rasterize_forward(self, projs, binning, settings):
        """rasterize_forward, src/rasterize.rs:305-337 -> (image, final_t, n_contrib)."""
        w, h = int(settings["width"]), int(settings["height"])
        image = np.zeros((h, w, 3), np.float32)
        final_t = np.zeros((h, w), np.float32)
        n_contrib = np.zeros((h, w), np.uint32)
        s = binning.as_struct()
        self.check(lib().splat_rasterize_forward(self._h, _p(projs), len(projs), C.byref(s),
                                                 _p(settings), _p(image), _p(final_t),
                                                 _p(n_contrib)), "rasterize_forward")
        return image, final_t, n_contrib

    def rasterize_backward(self, projs, binning, settings, final_t, n_contrib, dimage,
                           ordered=True):
        """rasterize_backward, src/rasterize.rs:459-505 -> (grads, abs_grads)."""
        n = len(projs)
        grads = np.zeros(n, rec.PROJ_GRAD)
        absg = np.zeros((n, 2), np.float32)
        s = binning.as_struct()
        dimage = np.ascontiguousarray(dimage, np.float32)
        self.check(lib().splat_rasterize_backward(self._h, _p(projs), n, C.byref(s),
                                                  _p(settings), _p(final_t), _p(n_contrib),
                                                  _p(dimage), 1 if ordered else 0, _p(grads),
                                                  _p(absg)), "rasterize_backward")
        return grads, absg

    def render(self, gaussians, cam, opts, background=(0.0, 0.0, 0.0)):
        """render, src/render.rs:14-23."""
        g = np.ascontiguousarray(gaussians, np.float32)
        bg = np.asarray(background, np.float32)
        img = np.zeros((int(cam["height"]), int(cam["width"]), 3), np.float32)
        self.check(lib().splat_render(self._h, _p(g), len(g), _p(cam), _p(opts), _p(bg),
                                      _p(img)), "render")
        return img

    def photometric_loss(self, pred, target, width, height, lambda_dssim):
        """photometric_loss, src/loss.rs:137-171 -> (loss, grad)."""
        pred = np.ascontiguousarray(pred, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        grad = np.zeros_like(pred)
        loss = np.zeros(1, np.float32)
        self.check(lib().splat_photometric_loss(self._h, _p(pred), _p(target), width, height,
                                                lambda_dssim, _p(loss), _p(grad)),
                   "photometric_loss")
        return float(loss[0]), grad

    def photometric_loss_device_noalloc(self, pred_ptr, target_ptr, width, height,
                                        lambda_dssim, loss_ptr, grad_ptr):
        """splat_photometric_loss_device_noalloc on device addresses (integers): enqueued on
        the context's stream; nothing is allocated.  Call sync() before reading."""
        self.check(lib().splat_photometric_loss_device_noalloc(
            self._h, C.c_void_p(pred_ptr), C.c_void_p(target_ptr), width, height,
            lambda_dssim, C.c_void_p(loss_ptr) if loss_ptr else None, C.c_void_p(grad_ptr)),
            "photometric_loss_device_noalloc")

    def resize_rgb8(self, image, out_width, out_height):
        """image::imageops::resize(.., Lanczos3) of an RGB8 image on the GPU."""
        image = np.ascontiguousarray(image, np.uint8)
        h, w = image.shape[:2]
        out = np.zeros((out_height, out_width, 3), np.uint8)
        self.check(lib().splat_resize_rgb8(self._h, _p(image), w, h, out_width, out_height,
                                           _p(out)), "resize_rgb8")
        return out

    def ssim_with_grad(self, x, y, width, height):
        """ssim_with_grad, src/loss.rs:75-134 -> (ssim, grad)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        grad = np.zeros_like(x)
        s = np.zeros(1, np.float32)
        self.check(lib().splat_ssim_with_grad(self._h, _p(x), _p(y), width, height, _p(s),
                                              _p(grad)), "ssim_with_grad")
        return float(s[0]), grad

    def blur(self, src, width, height, channels):
        """blur, src/loss.rs:31-72."""
        src = np.ascontiguousarray(src, np.float32)
        dst = np.zeros_like(src)
        self.check(lib().splat_blur(self._h, _p(src), width, height, channels, _p(dst)), "blur")
        return dst

    def adam_step(self, params, grads, m, v, state, lr):
        """Adam::step, src/adam.rs:77-161, in place on params, m, v, state."""
        grads = np.ascontiguousarray(grads, np.float32)
        self.check(lib().splat_adam_step(self._h, _p(params), _p(grads), _p(m), _p(v),
                                         len(params), _p(state), _p(lr)), "adam_step")

    def stats_accumulate(self, projs, grads, abs_grads, width, height, grad_accum, denom,
                         max_radii):
        """DensifyStats::accumulate, src/densify.rs:36-60, in place."""
        self.check(lib().splat_stats_accumulate(self._h, _p(projs), _p(grads), _p(abs_grads),
                                                len(projs), width, height, _p(grad_accum),
                                                _p(denom), _p(max_radii)), "stats_accumulate")

    def densify_and_prune(self, params, m, v, grad_accum, denom, max_radii, cfg, extent,
                          normals):
        """densify_and_prune, src/densify.rs:125-221 -> (params, m, v, report)."""
        n = len(params)
        normals = np.ascontiguousarray(normals, np.float32)
        cap = 3 * n + 16
        po = np.zeros((cap, rec.NUM_PARAMS), np.float32)
        mo = np.zeros_like(po)
        vo = np.zeros_like(po)
        n_out = np.zeros(1, np.uint64)
        report = np.zeros((), rec.DENSIFY_REPORT)
        self.check(lib().splat_densify_and_prune(
            self._h, _p(np.ascontiguousarray(params, np.float32)),
            _p(np.ascontiguousarray(m, np.float32)), _p(np.ascontiguousarray(v, np.float32)),
            _p(np.ascontiguousarray(grad_accum, np.float32)),
            _p(np.ascontiguousarray(denom, np.uint32)),
            _p(np.ascontiguousarray(max_radii, np.uint32)), n, _p(cfg), extent, _p(normals),
            normals.size, _p(po), _p(mo), _p(vo), cap, _p(n_out), _p(report)),
            "densify_and_prune")
        k = int(n_out[0])
        return po[:k].copy(), mo[:k].copy(), vo[:k].copy(), report

    def mean_nn_dist2(self, xyz):
        """mean_nn_dist2, src/model.rs:128-186 (binary64)."""
        x = np.ascontiguousarray(xyz, np.float32)
        out = np.zeros(len(x), np.float64)
        self.check(lib().splat_mean_nn_dist2(self._h, _p(x), len(x), _p(out)), "mean_nn_dist2")
        return out

    def from_points(self, xyz, rgb):
        """Model::from_points, src/model.rs:77-104 -> n x 59 records."""
        x = np.ascontiguousarray(xyz, np.float32)
        c = np.ascontiguousarray(rgb, np.uint8)
        out = np.zeros((len(x), rec.NUM_PARAMS), np.float32)
        self.check(lib().splat_from_points(self._h, _p(x), _p(c), len(x), _p(out)), "from_points")
        return out

    def reset_opacity(self, params, m, v, max_alpha):
        """reset_opacity, src/densify.rs:224-230 (+ Adam::reset_opacity_state)."""
        self.check(lib().splat_reset_opacity(self._h, _p(params), _p(m), _p(v), len(params),
                                             max_alpha), "reset_opacity")


class PlyError(SplatError):
    pass


def ply_write(path, gaussians, max_sh_degree):
    """ply::write, src/ply.rs:19-69."""
    g = np.ascontiguousarray(gaussians, np.float32)
    err = C.create_string_buffer(512)
    st = lib().splat_ply_write(os.fsencode(path), _p(g), len(g), max_sh_degree, err, len(err))
    if st != OK:
        raise PlyError(st, err.value.decode())


def ply_read(path):
    """ply::read, src/ply.rs:73-181 -> (gaussians n x 59, sh_degree)."""
    err = C.create_string_buffer(512)
    n = C.c_uint64()
    deg = C.c_int32()
    st = lib().splat_ply_read(os.fsencode(path), None, 0, C.byref(n), C.byref(deg), err,
                              len(err))
    if st not in (OK, ERR_CAPACITY):
        raise PlyError(st, err.value.decode())
    g = np.zeros((n.value, rec.NUM_PARAMS), np.float32)
    st = lib().splat_ply_read(os.fsencode(path), _p(g), n.value, C.byref(n), C.byref(deg), err,
                              len(err))
    if st != OK:
        raise PlyError(st, err.value.decode())
    return g, deg.value


class CheckpointError(SplatError):
    pass


def checkpoint_write(path, state, adam_step, max_sh_degree):
    """Resumable checkpoint (splat_checkpoint_write): state as Trainer.download(moments=True,
    stats=True) returns it (missing moments or statistics are written as zeros)."""
    g = np.ascontiguousarray(state["params"], np.float32)
    n = len(g)

    def arr(key, dtype, shape):
        a = state.get(key)
        return None if a is None else np.ascontiguousarray(a, dtype).reshape(shape)
    m, v = arr("m", np.float32, g.shape), arr("v", np.float32, g.shape)
    acc = arr("grad_accum", np.float32, (n,))
    den, mr = arr("denom", np.uint32, (n,)), arr("max_radii", np.uint32, (n,))
    err = C.create_string_buffer(512)
    st = lib().splat_checkpoint_write(os.fsencode(path), _p(g), _p(m), _p(v), _p(acc), _p(den),
                                      _p(mr), n, adam_step, max_sh_degree, err, len(err))
    if st != OK:
        raise CheckpointError(st, err.value.decode())


def checkpoint_read(path):
    """-> (state dict with params, m, v, grad_accum, denom, max_radii; adam_step;
    max_sh_degree)."""
    err = C.create_string_buffer(512)
    n, step, deg = C.c_uint64(), C.c_uint64(), C.c_int32()
    st = lib().splat_checkpoint_read(os.fsencode(path), None, None, None, None, None, None, 0,
                                     C.byref(n), C.byref(step), C.byref(deg), err, len(err))
    if st not in (OK, ERR_CAPACITY):
        raise CheckpointError(st, err.value.decode())
    k = n.value
    out = {"params": np.zeros((k, rec.NUM_PARAMS), np.float32),
           "m": np.zeros((k, rec.NUM_PARAMS), np.float32),
           "v": np.zeros((k, rec.NUM_PARAMS), np.float32),
           "grad_accum": np.zeros(k, np.float32), "denom": np.zeros(k, np.uint32),
           "max_radii": np.zeros(k, np.uint32)}
    st = lib().splat_checkpoint_read(os.fsencode(path), _p(out["params"]), _p(out["m"]),
                                     _p(out["v"]), _p(out["grad_accum"]), _p(out["denom"]),
                                     _p(out["max_radii"]), k, C.byref(n), C.byref(step),
                                     C.byref(deg), err, len(err))
    if st != OK:
        raise CheckpointError(st, err.value.decode())
    return out, step.value, deg.value


class ColmapError(SplatError):
    pass


def colmap_read(directory):
    """read_model (src/colmap/mod.rs:198-238) -> (cameras, images, points) as structured
    arrays of records.COLMAP_CAMERA / COLMAP_IMAGE / COLMAP_POINT."""
    err = C.create_string_buffer(512)
    nc, ni, npt = C.c_uint64(), C.c_uint64(), C.c_uint64()
    d = os.fsencode(directory)
    st = lib().splat_colmap_read(d, None, 0, C.byref(nc), None, 0, C.byref(ni), None, 0,
                                 C.byref(npt), err, len(err))
    if st not in (OK, ERR_CAPACITY):
        raise ColmapError(st, err.value.decode())
    cams = np.zeros(nc.value, rec.COLMAP_CAMERA)
    imgs = np.zeros(ni.value, rec.COLMAP_IMAGE)
    pts = np.zeros(npt.value, rec.COLMAP_POINT)
    st = lib().splat_colmap_read(d, _p(cams), len(cams), C.byref(nc), _p(imgs), len(imgs),
                                 C.byref(ni), _p(pts), len(pts), C.byref(npt), err, len(err))
    if st != OK:
        raise ColmapError(st, err.value.decode())
    return cams, imgs, pts


def colmap_write(directory, cameras, images, points, binary=True):
    """write_model_bin / write_model_txt (src/colmap/binary.rs:165-224, text.rs:132-190)."""
    c = np.ascontiguousarray(cameras, rec.COLMAP_CAMERA)
    i = np.ascontiguousarray(images, rec.COLMAP_IMAGE)
    p = np.ascontiguousarray(points, rec.COLMAP_POINT)
    err = C.create_string_buffer(512)
    st = lib().splat_colmap_write(os.fsencode(directory), int(bool(binary)), _p(c), len(c),
                                  _p(i), len(i), _p(p), len(p), err, len(err))
    if st != OK:
        raise ColmapError(st, err.value.decode())


def camera_from_colmap(cameras, image):
    """camera_from_colmap (src/dataset.rs:83-129) -> records.CAMERA."""
    c = np.ascontiguousarray(cameras, rec.COLMAP_CAMERA)
    im = np.ascontiguousarray(image, rec.COLMAP_IMAGE).reshape(())
    out = np.zeros((), rec.CAMERA)
    err = C.create_string_buffer(512)
    st = lib().splat_camera_from_colmap(_p(c), len(c), _p(im), _p(out), err, len(err))
    if st != OK:
        raise ColmapError(st, err.value.decode())
    return out


def match_image_size(camera, width, height):
    """match_image_size (src/dataset.rs:213-237) -> records.CAMERA."""
    cam = np.ascontiguousarray(camera, rec.CAMERA).reshape(())
    out = np.zeros((), rec.CAMERA)
    st = lib().splat_camera_match_image_size(_p(cam), width, height, _p(out))
    if st != OK:
        raise ColmapError(st, f"image size {width}x{height} does not match the camera's "
                              f"{int(cam['width'])}x{int(cam['height'])}")
    return out


def downscaled_size(width, height, downscale):
    """The view size of dataset::load for an integer downscale (src/dataset.rs:172-175)."""
    w, h = C.c_uint32(), C.c_uint32()
    st = lib().splat_downscaled_size(width, height, downscale, C.byref(w), C.byref(h))
    if st != OK:
        raise SplatError(st, "downscale must be at least 1")
    return w.value, h.value


def split_eval(n, hold):
    """split_eval (src/dataset.rs:239-250) -> (train indices, held-out indices)."""
    t = np.zeros(n, np.uint8)
    st = lib().splat_split_eval(n, hold, _p(t))
    if st != OK:
        raise SplatError(st, "split_eval")
    idx = np.arange(n)
    return idx[t == 0], idx[t == 1]


def colmap_sort_images(images):
    """The image order of dataset::load (src/dataset.rs:141): by name, stable."""
    out = np.ascontiguousarray(images, rec.COLMAP_IMAGE).copy()
    st = lib().splat_colmap_sort_images(_p(out), len(out))
    if st != OK:
        raise SplatError(st, "colmap_sort_images")
    return out


def locate_sparse(root, sparse=None):
    """locate_sparse (src/dataset.rs:40-59): root/sparse/0, then root/sparse."""
    if sparse is not None:
        return sparse
    for d in (os.path.join(root, "sparse", "0"), os.path.join(root, "sparse")):
        if os.path.isdir(d):
            return d
    raise SplatError(ERR_INVALID, f"no sparse/0 or sparse directory under {root}")


def locate(root, sparse=None, images=None):
    """locate (src/dataset.rs:61-80) -> (sparse dir, images dir)."""
    s = locate_sparse(root, sparse)
    i = images if images is not None else os.path.join(root, "images")
    if not os.path.isdir(i):
        raise SplatError(ERR_INVALID, f"images directory {i} does not exist")
    return s, i


def load_dataset(trainer, sparse, images, downscale, decode):
    """dataset::load (src/dataset.rs:131-205) into a trainer's views: the COLMAP model's
    images in name order, each decoded by `decode(path) -> (h, w, 3) uint8` (the reference
    uses the `image` crate), matched to its camera and downscaled on the GPU.  Returns
    (views' cameras, points xyz float32, points rgb uint8, extent)."""
    if downscale == 0:
        raise SplatError(ERR_INVALID, "downscale must be at least 1")
    cams, imgs, pts = colmap_read(sparse)
    if len(imgs) == 0:
        raise SplatError(ERR_INVALID, f"no registered images in {sparse}")
    used = []
    for im in colmap_sort_images(imgs):
        cam = camera_from_colmap(cams, im)
        name = im["name"].decode() if isinstance(im["name"], bytes) else str(im["name"])
        used.append(trainer.add_view_downscaled(cam, decode(os.path.join(images, name)),
                                                downscale))
    xyz = np.ascontiguousarray(pts["xyz"], np.float64).astype(np.float32)
    rgb = np.ascontiguousarray(pts["rgb"], np.uint8)
    from . import scenes
    return used, xyz, rgb, scenes.scene_extent(used)


def to_rgb8(image):
    """to_rgb8, src/render.rs:26-31."""
    image = np.ascontiguousarray(image, np.float32)
    out = np.zeros(image.shape, np.uint8)
    st = lib().splat_to_rgb8(_p(image), image.size, _p(out))
    if st != OK:
        raise SplatError(st, "to_rgb8")
    return out


def psnr(pred, target):
    """psnr, src/loss.rs:175-190."""
    pred = np.ascontiguousarray(pred, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    out = np.zeros(1, np.float32)
    st = lib().splat_psnr(_p(pred), _p(target), pred.size, _p(out))
    if st != OK:
        raise SplatError(st, "psnr")
    return float(out[0])


class Trainer:
    """Device-resident training state (splat_trainer_*): model, Adam moments, statistics
    and views in HBM; `step` runs one iteration of src/train.rs:211-247 on the GPU."""

    def __init__(self, ctx, gaussians, capacity=0):
        self.ctx = ctx
        g = np.ascontiguousarray(gaussians, np.float32)
        self._h = C.c_void_p()
        ctx.check(lib().splat_trainer_create(ctx.handle, _p(g), len(g), max(capacity, len(g)),
                                             C.byref(self._h)), "trainer_create")
        ctx._trainers.add(self)
        self.views = []

    def close(self):
        if self._h:
            lib().splat_trainer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_view(self, cam, rgb):
        rgb = np.ascontiguousarray(rgb, np.uint8)
        assert rgb.shape == (int(cam["height"]), int(cam["width"]), 3)
        idx = C.c_uint32()
        self.ctx.check(lib().splat_trainer_add_view(self._h, _p(cam), _p(rgb), C.byref(idx)),
                       "trainer_add_view")
        self.views.append(cam)
        return idx.value

    def add_view_downscaled(self, cam, rgb, downscale):
        """The per-image step of dataset::load (src/dataset.rs:146-185): cam at the model's
        resolution, rgb the decoded image; returns the view's camera."""
        rgb = np.ascontiguousarray(rgb, np.uint8)
        h, w = rgb.shape[:2]
        used = np.zeros((), rec.CAMERA)
        idx = C.c_uint32()
        self.ctx.check(lib().splat_trainer_add_view_downscaled(
            self._h, _p(cam), _p(rgb), w, h, downscale, _p(used), C.byref(idx)),
            "trainer_add_view_downscaled")
        self.views.append(used)
        return used

    def view(self, index):
        """A view's camera and RGB8 pixels as stored on the device."""
        cam = np.zeros((), rec.CAMERA)
        self.ctx.check(lib().splat_trainer_view(self._h, index, _p(cam), None, 0), "trainer_view")
        rgb = np.zeros((int(cam["height"]), int(cam["width"]), 3), np.uint8)
        self.ctx.check(lib().splat_trainer_view(self._h, index, _p(cam), _p(rgb), rgb.nbytes),
                       "trainer_view")
        return cam, rgb

    def step(self, view, params, result=True):
        """One iteration; with result=False the call returns without waiting (graph mode)."""
        r = np.zeros((), rec.STEP_RESULT) if result else None
        self.ctx.check(lib().splat_trainer_step(self._h, view, _p(params),
                                                _p(r) if result else None), "trainer_step")
        return r

    def set_graph(self, enabled):
        self.ctx.check(lib().splat_trainer_set_graph(self._h, int(enabled)), "trainer_set_graph")

    def sync(self):
        self.ctx.check(lib().splat_trainer_sync(self._h), "trainer_sync")

    def tune(self, view, params):
        """Autotune the compositing launch shapes on `view` (records.TUNE_RESULT)."""
        r = np.zeros((), rec.TUNE_RESULT)
        self.ctx.check(lib().splat_trainer_tune(self._h, view, _p(params), _p(r)), "trainer_tune")
        return r

    def set_shapes(self, shapes):
        self.ctx.check(lib().splat_trainer_set_shapes(self._h, _p(shapes)), "trainer_set_shapes")

    PAIR_COUNTERS = ("fwd_visits", "fwd_active_visits", "fwd_evaluated", "fwd_passing",
                     "fwd_composited", "bwd_visits", "bwd_active_visits", "bwd_evaluated",
                     "bwd_passing")

    def pair_counts(self, view, params):
        """(Gaussian, pixel) pair and (entry, wave) visit counts of the compositing kernels on
        `view` (splat_trainer_pair_counts), as a dict keyed by PAIR_COUNTERS."""
        c = np.zeros(9, np.uint64)
        self.ctx.check(lib().splat_trainer_pair_counts(self._h, view, _p(params), _p(c)),
                       "trainer_pair_counts")
        return {k: int(v) for k, v in zip(self.PAIR_COUNTERS, c)}

    def reserve_entries(self, entries):
        self.ctx.check(lib().splat_trainer_reserve_entries(self._h, entries),
                       "trainer_reserve_entries")

    def attach_scorer(self, scorer, cfg=None):
        """Learned densification in `train` (cfg: records.learn_config); None detaches."""
        self._scorer = scorer
        self.ctx.check(lib().splat_trainer_attach_scorer(
            self._h, None if scorer is None else scorer.handle, _p(cfg)), "trainer_attach_scorer")

    def learn_stats(self):
        s = np.zeros((), rec.LEARN_STATS)
        self.ctx.check(lib().splat_trainer_learn_stats(self._h, _p(s)), "trainer_learn_stats")
        return s

    def features(self, extent, grad_threshold=0.0002, percent_dense=0.01, iteration=0):
        n, _ = self.info()
        out = np.zeros((n, 12), np.float32)
        self.ctx.check(lib().splat_trainer_features(self._h, extent, grad_threshold,
                                                    percent_dense, iteration, _p(out)),
                       "trainer_features")
        return out

    def lineage(self):
        n, _ = self.info()
        mg = np.zeros(n, np.float32)
        st, bi, og = (np.zeros(n, np.uint32) for _ in range(3))
        self.ctx.check(lib().splat_trainer_download_lineage(self._h, _p(mg), _p(st), _p(bi),
                                                            _p(og), n), "trainer_lineage")
        return dict(max_grad=mg, strong=st, birth=bi, origin=og)

    def learn_samples(self):
        st = self.learn_stats()
        k = int(st["resolved"])
        f = np.zeros((k, 12), np.float32)
        y = np.zeros(k, np.uint8)
        w = np.zeros(k, np.float32)
        cnt = C.c_uint64()
        self.ctx.check(lib().splat_trainer_learn_samples(self._h, _p(f), _p(y), _p(w), k,
                                                         C.byref(cnt)), "trainer_learn_samples")
        return f, y, w

    def graph_stats(self):
        """(captures, replays, aborted steps replayed after growing the entry arrays)."""
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.check(lib().splat_trainer_graph_stats(self._h, C.byref(a), C.byref(b),
                                                       C.byref(c)), "trainer_graph_stats")
        return a.value, b.value, c.value

    def densify(self, cfg, extent, seed):
        rep = np.zeros((), rec.DENSIFY_REPORT)
        self.ctx.check(lib().splat_trainer_densify(self._h, _p(cfg), extent, seed, _p(rep)),
                       "trainer_densify")
        return rep

    def reset_opacity(self, max_alpha=0.01):
        self.ctx.check(lib().splat_trainer_reset_opacity(self._h, max_alpha),
                       "trainer_reset_opacity")

    def render(self, view, sh_degree, antialias=False, background=(0.0, 0.0, 0.0)):
        cam = self.views[view]
        img = np.zeros((int(cam["height"]), int(cam["width"]), 3), np.float32)
        bg = np.asarray(background, np.float32)
        self.ctx.check(lib().splat_trainer_render(self._h, view, sh_degree, int(antialias),
                                                  _p(bg), _p(img)), "trainer_render")
        return img

    def train(self, extent, cfg, on_progress=None, on_save=None):
        """train, src/train.rs:161-306 on the resident model; callbacks receive a
        Progress record and (iteration, trainer)."""
        def prog(p, _user):
            if on_progress:
                on_progress(p.contents)

        def save(it, _h, _user):
            try:
                return 0 if on_save is None or on_save(it, self) is not False else 1
            except Exception:
                return 1
        pf, sf = PROGRESS_FN(prog), SAVE_FN(save)
        self.ctx.check(lib().splat_train(self._h, extent, C.byref(cfg), pf, sf, None), "train")

    def info(self):
        n = C.c_uint64()
        s = C.c_uint64()
        self.ctx.check(lib().splat_trainer_info(self._h, C.byref(n), C.byref(s)), "trainer_info")
        return n.value, s.value

    def download(self, moments=False, stats=False):
        n, _ = self.info()
        g = np.zeros((n, rec.NUM_PARAMS), np.float32)
        m = np.zeros_like(g) if moments else None
        v = np.zeros_like(g) if moments else None
        acc = np.zeros(n, np.float32) if stats else None
        den = np.zeros(n, np.uint32) if stats else None
        mr = np.zeros(n, np.uint32) if stats else None
        self.ctx.check(lib().splat_trainer_download(self._h, _p(g), _p(m), _p(v), _p(acc),
                                                    _p(den), _p(mr), n), "trainer_download")
        out = {"params": g}
        if moments:
            out.update(m=m, v=v)
        if stats:
            out.update(grad_accum=acc, denom=den, max_radii=mr)
        return out

    def evaluate(self, views, sh_degree, antialias=False, background=(0.0, 0.0, 0.0)):
        """evaluate, src/train.rs:309-321 -> (mean PSNR, per-view PSNR array)."""
        vs = np.ascontiguousarray(views, np.uint32)
        bg = np.asarray(background, np.float32)
        out = np.zeros(len(vs), np.float32)
        mean = C.c_float()
        self.ctx.check(lib().splat_trainer_evaluate(self._h, _p(vs), len(vs), sh_degree,
                                                    int(bool(antialias)), _p(bg), _p(out),
                                                    C.byref(mean)), "trainer_evaluate")
        return mean.value, out

    def restore(self, state, adam_step):
        """Replace the model and training state (splat_trainer_restore): state as download
        returns it; missing moments or statistics restore as zeros."""
        g = np.ascontiguousarray(state["params"], np.float32)
        n = len(g)

        def arr(key, dtype, shape):
            a = state.get(key)
            return None if a is None else np.ascontiguousarray(a, dtype).reshape(shape)
        m, v = arr("m", np.float32, g.shape), arr("v", np.float32, g.shape)
        acc = arr("grad_accum", np.float32, (n,))
        den, mr = arr("denom", np.uint32, (n,)), arr("max_radii", np.uint32, (n,))
        self.ctx.check(lib().splat_trainer_restore(self._h, _p(g), _p(m), _p(v), _p(acc),
                                                   _p(den), _p(mr), n, adam_step),
                       "trainer_restore")


class Scorer:
    """Learned densification scorer (splat_scorer_*): 12 -> 3x64 GELU -> 3 logits on the
    matrix cores (bfloat16 operands, binary32 accumulation)."""

    def __init__(self, ctx, params=None, seed=0):
        self.ctx = ctx
        self._h = C.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        ctx.check(lib().splat_scorer_create(ctx.handle, _p(p), seed, C.byref(self._h)),
                  "scorer_create")
        ctx._trainers.add(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().splat_scorer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(rec.SCORER_PARAMS, np.float32)
        self.ctx.check(lib().splat_scorer_get_params(self._h, _p(out)), "scorer_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.ctx.check(lib().splat_scorer_set_params(self._h, _p(p)), "scorer_set_params")

    def forward(self, feats):
        x = np.ascontiguousarray(feats, np.float32)
        out = np.zeros((len(x), 3), np.float32)
        self.ctx.check(lib().splat_scorer_forward(self._h, _p(x), len(x), _p(out)),
                       "scorer_forward")
        return out

    def loss_grad(self, feats, labels, weights):
        x = np.ascontiguousarray(feats, np.float32)
        y = np.ascontiguousarray(labels, np.uint8)
        w = np.ascontiguousarray(weights, np.float32)
        g = np.zeros(rec.SCORER_PARAMS, np.float32)
        loss = C.c_float()
        self.ctx.check(lib().splat_scorer_loss_grad(self._h, _p(x), _p(y), _p(w), len(x), _p(g),
                                                    C.byref(loss)), "scorer_loss_grad")
        return loss.value, g

    def fit(self, feats, labels, weights, cfg):
        x = np.ascontiguousarray(feats, np.float32)
        y = np.ascontiguousarray(labels, np.uint8)
        w = np.ascontiguousarray(weights, np.float32)
        loss = C.c_float()
        self.ctx.check(lib().splat_scorer_fit(self._h, _p(x), _p(y), _p(w), len(x), _p(cfg),
                                              C.byref(loss)), "scorer_fit")
        return loss.value


class Batch:
    """Batched training of many scenes and views (splat_batch_*): one launch per stage over
    every (scene, view) pair, the step captured in a HIP graph."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._h = C.c_void_p()
        ctx.check(lib().splat_batch_create(ctx.handle, C.byref(self._h)), "batch_create")
        ctx._trainers.add(self)
        self.scene_sizes = []
        self.views = []

    def close(self):
        if self._h:
            lib().splat_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, gaussians):
        g = np.ascontiguousarray(gaussians, np.float32)
        idx = C.c_uint32()
        self.ctx.check(lib().splat_batch_add_scene(self._h, _p(g), len(g), C.byref(idx)),
                       "batch_add_scene")
        self.scene_sizes.append(len(g))
        return idx.value

    def add_view(self, scene, cam, rgb):
        rgb = np.ascontiguousarray(rgb, np.uint8)
        assert rgb.shape == (int(cam["height"]), int(cam["width"]), 3)
        idx = C.c_uint32()
        self.ctx.check(lib().splat_batch_add_view(self._h, scene, _p(cam), _p(rgb),
                                                  C.byref(idx)), "batch_add_view")
        self.views.append((scene, cam))
        return idx.value

    def step(self, views, params, lr_mean, want_losses=True):
        v = np.ascontiguousarray(views, np.uint32)
        lm = np.ascontiguousarray(lr_mean, np.float32)
        losses = np.zeros(len(v), np.float32) if want_losses else None
        self.ctx.check(lib().splat_batch_step(self._h, _p(v), len(v), _p(params), _p(lm),
                                              _p(losses)), "batch_step")
        return losses

    def tune(self, views, params):
        v = np.ascontiguousarray(views, np.uint32)
        r = np.zeros((), rec.TUNE_RESULT)
        self.ctx.check(lib().splat_batch_tune(self._h, _p(v), len(v), _p(params), _p(r)),
                       "batch_tune")
        return r

    def set_shapes(self, shapes):
        self.ctx.check(lib().splat_batch_set_shapes(self._h, _p(shapes)), "batch_set_shapes")

    def download(self, scene, moments=False):
        n = self.scene_sizes[scene]
        g = np.zeros((n, rec.NUM_PARAMS), np.float32)
        m = np.zeros_like(g) if moments else None
        v = np.zeros_like(g) if moments else None
        self.ctx.check(lib().splat_batch_download(self._h, scene, _p(g), _p(m), _p(v), n),
                       "batch_download")
        return {"params": g, "m": m, "v": v} if moments else {"params": g}

    def info(self):
        vals = [C.c_uint64() for _ in range(5)]
        self.ctx.check(lib().splat_batch_info(self._h, *[C.byref(x) for x in vals]), "batch_info")
        return dict(zip(["rows", "adam_step", "captures", "replays", "aborts"],
                        [x.value for x in vals]))
