"""numpy layouts of the plain records in include/splat.h.

The records keep the field order of the reference's structs (src/project.rs:47-93,
src/camera.rs:13-32), so a numpy structured array of these dtypes can be handed to the
C ABI as a pointer.
"""
import numpy as np

NUM_PARAMS = 59  # tests/common/mod.rs:109
TILE = 16        # src/project.rs:18
SLOTS = 11       # src/rasterize.rs:341

CAMERA = np.dtype([
    ("width", "<u4"), ("height", "<u4"),
    ("fx", "<f4"), ("fy", "<f4"), ("cx", "<f4"), ("cy", "<f4"),
    ("rot", "<f4", (9,)), ("trans", "<f4", (3,)),
])
PROJ_OPTS = np.dtype([("near_plane", "<f4"), ("eps2d", "<f4"), ("antialias", "<i4"),
                      ("sh_degree", "<i4")])
PROJ = np.dtype([
    ("mean2d", "<f4", (2,)), ("conic", "<f4", (3,)), ("depth", "<f4"), ("radius", "<u4"),
    ("opacity", "<f4"), ("color", "<f4", (3,)), ("tile_min", "<u4", (2,)),
    ("tile_max", "<u4", (2,)),
])
PROJ_GRAD = np.dtype([("mean2d", "<f4", (2,)), ("conic", "<f4", (3,)), ("opacity", "<f4"),
                      ("color", "<f4", (3,))])
RASTER_SETTINGS = np.dtype([("width", "<u4"), ("height", "<u4"), ("background", "<f4", (3,))])
LEARNING_RATES = np.dtype([(k, "<f4") for k in
                           ("mean", "log_scale", "quat", "opacity", "sh_dc", "sh_rest")])
ADAM_STATE = np.dtype([("beta1", "<f4"), ("beta2", "<f4"), ("eps", "<f4"), ("pad_", "<u4"),
                       ("step", "<u8")])
DENSIFY_CONFIG = np.dtype([
    ("grad_threshold", "<f4"), ("percent_dense", "<f4"), ("min_opacity", "<f4"),
    ("prune_big_world", "<i4"), ("max_world_size_frac", "<f4"),
    ("has_max_screen_size", "<i4"), ("max_screen_size", "<u4"), ("split_children", "<u4"),
])
DENSIFY_REPORT = np.dtype([("cloned", "<u8"), ("split", "<u8"), ("pruned", "<u8")])
SCORER_FIT_CONFIG = np.dtype([("epochs", "<u4"), ("batch", "<u4"), ("lr", "<f4"), ("pad_", "<u4"),
                              ("seed", "<u8")])
LEARN_CONFIG = np.dtype([("mode", "<i4"), ("samples_per_event", "<u4"), ("max_events", "<u4"),
                         ("outcome_delay", "<u4"), ("threshold", "<f4"),
                         ("prune_threshold", "<f4"), ("eval_views", "<u4"),
                         ("useful_opacity", "<f4"), ("useful_ratio", "<f4"),
                         ("credit_views", "<u4"), ("credit_ratio", "<f4"),
                         ("share_prune", "<f4"), ("share_prune_from", "<u4"), ("pad_", "<u4"),
                         ("fit", SCORER_FIT_CONFIG)])
LEARN_STATS = np.dtype([("logged", "<u8"), ("resolved", "<u8"), ("updates", "<u8"),
                        ("positives", "<u8"), ("events", "<u8"), ("last_loss", "<f4"),
                        ("pad_", "<u4")])
BATCH_STEP_PARAMS = np.dtype([("sh_degree", "<i4"), ("ordered_backward", "<i4"),
                              ("antialias", "<i4"), ("near_plane", "<f4"),
                              ("lambda_dssim", "<f4"), ("background", "<f4", (3,)),
                              ("lr", LEARNING_RATES)])
TUNE_RESULT = np.dtype([("fwd_batch", "<i4"), ("fwd_persistent", "<i4"), ("fwd_per_cu", "<i4"),
                        ("bwd_batch", "<i4"), ("bwd_persistent", "<i4"), ("bwd_per_cu", "<i4"),
                        ("fwd_ms", "<f4"), ("bwd_ms", "<f4"), ("default_fwd_ms", "<f4"),
                        ("default_bwd_ms", "<f4"), ("candidates", "<u4"),
                        ("bwd_row_lists", "<i4")])
SCORER_PARAMS = 9347
SCORER_FEATURES = 12

STEP_PARAMS = np.dtype([
    ("sh_degree", "<i4"), ("accumulate_stats", "<i4"), ("absgrad", "<i4"),
    ("ordered_backward", "<i4"), ("want_loss", "<i4"), ("antialias", "<i4"),
    ("near_plane", "<f4"), ("lambda_dssim", "<f4"), ("background", "<f4", (3,)),
    ("lr", LEARNING_RATES),
])
STEP_RESULT = np.dtype([("loss", "<f4"), ("pad_", "<u4"), ("num_entries", "<u8"),
                        ("num_visible", "<u8")])

COLMAP_CAMERA = np.dtype([("id", "<u4"), ("model", "<i4"), ("width", "<u8"), ("height", "<u8"),
                          ("num_params", "<u4"), ("pad", "<u4"), ("params", "<f8", (16,))])
COLMAP_IMAGE = np.dtype([("id", "<u4"), ("camera_id", "<u4"), ("qvec", "<f8", (4,)),
                         ("tvec", "<f8", (3,)), ("name", "S256")])
COLMAP_POINT = np.dtype([("id", "<u8"), ("xyz", "<f8", (3,)), ("rgb", "u1", (3,)),
                         ("pad", "u1", (5,)), ("error", "<f8")])
# COLMAP camera models in id order with their parameter counts (src/colmap/mod.rs:17-135)
COLMAP_MODELS = [("SIMPLE_PINHOLE", 3), ("PINHOLE", 4), ("SIMPLE_RADIAL", 4), ("RADIAL", 5),
                 ("OPENCV", 8), ("OPENCV_FISHEYE", 8), ("FULL_OPENCV", 12), ("FOV", 5),
                 ("SIMPLE_RADIAL_FISHEYE", 4), ("RADIAL_FISHEYE", 5),
                 ("THIN_PRISM_FISHEYE", 12), ("RAD_TAN_THIN_PRISM_FISHEYE", 16)]
assert COLMAP_CAMERA.itemsize == 160 and COLMAP_IMAGE.itemsize == 320
assert COLMAP_POINT.itemsize == 48
assert CAMERA.itemsize == 72 and PROJ.itemsize == 60 and PROJ_GRAD.itemsize == 36
assert STEP_PARAMS.itemsize == 68 and STEP_RESULT.itemsize == 24
assert ADAM_STATE.itemsize == 24 and DENSIFY_CONFIG.itemsize == 32


def proj_opts(sh_degree=3, antialias=False, near_plane=0.2, eps2d=0.3):
    """ProjOpts (src/project.rs:34-43 defaults)."""
    o = np.zeros((), PROJ_OPTS)
    o["near_plane"] = near_plane
    o["eps2d"] = eps2d
    o["antialias"] = int(bool(antialias))
    o["sh_degree"] = sh_degree
    return o


def raster_settings(width, height, background=(0.0, 0.0, 0.0)):
    s = np.zeros((), RASTER_SETTINGS)
    s["width"] = width
    s["height"] = height
    s["background"] = np.asarray(background, np.float32)
    return s


def learning_rates(mean, log_scale=0.005, quat=0.001, opacity=0.025, sh_dc=0.0025,
                   sh_rest=None):
    """LearningRates with the reference's TrainConfig defaults (src/train.rs:84-90)."""
    lr = np.zeros((), LEARNING_RATES)
    lr["mean"] = mean
    lr["log_scale"] = log_scale
    lr["quat"] = quat
    lr["opacity"] = opacity
    lr["sh_dc"] = sh_dc
    lr["sh_rest"] = np.float32(sh_dc) / np.float32(20.0) if sh_rest is None else sh_rest
    return lr


def step_params(sh_degree, lr, accumulate_stats=False, absgrad=False, ordered_backward=False,
                want_loss=False, antialias=False, near_plane=0.2, lambda_dssim=0.2,
                background=(0.0, 0.0, 0.0)):
    p = np.zeros((), STEP_PARAMS)
    p["sh_degree"] = sh_degree
    p["accumulate_stats"] = int(bool(accumulate_stats))
    p["absgrad"] = int(bool(absgrad))
    p["ordered_backward"] = int(bool(ordered_backward))
    p["want_loss"] = int(bool(want_loss))
    p["antialias"] = int(bool(antialias))
    p["near_plane"] = near_plane
    p["lambda_dssim"] = lambda_dssim
    p["background"] = np.asarray(background, np.float32)
    p["lr"] = lr
    return p


def position_lr(iteration, extent, init=0.00016, final=0.0000016, max_steps=30000):
    """TrainConfig::position_lr, src/train.rs:132-137, in binary32."""
    f = np.float32
    t = f(f(iteration) / f(max(max_steps, 1)))
    t = min(max(t, f(0.0)), f(1.0))
    a = f(np.log(f(f(init) * f(extent))))
    b = f(np.log(f(f(final) * f(extent))))
    return f(np.exp(f(f(a * f(f(1.0) - t)) + f(b * t))))


def adam_state(step=0, beta1=0.9, beta2=0.999, eps=1e-15):
    """Adam state (src/adam.rs:65-74 defaults)."""
    s = np.zeros((), ADAM_STATE)
    s["beta1"] = beta1
    s["beta2"] = beta2
    s["eps"] = eps
    s["step"] = step
    return s


def densify_config(grad_threshold=0.0002, percent_dense=0.01, min_opacity=0.005,
                   prune_big_world=False, max_world_size_frac=0.1, max_screen_size=None,
                   split_children=2):
    """DensifyConfig (src/densify.rs:91-103 defaults)."""
    c = np.zeros((), DENSIFY_CONFIG)
    c["grad_threshold"] = grad_threshold
    c["percent_dense"] = percent_dense
    c["min_opacity"] = min_opacity
    c["prune_big_world"] = int(bool(prune_big_world))
    c["max_world_size_frac"] = max_world_size_frac
    c["has_max_screen_size"] = 0 if max_screen_size is None else 1
    c["max_screen_size"] = 0 if max_screen_size is None else max_screen_size
    c["split_children"] = split_children
    return c


def scorer_fit_config(epochs=4, batch=4096, lr=1e-3, seed=0):
    c = np.zeros((), SCORER_FIT_CONFIG)
    c["epochs"] = epochs
    c["batch"] = batch
    c["lr"] = lr
    c["seed"] = seed
    return c


def learn_config(mode=1, samples_per_event=16384, max_events=64, outcome_delay=500,
                 threshold=0.5, prune_threshold=0.5, eval_views=8, useful_opacity=0.05,
                 useful_ratio=0.5, credit_views=0, credit_ratio=1.0, share_prune=0.0,
                 share_prune_from=0, fit=None):
    c = np.zeros((), LEARN_CONFIG)
    c["mode"] = mode
    c["samples_per_event"] = samples_per_event
    c["max_events"] = max_events
    c["outcome_delay"] = outcome_delay
    c["threshold"] = threshold
    c["prune_threshold"] = prune_threshold
    c["eval_views"] = eval_views
    c["useful_opacity"] = useful_opacity
    c["useful_ratio"] = useful_ratio
    c["credit_views"] = credit_views
    c["credit_ratio"] = credit_ratio
    c["share_prune"] = share_prune
    c["share_prune_from"] = share_prune_from
    c["fit"] = fit if fit is not None else scorer_fit_config()
    return c


def batch_step_params(sh_degree, lr, ordered_backward=False, antialias=False, near_plane=0.2,
                      lambda_dssim=0.2, background=(0.0, 0.0, 0.0)):
    p = np.zeros((), BATCH_STEP_PARAMS)
    p["sh_degree"] = sh_degree
    p["ordered_backward"] = int(bool(ordered_backward))
    p["antialias"] = int(bool(antialias))
    p["near_plane"] = near_plane
    p["lambda_dssim"] = lambda_dssim
    p["background"] = background
    p["lr"] = lr
    return p
