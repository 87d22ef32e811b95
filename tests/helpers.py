"""Shared test helpers: the reference test fixtures restated with numpy's generator
(tests/common/mod.rs, tests/common/scene.rs).  Seeds name numpy PCG64 streams; the
reference's ChaCha8 draws are not reproducible here."""
import math

import numpy as np

from splatrs_amd import records as rec
from splatrs_amd import scenes

# Finite-difference settings of tests/common/mod.rs:16-34.
F64 = dict(H=1e-6, RTOL=1e-5, ATOL=1e-8)
F32 = dict(H=1e-2, RTOL=5e-2, ATOL=1e-3)


def check(what, analytic, numeric, scale, tol):
    """tests/common/mod.rs:37-43: |a - n| <= ATOL * scale + RTOL * max(|a|, |n|)."""
    t = tol["ATOL"] * scale + tol["RTOL"] * max(abs(analytic), abs(numeric))
    assert abs(analytic - numeric) <= t, \
        f"{what}: analytic {analytic:.9e} vs finite difference {numeric:.9e} (tol {t:.3e})"


def central(x, h, f):
    """tests/common/mod.rs:46-48."""
    return (f(x + h) - f(x - h)) / (2.0 * h)


def rnd(rng, lo, hi):
    return lo + (hi - lo) * rng.random()


def random_gaussian(rng):
    """tests/common/mod.rs:61-92: a Gaussian near the origin with positive colours."""
    g = np.zeros(59, np.float64)
    for k in range(16):
        for c in range(3):
            g[11 + 3 * k + c] = rnd(rng, 0.2, 1.5) if k == 0 else rnd(rng, -0.15, 0.15)
    g[0:3] = [rnd(rng, -0.6, 0.6), rnd(rng, -0.4, 0.4), rnd(rng, -0.5, 0.5)]
    g[3:6] = [rnd(rng, -2.5, -1.2) for _ in range(3)]
    g[6:10] = [rnd(rng, 0.5, 1.5), rnd(rng, -0.7, 0.7), rnd(rng, -0.7, 0.7), rnd(rng, -0.7, 0.7)]
    g[10] = rnd(rng, -1.0, 2.0)
    return g


def ref_test_camera(width, height):
    """tests/common/mod.rs:95-106."""
    return scenes.look_at(width, height, width * 0.9, width * 0.9, (0.3, -0.2, -3.0),
                          (0.0, 0.0, 0.0))


def ring_cameras(n, width, height, radius):
    """tests/common/scene.rs:32-53."""
    cams = []
    for i in range(n):
        a = np.float32(np.float32(i) / np.float32(n) * np.float32(2 * math.pi))
        eye = (radius * math.cos(a), -0.6 + 0.3 * math.sin(3.0 * a), radius * math.sin(a))
        cams.append(scenes.look_at(width, height, width * 1.1, width * 1.1, eye,
                                   (0.0, 0.0, 0.0)))
    return cams


def ground_truth(rng, n):
    """tests/common/scene.rs:56-94: random blobs, SH degree 1."""
    sh_c0 = 0.28209479177387814
    out = np.zeros((n, 59), np.float32)
    for i in range(n):
        g = np.zeros(59)
        for c in range(3):
            g[11 + c] = (rnd(rng, 0.1, 0.9) - 0.5) / sh_c0
        for k in range(1, 4):
            for c in range(3):
                g[11 + 3 * k + c] = rnd(rng, -0.2, 0.2)
        g[0:3] = [rnd(rng, -0.7, 0.7), rnd(rng, -0.5, 0.5), rnd(rng, -0.7, 0.7)]
        g[3:6] = [rnd(rng, -2.6, -1.6) for _ in range(3)]
        g[6:10] = [rnd(rng, 0.5, 1.0), rnd(rng, -0.5, 0.5), rnd(rng, -0.5, 0.5),
                   rnd(rng, -0.5, 0.5)]
        g[10] = rnd(rng, 0.5, 3.0)
        out[i] = g
    return out


def init_points(gt, rng):
    """tests/common/scene.rs:111-127: the true means jittered by U(-0.03, 0.03), with the
    DC colour as RGB8."""
    sh_c0 = np.float32(0.28209479177387814)
    xyz = np.zeros((len(gt), 3), np.float32)
    rgb = np.zeros((len(gt), 3), np.uint8)
    for i, g in enumerate(gt):
        xyz[i] = [np.float32(g[k]) + np.float32(rnd(rng, -0.03, 0.03)) for k in range(3)]
        for c in range(3):
            v = np.clip(np.float32(g[11 + c]) * sh_c0 + np.float32(0.5), 0, 1) * np.float32(255)
            rgb[i, c] = int(np.floor(v + 0.5))
    return xyz, rgb


def naive_render(projs, width, height, bg):
    """tests/pipeline.rs:134-186: per-pixel compositor over every visible Gaussian whose
    tile rectangle covers the pixel's tile, in (depth bits, index) order."""
    vis = [i for i in range(len(projs)) if projs[i]["radius"] > 0]
    depth_bits = projs["depth"].astype(np.float32).view(np.uint32)
    order = sorted(vis, key=lambda i: (int(depth_bits[i]), i))
    f = np.float32
    image = np.zeros((height, width, 3), np.float32)
    P = [(projs[i]["tile_min"], projs[i]["tile_max"], projs[i]["mean2d"].astype(np.float32),
          projs[i]["conic"].astype(np.float32), f(projs[i]["opacity"]),
          projs[i]["color"].astype(np.float32)) for i in order]
    for y in range(height):
        for x in range(width):
            tx, ty = x // 16, y // 16
            px, py = f(x) + f(0.5), f(y) + f(0.5)
            t = f(1.0)
            c = [f(0.0)] * 3
            for tmin, tmax, m, cn, o, col in P:
                if tx < tmin[0] or tx >= tmax[0] or ty < tmin[1] or ty >= tmax[1]:
                    continue
                dx = f(px - m[0])
                dy = f(py - m[1])
                sigma = f(f(f(0.5) * f(f(f(cn[0] * dx) * dx) + f(f(cn[2] * dy) * dy))) +
                          f(f(cn[1] * dx) * dy))
                if sigma < 0:
                    continue
                og = f(o * np.exp(f(-sigma), dtype=np.float32))
                if og < f(f(1.0) / f(255.0)):
                    continue
                alpha = min(og, f(0.99))
                test_t = f(t * f(f(1.0) - alpha))
                if test_t < f(1e-4):
                    break
                for ch in range(3):
                    c[ch] = f(c[ch] + f(f(alpha * t) * col[ch]))
                t = test_t
            for ch in range(3):
                image[y, x, ch] = f(c[ch] + f(t * f(bg[ch])))
    return image


ERR_QUANTILES = (0.5, 0.9, 0.99, 0.999)
# A quantile is compared only when at least this many samples lie beyond it: the p99.9 of
# 120 values is a single sample, whose ranking between two summation orders is chance.
TAIL_SAMPLES = 10


def error_quantiles(x, truth):
    """Quantiles of |x - truth| / |truth| over the entries with a non-zero truth; a quantile
    with fewer than TAIL_SAMPLES samples beyond it is left out."""
    x = np.asarray(x, np.float64).ravel()
    t = np.asarray(truth, np.float64).ravel()
    nz = t != 0
    n = int(nz.sum())
    qs = [q for q in ERR_QUANTILES if n * (1.0 - q) >= TAIL_SAMPLES]
    if not qs:
        return {}
    rel = np.abs(x[nz] - t[nz]) / np.abs(t[nz])
    return {q: float(np.quantile(rel, q)) for q in qs}


# Two binary32 summation orders of the same terms have error distributions that differ
# only by sampling noise, and where a gradient's error is dominated by the chain both
# share (the projection backward), their quantiles agree to a few parts in 10^4.  The 5%
# allowance absorbs that noise; a reduction mistake (a dropped, doubled or misrouted
# partial) moves the upper quantiles by orders of magnitude.
ACCURACY_SLACK = 1.05


def no_less_accurate(name, fast, ref, truth, say=print, slack=ACCURACY_SLACK):
    """Asserts that `fast` (a reassociated binary32 computation) is no less accurate than
    `ref` (the reference's binary32 order) against the binary64 `truth`: at the median,
    p99 and p99.9 of the relative errors, fast's error is at most `slack` times ref's.
    Returns both quantile tables."""
    qf, qr = error_quantiles(fast, truth), error_quantiles(ref, truth)
    say(f"{name}: relative error vs binary64 truth, fast / reference order: " +
        ", ".join(f"p{100 * q:g} {qf[q]:.2e} / {qr[q]:.2e}" for q in qf))
    for q in qf:
        assert qf[q] <= slack * qr[q] + 1e-12, (name, q, qf[q], qr[q])
    return qf, qr


class DeviceArrays:
    """Device buffers for tests of entry points that take device pointers.  The HIP calls
    are looked up through libsplat's own handle (dlsym searches its dependencies), so
    buffers and kernels share the runtime libsplat is bound to, whichever copy of it the
    process loaded first.  upload() returns a device address (int)."""

    def __init__(self):
        import ctypes
        from splatrs_amd import abi
        self.C = ctypes
        self.hip = abi.lib()
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []
