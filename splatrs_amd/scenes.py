"""Cameras and synthetic scenes.

`look_at` and `scene_extent` restate src/camera.rs:66-118 in binary32 (numpy float32
scalars, left-to-right sums, no fused operations) so a camera built here is bit-identical
to the reference's.  The scene generators draw from numpy's PCG64 (`default_rng(seed)`):
the reference's ChaCha8 streams are not reproducible here, so seeds name this generator.
"""
import math

import numpy as np

from .records import CAMERA, NUM_PARAMS

f32 = np.float32


def _v(x, y, z):
    return [f32(x), f32(y), f32(z)]


def _dot(a, b):
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def _norm(a):
    return f32(np.sqrt(_dot(a, a)))


def _normalized(a):  # math.rs:70-77
    n = _norm(a)
    if n > 0:
        inv = f32(f32(1.0) / n)
        return [f32(a[0] * inv), f32(a[1] * inv), f32(a[2] * inv)]
    return list(a)


def _cross(a, b):  # math.rs:51-57
    return [f32(f32(a[1] * b[2]) - f32(a[2] * b[1])),
            f32(f32(a[2] * b[0]) - f32(a[0] * b[2])),
            f32(f32(a[0] * b[1]) - f32(a[1] * b[0]))]


def look_at(width, height, fx, fy, eye, target, up=(0.0, -1.0, 0.0)):
    """Camera::look_at, src/camera.rs:66-94 (cx = width / 2, cy = height / 2)."""
    eye = _v(*eye)
    target = _v(*target)
    up = _v(*up)
    forward = _normalized([f32(target[i] - eye[i]) for i in range(3)])
    right = _normalized(_cross(forward, up))
    down = _cross(forward, right)
    rot = [right, down, forward]
    rv = [_dot(rot[r], eye) for r in range(3)]
    trans = [f32(-rv[0]), f32(-rv[1]), f32(-rv[2])]
    cam = np.zeros((), CAMERA)
    cam["width"] = width
    cam["height"] = height
    cam["fx"] = f32(fx)
    cam["fy"] = f32(fy)
    cam["cx"] = f32(f32(width) / f32(2.0))
    cam["cy"] = f32(f32(height) / f32(2.0))
    cam["rot"] = np.array([rot[r][c] for r in range(3) for c in range(3)], np.float32)
    cam["trans"] = np.array(trans, np.float32)
    return cam


def camera_center(cam):
    """Camera::center, src/camera.rs:41-43: -(rot^T trans)."""
    r = np.asarray(cam["rot"], np.float32).reshape(3, 3)
    t = np.asarray(cam["trans"], np.float32)
    out = []
    for c in range(3):
        col = [r[0, c], r[1, c], r[2, c]]
        out.append(f32(-_dot(col, t)))
    return out


def scene_extent(cams):
    """scene_extent, src/camera.rs:99-118."""
    if len(cams) == 0:
        return f32(1.0)
    center = [f32(0.0)] * 3
    for c in cams:
        cc = camera_center(c)
        center = [f32(center[i] + cc[i]) for i in range(3)]
        # V3 += V3
    inv = f32(f32(1.0) / f32(len(cams)))
    center = [f32(center[i] * inv) for i in range(3)]
    radius = f32(0.0)
    for c in cams:
        cc = camera_center(c)
        d = _norm([f32(cc[i] - center[i]) for i in range(3)])
        radius = f32(max(radius, d))
    ext = f32(radius * f32(1.1))
    return ext if ext > 0 else f32(1.0)


def camera_array(cams):
    out = np.zeros(len(cams), CAMERA)
    for i, c in enumerate(cams):
        out[i] = c
    return out


def random_unit_quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def random_gaussians(rng, n, mean_lo=-1.0, mean_hi=1.0, log_scale_mu=-4.0,
                     log_scale_sigma=0.5, dc_mu=0.5, dc_sigma=0.2, rest_sigma=0.05):
    """Gaussian records (n x 59 float32) drawn as the synthetic configurations specify:
    means uniform in [lo, hi)^3, log-scales N(mu, sigma), random unit quaternions,
    opacity logits N(0, 1), SH DC coefficients N(dc_mu, dc_sigma), higher bands
    N(0, rest_sigma)."""
    g = np.zeros((n, NUM_PARAMS), np.float64)
    g[:, 0:3] = rng.uniform(mean_lo, mean_hi, (n, 3))
    g[:, 3:6] = rng.normal(log_scale_mu, log_scale_sigma, (n, 3))
    g[:, 6:10] = random_unit_quats(rng, n)
    g[:, 10] = rng.standard_normal(n)
    g[:, 11:14] = rng.normal(dc_mu, dc_sigma, (n, 3))
    g[:, 14:] = rng.normal(0.0, rest_sigma, (n, 45))
    return g.astype(np.float32)


def fibonacci_sphere(n):
    """Deterministic, roughly uniform unit directions."""
    out = []
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        y = 1.0 - 2.0 * (i + 0.5) / n
        r = math.sqrt(max(0.0, 1.0 - y * y))
        th = golden * i
        out.append((r * math.cos(th), y, r * math.sin(th)))
    return out


def small_scene_cameras(n=8, size=256, f=300.0, radius=4.0):
    """Eight pinhole cameras on a radius-4 sphere looking at the origin, 256x256,
    fx = fy = 300 (the small synthetic CPU-runnable configuration)."""
    cams = []
    for d in fibonacci_sphere(n):
        eye = (radius * d[0], radius * d[1], radius * d[2])
        up = (0.0, -1.0, 0.0) if abs(d[1]) < 0.99 else (0.0, 0.0, 1.0)
        cams.append(look_at(size, size, f, f, eye, (0.0, 0.0, 0.0), up))
    return cams


def small_scene(seed=0, n=10_000):
    """10k Gaussians: means U[-1,1]^3, log-scales N(-4,0.5), SH degree 3."""
    return random_gaussians(np.random.default_rng(seed), n)


def large_scene(seed=0, n=1_000_000):
    """1M Gaussians: means U[-5,5]^3, log-scales N(-5,0.7), SH degree 3."""
    return random_gaussians(np.random.default_rng(seed), n, -5.0, 5.0, -5.0, 0.7)


def hd_camera():
    """One 1920x1080 camera, fx = fy = 1600, at distance 8 looking at the origin."""
    return look_at(1920, 1080, 1600.0, 1600.0, (0.0, 0.0, -8.0), (0.0, 0.0, 0.0))


# ---- ellipsoid-surface scene (full training with densification) ------------------------

SH_C0 = 0.28209479177387814


def _quat_mats(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def ellipsoids(rng, k=20):
    """k random ellipsoids: centres U[-2.2, 2.2]^3, semi-axes U[0.25, 1.0], random
    orientations, base colours U[0.1, 0.9]^3."""
    return dict(center=rng.uniform(-2.2, 2.2, (k, 3)), radii=rng.uniform(0.25, 1.0, (k, 3)),
                rot=_quat_mats(random_unit_quats(rng, k)), color=rng.uniform(0.1, 0.9, (k, 3)))


def sample_surface(rng, ell, n):
    """n points on the ellipsoid surfaces (ellipsoid chosen in proportion to its area, by
    Thomsen's approximation), with their ellipsoid index."""
    a, b, c = ell["radii"].T
    p = 1.6075
    area = 4 * np.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)
    idx = rng.choice(len(a), n, p=area / area.sum())
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    local = d * ell["radii"][idx]
    pts = np.einsum("nij,nj->ni", ell["rot"][idx], local) + ell["center"][idx]
    return pts, idx


def surface_color(ell, idx, pts):
    """Each ellipsoid's base colour with a smooth modulation and a fine 3D checker
    (cells of 1/32 unit, a few pixels on screen at distance 6), so the surfaces carry
    detail the initial point cloud cannot represent."""
    smooth = 0.12 * np.sin(4.0 * pts[:, 0] + 3.0 * pts[:, 1]) * np.cos(5.0 * pts[:, 2])
    cells = np.floor(pts * 32.0).astype(np.int64)
    checker = np.where((cells.sum(axis=1) & 1) == 0, 0.3, -0.3)
    rot = (idx % 3)[:, None] == np.arange(3)[None, :]   # the checker tints one channel
    col = ell["color"][idx] + smooth[:, None] + checker[:, None] * (0.5 + 0.5 * rot)
    return np.clip(col, 0.02, 0.98)


def ellipsoid_scene(seed=0, n_gt=2_000_000, n_points=300_000, k=20):
    """Ground-truth Gaussians on k ellipsoid surfaces (log-scales N(-4.6, 0.25), opacity
    logits N(2, 0.5), coloured DC, higher bands N(0, 0.02)) and an initial point cloud of
    n_points surface samples with per-point RGB8 colour."""
    rng = np.random.default_rng(seed)
    ell = ellipsoids(rng, k)
    pos, idx = sample_surface(rng, ell, n_gt)
    g = np.zeros((n_gt, NUM_PARAMS), np.float32)
    g[:, 0:3] = pos
    g[:, 3:6] = rng.normal(-4.6, 0.25, (n_gt, 3))
    g[:, 6:10] = random_unit_quats(rng, n_gt)
    g[:, 10] = rng.normal(2.0, 0.5, n_gt)
    g[:, 11:14] = (surface_color(ell, idx, pos) - 0.5) / SH_C0
    g[:, 14:] = rng.normal(0.0, 0.02, (n_gt, 45))
    pts, pidx = sample_surface(np.random.default_rng(seed + 1), ell, n_points)
    rgb = np.floor(surface_color(ell, pidx, pts) * 255 + 0.5).astype(np.uint8)
    return g, pts.astype(np.float32), rgb


def hemisphere_cameras(n=100, width=1297, height=840, f=1100.0, radius=6.0):
    """n cameras on the upper hemisphere (world up is -y) at `radius`, looking at the
    origin."""
    cams = []
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        y = -(0.05 + 0.9 * (i + 0.5) / n)
        r = math.sqrt(max(0.0, 1.0 - y * y))
        th = golden * i
        eye = (radius * r * math.cos(th), radius * y, radius * r * math.sin(th))
        cams.append(look_at(width, height, f, f, eye, (0.0, 0.0, 0.0)))
    return cams


# ---- batched configuration: many scenes, many views of uneven resolution ------------------

RESOLUTIONS = [(320, 240), (480, 360), (640, 480), (800, 600), (1024, 768), (1280, 960),
               (1600, 1200)]


def batch_scene_sizes(seed=0, scenes=16, lo=50_000, hi=800_000):
    """Scene sizes log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(scenes)]


def batch_views(seed=0, scenes=16, per_scene=4, radius=4.0):
    """per_scene cameras per scene on a radius-4 sphere looking at the origin, resolutions
    drawn from 320x240 .. 1600x1200 (fx = fy = 1.1 width).  Returns [(scene, camera)]."""
    rng = np.random.default_rng(seed + 1000)
    dirs = fibonacci_sphere(per_scene * scenes)
    out = []
    for s in range(scenes):
        for k in range(per_scene):
            d = dirs[(k * scenes + s) % len(dirs)]
            w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
            eye = (radius * d[0], radius * d[1], radius * d[2])
            up = (0.0, -1.0, 0.0) if abs(d[1]) < 0.99 else (0.0, 0.0, 1.0)
            out.append((s, look_at(w, h, 1.1 * w, 1.1 * w, eye, (0.0, 0.0, 0.0), up)))
    return out
