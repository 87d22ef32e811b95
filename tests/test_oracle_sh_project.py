"""Pins of the oracle's SH basis and projection (src/sh.rs, src/project.rs).

- basis: orthonormality on the sphere by exact Gauss-Legendre x trapezoid quadrature (a
  wrong constant, sign pattern or cross term breaks it) and closed-form values;
- basis_grad: central differences (port of tests/gradcheck.rs:21-47) in f64 and f32;
- projection forward: closed form for an isotropic Gaussian on the optical axis, and an
  independent numpy f64 EWA computation;
- projection backward: central differences over all 59 parameters (port of
  tests/gradcheck.rs:79-182: plain, antialiased, lower degrees, clamped Jacobian).
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import F32, F64, central, check, random_gaussian, rnd, ref_test_camera
from splatrs_amd import records as rec
from splatrs_amd import scenes


def basis64(degree, d):
    out = np.zeros(16)
    O.lib().oracle_sh_basis_f64(degree, O.ptr(np.asarray(d, np.float64)), O.ptr(out))
    return out


def basis_grad64(degree, d):
    out = np.zeros(48)
    O.lib().oracle_sh_basis_grad_f64(degree, O.ptr(np.asarray(d, np.float64)), O.ptr(out))
    return out.reshape(16, 3)


def basis32(degree, d):
    out = np.zeros(16, np.float32)
    O.lib().oracle_sh_basis_f32(degree, O.ptr(np.asarray(d, np.float32)), O.ptr(out))
    return out


def basis_grad32(degree, d):
    out = np.zeros(48, np.float32)
    O.lib().oracle_sh_basis_grad_f32(degree, O.ptr(np.asarray(d, np.float32)), O.ptr(out))
    return out.reshape(16, 3)


def test_sh_basis_orthonormal():
    # Products of degree <= 3 basis functions are polynomials of degree <= 6 on the
    # sphere: 8 Gauss-Legendre nodes in z and 16 equispaced azimuths integrate them exactly.
    zs, wz = np.polynomial.legendre.leggauss(8)
    phis = np.arange(16) * (2 * np.pi / 16)
    gram = np.zeros((16, 16))
    for z, w in zip(zs, wz):
        r = np.sqrt(1 - z * z)
        for ph in phis:
            b = basis64(3, [r * np.cos(ph), r * np.sin(ph), z])
            gram += w * (2 * np.pi / 16) * np.outer(b, b)
    assert np.abs(gram - np.eye(16)).max() < 1e-12


def test_sh_basis_closed_form():
    # Real spherical harmonics with the 3DGS sign convention, written independently.
    rng = np.random.default_rng(3)
    for _ in range(10):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        x, y, z = d
        pi = np.pi
        want = [
            0.5 / np.sqrt(pi),
            -np.sqrt(3 / (4 * pi)) * y, np.sqrt(3 / (4 * pi)) * z, -np.sqrt(3 / (4 * pi)) * x,
            0.5 * np.sqrt(15 / pi) * x * y, -0.5 * np.sqrt(15 / pi) * y * z,
            0.25 * np.sqrt(5 / pi) * (3 * z * z - 1), -0.5 * np.sqrt(15 / pi) * x * z,
            0.25 * np.sqrt(15 / pi) * (x * x - y * y),
            -0.25 * np.sqrt(35 / (2 * pi)) * y * (3 * x * x - y * y),
            0.5 * np.sqrt(105 / pi) * x * y * z,
            -0.25 * np.sqrt(21 / (2 * pi)) * y * (5 * z * z - 1),
            0.25 * np.sqrt(7 / pi) * z * (5 * z * z - 3),
            -0.25 * np.sqrt(21 / (2 * pi)) * x * (5 * z * z - 1),
            0.25 * np.sqrt(105 / pi) * z * (x * x - y * y),
            -0.25 * np.sqrt(35 / (2 * pi)) * x * (x * x - 3 * y * y),
        ]
        assert np.allclose(basis64(3, d), want, rtol=1e-12, atol=1e-14)
        # lower degrees leave the higher entries at zero
        b1 = basis64(1, d)
        assert np.all(b1[4:] == 0) and np.allclose(b1[:4], want[:4])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_sh_basis_gradient(prec):
    rng = np.random.default_rng(7)
    tol = F64 if prec == "f64" else F32
    bf, gf = (basis64, basis_grad64) if prec == "f64" else (basis32, basis_grad32)
    dt = np.float64 if prec == "f64" else np.float32
    for _ in range(20):
        d = np.array([rnd(rng, -1, 1) for _ in range(3)])
        d = (d / np.linalg.norm(d)).astype(dt)
        g = gf(3, d)
        for k in range(16):
            for axis in range(3):
                def f(e):
                    p = d.copy()
                    p[axis] = dt(p[axis] + dt(e))
                    return float(bf(3, p)[k])
                num = central(0.0, tol["H"], f)
                check(f"basis[{k}] d/d{axis}", float(g[k, axis]), num, 1.0, tol)


# ---- projection -------------------------------------------------------------------------

def proj15_fields(p):
    """(mean2d 2, conic 3, opacity, color 3) of a 15-double projection."""
    return np.concatenate([p[0:2], p[2:5], [p[7]], p[8:11]])


def test_project_isotropic_closed_form():
    cam = scenes.look_at(640, 480, 500.0, 500.0, (0.0, 0.0, -4.0), (0.0, 0.0, 0.0))
    c64 = O.cam64(cam)
    for s in [0.01, 0.05, 0.2]:
        g = np.zeros((1, 59))
        g[0, 3:6] = np.log(s)
        g[0, 6] = 1.0
        g[0, 10] = 0.7
        g[0, 11:14] = [0.4, -0.2, 1.0]
        p = O.project_all_f64(g, c64, O.opts64(0))[0]
        v = (500.0 * s / 4.0) ** 2 + 0.3
        assert np.allclose(p[0:2], [320.0, 240.0], rtol=0, atol=1e-12)
        assert np.allclose(p[2:5], [1 / v, 0.0, 1 / v], rtol=1e-12, atol=1e-15)
        assert p[5] == 4.0
        assert p[6] == np.ceil(3 * np.sqrt(v + np.sqrt(0.1)))
        assert np.isclose(p[7], 1 / (1 + np.exp(-0.7)), rtol=1e-14)
        sh_c0 = 0.28209479177387814
        assert np.allclose(p[8:11], np.maximum(0.5 + sh_c0 * np.array([0.4, -0.2, 1.0]), 0),
                           rtol=1e-14)
        # antialias: comp = sqrt(det0 / det)
        pa = O.project_all_f64(g, c64, O.opts64(0, antialias=True))[0]
        v0 = (500.0 * s / 4.0) ** 2
        assert np.isclose(pa[7], p[7] * np.sqrt(v0 * v0 / (v * v)), rtol=1e-12)


def ewa_reference(g, cam, eps2d=0.3):
    """Independent numpy f64 EWA: Sigma' = J W Sigma W^T J^T + eps I, J at the clamped
    point of the enlarged frustum."""
    W = np.asarray(cam["rot"], np.float64).reshape(3, 3)
    t = W @ g[0:3] + np.asarray(cam["trans"], np.float64)
    q = g[6:10] / np.linalg.norm(g[6:10])
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    S = np.diag(np.exp(g[3:6]))
    Sig = R @ S @ S @ R.T
    fx, fy, cx, cy = (float(cam[k]) for k in ("fx", "fy", "cx", "cy"))
    Wd, Hd = float(cam["width"]), float(cam["height"])
    tx = 0.5 * Wd / fx
    ty = 0.5 * Hd / fy
    u = np.clip(t[0] / t[2], -(cx / fx + 0.3 * tx), (Wd - cx) / fx + 0.3 * tx) * t[2]
    v = np.clip(t[1] / t[2], -(cy / fy + 0.3 * ty), (Hd - cy) / fy + 0.3 * ty) * t[2]
    J = np.array([[fx / t[2], 0, -fx * u / t[2] ** 2], [0, fy / t[2], -fy * v / t[2] ** 2]])
    cov = J @ W @ Sig @ W.T @ J.T + eps2d * np.eye(2)
    conic = np.linalg.inv(cov)
    return np.array([conic[0, 0], conic[0, 1], conic[1, 1]]), \
        np.array([fx * t[0] / t[2] + cx, fy * t[1] / t[2] + cy]), t[2]


def test_project_matches_independent_ewa():
    rng = np.random.default_rng(5)
    cam = ref_test_camera(96, 64)
    c64 = O.cam64(cam)
    for _ in range(50):
        g = random_gaussian(rng)
        p = O.project_all_f64(g[None], c64, O.opts64(3))[0]
        conic, mean, depth = ewa_reference(g, cam)
        assert np.allclose(p[2:5], conic, rtol=1e-10, atol=1e-14)
        assert np.allclose(p[0:2], mean, rtol=1e-13)
        assert np.isclose(p[5], depth, rtol=1e-14)


def proj_loss(p15, w9):
    return float(np.dot(proj15_fields(p15), w9))


def check_projection(cam, g, degree, antialias, rng, label, prec):
    tol = F64 if prec == "f64" else F32
    if prec == "f64":
        c, o = O.cam64(cam), O.opts64(degree, antialias)
        base = O.project_all_f64(g[None], c, o)[0]
        assert base[6] > 0, f"{label}: test Gaussian must be visible"
        w = np.array([rnd(rng, -1, 1) for _ in range(9)])
        analytic = O.project_backward_all_f64(g[None], c, o, w[None])[0]

        def fwd(gg):
            p = O.project_all_f64(gg[None], c, o)[0]
            assert p[6] > 0, f"{label}: perturbed Gaussian left the view"
            return proj_loss(p, w)
        scale = 1.0 + abs(base[2]) + abs(base[4])
        dt = np.float64
    else:
        o = rec.proj_opts(degree, antialias)
        g32 = g.astype(np.float32)
        base = O.project_all(g32[None], cam, o)[0]
        assert base["radius"] > 0
        w = np.array([rnd(rng, -1, 1) for _ in range(9)], np.float32)
        wg = np.zeros(1, rec.PROJ_GRAD)
        wg["mean2d"], wg["conic"], wg["opacity"], wg["color"] = w[0:2], w[2:5], w[5], w[6:9]
        analytic = O.project_backward_all(g32[None], cam, o, wg)[0]

        def fwd(gg):
            p = O.project_all(gg[None].astype(np.float32), cam, o)[0]
            assert p["radius"] > 0
            vals = np.concatenate([p["mean2d"], p["conic"], [p["opacity"]], p["color"]])
            return float(np.dot(vals.astype(np.float32), w))
        scale = 1.0 + abs(float(base["conic"][0])) + abs(float(base["conic"][2]))
        dt = np.float32
        g = g32
    for i in range(59):
        def f(e):
            gg = g.copy()
            gg[i] = dt(gg[i] + dt(e))
            return fwd(gg)
        num = central(0.0, tol["H"], f)
        check(f"{label} param {i}", float(analytic[i]), num, scale, tol)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_projection_gradient_plain(prec):
    rng = np.random.default_rng(11)
    cam = ref_test_camera(96, 64)
    for n in range(12):
        check_projection(cam, random_gaussian(rng), 3, False, rng, f"plain #{n}", prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_projection_gradient_antialiased(prec):
    rng = np.random.default_rng(12)
    cam = ref_test_camera(96, 64)
    for n in range(12):
        check_projection(cam, random_gaussian(rng), 3, True, rng, f"antialiased #{n}", prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_projection_gradient_lower_degrees(prec):
    rng = np.random.default_rng(13)
    cam = ref_test_camera(96, 64)
    for degree in range(3):
        check_projection(cam, random_gaussian(rng), degree, False, rng, f"degree {degree}", prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_projection_gradient_with_jacobian_clamping(prec):
    rng = np.random.default_rng(14)
    cam = scenes.look_at(64, 48, 60.0, 60.0, (0.0, 0.0, -2.0), (0.0, 0.0, 0.0))
    g = random_gaussian(rng)
    g[0:3] = [3.2, 0.1, 0.0]
    g[3:6] = 0.4
    W = np.asarray(cam["rot"], np.float64).reshape(3, 3)
    t = W @ g[0:3] + np.asarray(cam["trans"], np.float64)
    lim = (64 - float(cam["cx"])) / 60.0 + 0.3 * 0.5 * 64 / 60.0
    assert t[0] / t[2] > lim, "test setup: Gaussian must be past the clamp limit"
    check_projection(cam, g, 2, True, rng, "clamped", prec)


def test_project_culls_behind_near_plane_and_nan():
    cam = ref_test_camera(96, 64)
    g = np.zeros((3, 59), np.float32)
    g[:, 6] = 1.0
    g[0, 0:3] = [0.0, 0.0, -2.9]   # depth 0.1 < near 0.2
    g[1, 0:3] = [np.nan, 0.0, 0.0]
    g[2, 0:3] = [0.0, 0.0, 0.0]
    g[2, 3:6] = -2.0
    p = O.project_all(g, cam, rec.proj_opts())
    assert p["radius"][0] == 0 and p["radius"][1] == 0 and p["radius"][2] > 0
    assert np.all(p[0]["conic"] == 0)
