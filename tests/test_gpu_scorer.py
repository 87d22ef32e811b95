"""The learned densification scorer's matrix-core kernels against the oracle (needs a
gfx950 GPU).

Forward logits and loss gradients of the bfloat16 MLP are compared with the oracle's
restatement, which rounds at the same points (weights, features, activations and error
signals to bfloat16) and sums in binary32 in index order.  The matrix cores sum in their
own order, and a sum that lands near a bfloat16 rounding boundary can round to the
neighbouring value, so the bars are tolerances: logits within 1e-2 of the logit scale and
gradients within 2e-2 relative (L2, per parameter block), with a cosine >= 0.999.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from splatrs_amd import abi
from splatrs_amd import records as rec

pytestmark = pytest.mark.gpu
BLOCKS = [(0, 768), (768, 832), (832, 4928), (4928, 4992), (4992, 9088), (9088, 9152),
          (9152, 9344), (9344, 9347)]


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


def problem(seed, n):
    rng = np.random.default_rng(seed)
    p = (rng.normal(0, 1, rec.SCORER_PARAMS) * 0.25).astype(np.float32)
    x = rng.normal(0, 1, (n, 12)).astype(np.float32)
    y = rng.integers(0, 8, n).astype(np.uint8)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return p, x, y, w


@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 70000])
def test_forward_matches_oracle(ctx, n):
    p, x, _, _ = problem(n, n)
    say(f"scorer forward n={n}")
    s = abi.Scorer(ctx, p)
    got = s.forward(x)
    want = O.mlp_forward(p, x, bf16=True)
    scale = np.abs(want).max()
    err = np.abs(got - want)
    assert err.max() <= 1e-2 * scale, (err.max(), scale)
    assert err.mean() <= 1e-3 * scale
    s.close()


@pytest.mark.parametrize("n", [1, 64, 100, 4096, 20000])
def test_gradient_matches_oracle(ctx, n):
    p, x, y, w = problem(100 + n, n)
    say(f"scorer gradient n={n}")
    s = abi.Scorer(ctx, p)
    loss, g = s.loss_grad(x, y, w)
    wl, wg = O.mlp_loss_grad(p, x, y, w, bf16=True)
    assert abs(loss - wl) <= 1e-4 * abs(wl) + 1e-6
    for a, b in BLOCKS:
        ga, wa = g[a:b].astype(np.float64), wg[a:b].astype(np.float64)
        rel = np.linalg.norm(ga - wa) / (np.linalg.norm(wa) + 1e-30)
        cos = ga @ wa / (np.linalg.norm(ga) * np.linalg.norm(wa) + 1e-30)
        assert rel <= 2e-2 and cos >= 0.999, (a, b, rel, cos)
    s.close()


def test_params_round_trip_and_seeded_init(ctx):
    a = abi.Scorer(ctx, seed=5)
    b = abi.Scorer(ctx, seed=5)
    c = abi.Scorer(ctx, seed=6)
    pa, pb, pc = a.params(), b.params(), c.params()
    assert np.array_equal(pa, pb) and not np.array_equal(pa, pc)
    assert not pa[768:832].any() and np.abs(pa[:768]).std() > 0.1
    q = pa * 2
    a.set_params(q)
    assert np.array_equal(a.params(), q)
    for s in (a, b, c):
        s.close()


def test_fit_learns_a_rule_and_matches_oracle_adam(ctx):
    """Labels from a fixed rule on the features; minibatch Adam on the GPU drives the loss
    down and classifies held-out samples; one full-batch step equals the oracle's Adam on
    the oracle's gradient to the gradient tolerance."""
    rng = np.random.default_rng(9)
    n = 32768
    x = rng.normal(0, 1, (n, 12)).astype(np.float32)
    y = ((x[:, 0] + x[:, 1] > 0.5).astype(np.uint8)
         | ((x[:, 2] * x[:, 3] > 0.2).astype(np.uint8) << 1)
         | ((x[:, 4] < -0.8).astype(np.uint8) << 2))
    w = np.ones(n, np.float32)
    s = abi.Scorer(ctx, seed=1)
    l0, _ = s.loss_grad(x[:4096], y[:4096], w[:4096])
    loss = s.fit(x[:28672], y[:28672], w[:28672], rec.scorer_fit_config(epochs=8, batch=1024,
                                                                       lr=3e-3, seed=3))
    say(f"scorer fit: loss {l0:.4f} -> {loss:.4f}")
    assert loss < 0.5 * l0
    z = s.forward(x[28672:])
    pred = ((z[:, 0] > 0).astype(np.uint8) | ((z[:, 1] > 0).astype(np.uint8) << 1)
            | ((z[:, 2] > 0).astype(np.uint8) << 2))
    acc = np.mean([np.mean(((pred >> o) & 1) == ((y[28672:] >> o) & 1)) for o in range(3)])
    say(f"held-out accuracy {acc:.4f}")
    assert acc > 0.9
    # one Adam step on a single full minibatch vs the oracle
    p0 = s.params()
    s.set_params(p0)
    s.fit(x[:2048], y[:2048], w[:2048], rec.scorer_fit_config(epochs=1, batch=2048, lr=1e-3))
    _, g = O.mlp_loss_grad(p0, x[:2048], y[:2048], w[:2048], bf16=True)
    po, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    O.mlp_adam(po, g, m, v, 1e-3, np.float32(0.1), np.float32(1 - 0.999))
    step_gpu, step_cpu = s.params() - p0, po - p0
    cos = step_gpu.astype(np.float64) @ step_cpu / (np.linalg.norm(step_gpu) *
                                                     np.linalg.norm(step_cpu))
    assert cos > 0.99, cos
    s.close()
