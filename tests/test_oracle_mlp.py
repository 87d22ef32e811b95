"""Pins the oracle of the learned densification scorer (CPU only).

The scorer has no counterpart in the reference; its specification is 12 features -> 3 x 64
tanh-GELU -> 3 logits with a weighted binary cross-entropy.  The oracle's forward and
backward are checked against
- torch autograd on the same network in binary64 (an independent implementation),
- central finite differences of the oracle's own binary64 loss,
- torch's bfloat16 conversion for the rounding helper, and its tanh GELU,
so a dropped term, a transposed weight or a wrong derivative fails here.
"""
import numpy as np
import torch

import oracle_lib as O

N_PARAMS = O.MLP_PARAMS


def random_params(rng, scale=0.3):
    p = rng.normal(0, scale, N_PARAMS)
    return p


def torch_loss(p, x, y, w):
    """The same network written with torch ops (binary64)."""
    t = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1 = t[0:768].view(64, 12)
    b1 = t[768:832]
    W2 = t[832:4928].view(64, 64)
    b2 = t[4928:4992]
    W3 = t[4992:9088].view(64, 64)
    b3 = t[9088:9152]
    W4 = t[9152:9344].view(3, 64)
    b4 = t[9344:9347]
    xt = torch.tensor(x, dtype=torch.float64)
    g = torch.nn.functional.gelu
    h = g(xt @ W1.T + b1, approximate="tanh")
    h = g(h @ W2.T + b2, approximate="tanh")
    h = g(h @ W3.T + b3, approximate="tanh")
    z = h @ W4.T + b4
    yt = torch.tensor(np.stack([(y >> o) & 1 for o in range(3)], 1), dtype=torch.float64)
    wt = torch.tensor(w, dtype=torch.float64)[:, None]
    loss = torch.nn.functional.binary_cross_entropy_with_logits(z, yt, weight=wt,
                                                                reduction="sum") / len(x)
    loss.backward()
    return float(loss.detach()), t.grad.numpy(), z.detach().numpy()


def test_backward_matches_torch_autograd():
    rng = np.random.default_rng(0)
    p = random_params(rng)
    x = rng.normal(0, 1, (37, 12))
    y = rng.integers(0, 8, 37).astype(np.uint8)
    w = rng.uniform(0.5, 2.0, 37)
    loss, grad = O.mlp_loss_grad_f64(p, x, y, w)
    tl, tg, tz = torch_loss(p, x, y, w)
    z = O.mlp_forward_f64(p, x)
    assert np.allclose(z, tz, rtol=1e-12, atol=1e-12)
    assert abs(loss - tl) <= 1e-12 * abs(tl)
    err = np.abs(grad - tg).max() / np.abs(tg).max()
    assert err < 1e-11, err


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = random_params(rng)
    x = rng.normal(0, 1, (5, 12))
    y = rng.integers(0, 8, 5).astype(np.uint8)
    w = rng.uniform(0.5, 2.0, 5)
    _, grad = O.mlp_loss_grad_f64(p, x, y, w)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False)
                          for a, b in [(0, 768), (768, 832), (832, 4928), (4928, 4992),
                                       (4992, 9088), (9088, 9152), (9152, 9344)]]
                         + [np.arange(9344, 9347)])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.mlp_loss_grad_f64(pp, x, y, w)[0] - O.mlp_loss_grad_f64(pm, x, y, w)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-7 * max(1.0, abs(grad[k])) + 1e-9, (k, fd, grad[k])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(2)
    vals = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-3, 1000),
                           rng.normal(0, 1e4, 1000)]).astype(np.float32)
    # exact ties (low 16 bits = 0x8000) with even and odd kept parts
    ties = ((np.arange(200, dtype=np.uint32) << 16) | 0x3F800000 | 0x8000).view(np.float32)
    vals = np.concatenate([vals, ties])
    want = torch.tensor(vals).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(O.bf16_round(vals).view(np.uint32), want.view(np.uint32))


def test_gelu_matches_torch_tanh_form():
    rng = np.random.default_rng(3)
    p = random_params(rng, 0.5)
    x = rng.normal(0, 2, (50, 12))
    _, _, tz = torch_loss(p, x, np.zeros(50, np.uint8), np.ones(50))
    assert np.allclose(O.mlp_forward_f64(p, x), tz, rtol=1e-12, atol=1e-12)


def test_bf16_forward_close_to_exact():
    """The bf16-rounded f32 forward stays near the exact f64 network (bf16 has 8 bits)."""
    rng = np.random.default_rng(4)
    p = random_params(rng, 0.2).astype(np.float32)
    x = rng.normal(0, 1, (200, 12)).astype(np.float32)
    a = O.mlp_forward(p, x, bf16=True)
    b = O.mlp_forward_f64(p, x)
    assert np.abs(a - b).max() < 0.05 * np.abs(b).max()
    c = O.mlp_forward(p, x, bf16=False)
    assert np.abs(c - b).max() < 1e-5 * np.abs(b).max()


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1, N_PARAMS).astype(np.float32)
    tp = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 6):
        g = rng.normal(0, 1, N_PARAMS).astype(np.float32)
        tp.grad = torch.tensor(g)
        opt.step()
        c1 = np.float32(1) - np.float32(0.9) ** step
        c2 = np.float32(1) - np.float32(0.999) ** step
        O.mlp_adam(p, g, m, v, 1e-3, c1, c2)
    assert np.abs(p - tp.detach().numpy()).max() < 1e-6
