"""Pins of the oracle's relaxation-time MLP (6 -> 32 -> 32 -> 1, SiLU,
tau = softplus + 0.5): forward and gradient against torch autograd in FP64,
gradient against central finite differences, features against their
definitions."""
import math

import numpy as np
import pytest

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def random_blob(rng, oracle):
    p = np.zeros(oracle.MLP_BLOB)
    n = oracle.MLP_NPARAM
    p[:n] = rng.normal(0, 0.4, n)
    p[n:n + 6] = rng.normal(0, 0.1, 6)         # mu
    p[n + 6:n + 12] = rng.uniform(0.5, 2.0, 6)  # isig
    return p


def torch_forward(p, x):
    import torch
    P = torch.tensor(p[:1313], dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(p[1313:1319], dtype=torch.float64)
    isg = torch.tensor(p[1319:1325], dtype=torch.float64)
    W1 = P[0:192].reshape(32, 6)
    b1 = P[192:224]
    W2 = P[224:1248].reshape(32, 32)
    b2 = P[1248:1280]
    W3 = P[1280:1312]
    b3 = P[1312]
    X = (torch.tensor(x, dtype=torch.float64) - mu) * isg
    h1 = torch.nn.functional.silu(X @ W1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
    o = h2 @ W3 + b3
    return torch.nn.functional.softplus(o) + 0.5, P


def test_forward_matches_torch(oracle):
    rng = np.random.default_rng(0)
    p = random_blob(rng, oracle)
    xs = rng.normal(0, 1, (20, 6))
    t, _ = torch_forward(p, xs)
    for i in range(20):
        assert abs(oracle.mlp_forward(p, xs[i]) - float(t[i])) < 1e-13


def test_gradient_matches_torch(oracle):
    rng = np.random.default_rng(1)
    p = random_blob(rng, oracle)
    xs = rng.normal(0, 1, (16, 6))
    ys = rng.uniform(1.4, 1.99, 16)
    loss, g = oracle.mlp_loss_grad(p, xs, ys)
    import torch
    t, P = torch_forward(p, xs)
    L = 0.5 * ((1.0 / t - torch.tensor(ys)) ** 2).mean()
    L.backward()
    assert abs(loss - float(L)) < 1e-14
    assert np.allclose(g, P.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_gradient_central_differences(oracle):
    rng = np.random.default_rng(2)
    p = random_blob(rng, oracle)
    xs = rng.normal(0, 1, (8, 6))
    ys = rng.uniform(1.4, 1.99, 8)
    _, g = oracle.mlp_loss_grad(p, xs, ys)
    h = 1e-6
    for i in rng.choice(oracle.MLP_NPARAM, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.mlp_loss_grad(pp, xs, ys)[0] - oracle.mlp_loss_grad(pm, xs, ys)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), i


def test_features_definition(oracle):
    rng = np.random.default_rng(3)
    for _ in range(10):
        rho, ux, uy = rng.uniform(0.95, 1.05), rng.uniform(-.1, .1), rng.uniform(-.1, .1)
        f = oracle.feq(rho, ux, uy) + W * rng.uniform(-1e-2, 1e-2, 9)
        feat = oracle.mlp_features(f, rho, ux, uy)
        neq = f - oracle.feq(rho, ux, uy)
        pxx, pxy, pyy = (neq * EX * EX).sum(), (neq * EX * EY).sum(), (neq * EY * EY).sum()
        c = 1.5 / rho
        S = math.sqrt(2 * c * c * (pxx ** 2 + 2 * pxy ** 2 + pyy ** 2))
        assert np.allclose(feat, [rho, ux, uy, S, pxx - pyy, pxy], rtol=1e-12, atol=1e-16)


def test_tau_at_least_half(oracle):
    rng = np.random.default_rng(4)
    p = random_blob(rng, oracle)
    for _ in range(50):
        assert oracle.mlp_forward(p, rng.normal(0, 3, 6)) > 0.5
