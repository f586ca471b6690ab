"""The CPU oracle of the learned smoother's network, pinned by a float64 numpy
restatement whose gradient is checked against central finite differences."""
import math

import numpy as np
import pytest

NP, NPP = 8385, 8448
OW1, OB1, OW2, OB2, OW3, OB3 = 0, 4096, 4160, 8256, 8320, 8384


def bf16(a):
    """round float32 values to bfloat16 (nearest even) and return the uint16 bits"""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(64, 64), p[OB1:OW2], p[OW2:OB2].reshape(64, 64), p[OB2:OW3], p[OW3:OB3], p[OB3])


def np_loss_grad(p, X, t):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = X.shape[0]
    z1 = X @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    y = h2 @ w3 + b3
    d = y - t
    loss = np.mean(d * d)
    dy = 2.0 * d / n
    g = np.zeros(NPP)
    g[OB3] = dy.sum()
    g[OW3:OB3] = h2.T @ dy
    dz2 = np.outer(dy, w3) * gelu_d(z2)
    g[OB2:OW3] = dz2.sum(0)
    g[OW2:OB2] = (dz2.T @ h1).ravel()
    dz1 = (dz2 @ W2) * gelu_d(z1)
    g[OB1:OW2] = dz1.sum(0)
    gw1 = dz1.T @ X
    gw1[:, 54:] = 0.0
    g[OW1:OB1] = gw1.ravel()
    return loss, g, y


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    W1 = rng.uniform(-0.3, 0.3, (64, 64)).astype(np.float32)
    W1[:, 54:] = 0.0
    p[OW1:OB1] = W1.ravel()
    p[OB1:OW2] = rng.uniform(-0.1, 0.1, 64)
    p[OW2:OB2] = rng.uniform(-0.3, 0.3, 64 * 64)
    p[OB2:OW3] = rng.uniform(-0.1, 0.1, 64)
    p[OW3:OB3] = rng.uniform(-0.2, 0.2, 64)
    p[OB3] = 0.05
    x = rng.uniform(-1, 1, (n, 64)).astype(np.float32)
    x[:, 54:] = 0.0
    Xb = bf16(x)
    t = rng.uniform(-1, 1, n).astype(np.float32)
    return p, Xb, t


def test_numpy_gradient_matches_finite_differences():
    p, Xb, t = random_problem(16, 0)
    X = from_bf16(Xb).astype(np.float64)
    t64 = t.astype(np.float64)
    _, g, _ = np_loss_grad(p, X, t64)
    rng = np.random.default_rng(1)
    idx = list(rng.choice(np.arange(OW1, OB1).reshape(64, 64)[:, :54].ravel(), 6)) + \
        list(rng.integers(OB1, NP, 10)) + [OB3]
    for k in idx:
        h = 1e-5
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[k] += h
        pm[k] -= h
        fd = (np_loss_grad(pp, X, t64)[0] - np_loss_grad(pm, X, t64)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_oracle_plain_network_matches_numpy(oracle):
    p, Xb, t = random_problem(200, 2)
    X = from_bf16(Xb).astype(np.float64)
    loss, g, y = np_loss_grad(p, X, t.astype(np.float64))
    go = oracle.mlp_grad(p, Xb, t, round_bf16=False)
    yo = oracle.mlp_forward(p, Xb, round_bf16=False)
    assert np.allclose(yo, y, rtol=1e-5, atol=1e-6)
    assert np.linalg.norm(go[:NP] - g[:NP]) <= 1e-5 * np.linalg.norm(g[:NP])
    assert np.isclose(go[NPP - 1] / 200, loss, rtol=1e-5)


def test_oracle_bf16_rounding_is_small(oracle):
    p, Xb, t = random_problem(200, 3)
    g0 = oracle.mlp_grad(p, Xb, t, round_bf16=False)
    g1 = oracle.mlp_grad(p, Xb, t, round_bf16=True)
    rel = np.linalg.norm(g1[:NP] - g0[:NP]) / np.linalg.norm(g0[:NP])
    assert 0 < rel < 2e-2
