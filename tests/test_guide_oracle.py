"""Pins of the guide oracle (FP64 MLP): analytic gradient vs finite
differences, mixture normalisation, vMF sampling density."""
import math

import numpy as np
import pytest

import oracle


def random_params(seed=0, out_scale=0.3):
    rng = np.random.default_rng(seed)
    K = [224, 64, 64, 64, 64]
    N = [64, 64, 64, 64, 48]
    parts = []
    for l in range(5):
        a = math.sqrt(6.0 / (K[l] + N[l])) * (out_scale if l == 4 else 1.0)
        parts.append(rng.uniform(-a, a, K[l] * N[l]))
        parts.append(rng.uniform(-0.1, 0.1, N[l]))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == oracle.mlp_num_params()
    return p


def random_samples(n, seed=1):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 11), np.float32)
    s[:, :7] = rng.random((n, 7))
    d = rng.standard_normal((n, 3))
    s[:, 7:10] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 10] = rng.exponential(1.0, n)
    return s


def test_analytic_gradient_matches_finite_differences():
    P = random_params(3)
    S = random_samples(6, 4)
    _, g = oracle.mlp_loss_grad(P, S, bf16=False)
    rng = np.random.default_rng(5)
    offs = [0, 14336, 14400, 18496, 18560, 22656, 22720, 26816, 26880, 29952]
    idx = np.concatenate([rng.integers(offs[i], offs[i + 1], 8) for i in range(len(offs) - 1)] +
                         [np.arange(29952, 29992)])
    checked = 0
    for p in idx:
        h = np.float32(1e-4)  # small enough not to cross ReLU kinks
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        d = float(Pp[p]) - float(Pm[p])
        lp, _ = oracle.mlp_loss_grad(Pp, S, bf16=False)
        lm, _ = oracle.mlp_loss_grad(Pm, S, bf16=False)
        fd = (lp - lm) / d
        if abs(g[p]) < 1e-7 and abs(fd) < 1e-7:
            continue
        assert abs(fd - g[p]) <= 1e-3 * max(abs(g[p]), abs(fd)) + 1e-9, (p, fd, g[p])
        checked += 1
    assert checked > 60


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mixture_pdf_integrates_to_one(seed):
    rng = np.random.default_rng(seed)
    o = rng.standard_normal(48)
    o[1::5] = np.clip(o[1::5], -3, 3)      # kappa up to e^3 for a quadrature-friendly check
    n = 400_000  # Fibonacci sphere
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = math.pi * (1 + 5 ** 0.5) * i
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    q = oracle.mlp_mixture_pdf(o, d)
    assert abs(q.mean() * 4 * math.pi - 1.0) < 2e-3


def test_bf16_emulation_close_to_fp64():
    P = random_params(7)
    X = np.random.default_rng(8).random((32, 7)).astype(np.float32)
    a = oracle.mlp_forward(P, X, bf16=True)
    b = oracle.mlp_forward(P, X, bf16=False)
    assert np.abs(a - b).max() < 0.05 * np.abs(b).max()
