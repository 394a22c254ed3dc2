"""Pins of the guide oracle (FP64 MLP): analytic gradient vs finite
differences, mixture normalisation, vMF sampling density."""
import math

import numpy as np
import pytest

import oracle


def net_shape(width=64):
    """Layer fan-ins, fan-outs and (W, b) offsets of the guide network: width 64
    with 8 lobes (48 outputs) or width 128 with 16 lobes (80 outputs)."""
    nout = 48 if width == 64 else 80
    K = [224, width, width, width, width]
    N = [width, width, width, width, nout]
    offs, o = [], 0
    for l in range(5):
        offs.append((o, o + K[l] * N[l]))
        o += K[l] * N[l] + N[l]
    return K, N, offs, o


def random_params(seed=0, out_scale=0.3, width=64):
    rng = np.random.default_rng(seed)
    K, N, _, _ = net_shape(width)
    parts = []
    for l in range(5):
        a = math.sqrt(6.0 / (K[l] + N[l])) * (out_scale if l == 4 else 1.0)
        parts.append(rng.uniform(-a, a, K[l] * N[l]))
        parts.append(rng.uniform(-0.1, 0.1, N[l]))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == oracle.mlp_num_params(width)
    return p


def random_samples(n, seed=1):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 11), np.float32)
    s[:, :7] = rng.random((n, 7))
    d = rng.standard_normal((n, 3))
    s[:, 7:10] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 10] = rng.exponential(1.0, n)
    return s


@pytest.mark.parametrize("width", [64, 128])
def test_analytic_gradient_matches_finite_differences(width):
    P = random_params(3, width=width)
    S = random_samples(6, 4)
    _, g = oracle.mlp_loss_grad(P, S, bf16=False, width=width)
    rng = np.random.default_rng(5)
    _, N, offs, total = net_shape(width)
    bounds = [b for wo, bo in offs for b in (wo, bo)]
    idx = np.concatenate([rng.integers(bounds[i], bounds[i + 1], 8) for i in range(len(bounds) - 1)] +
                         [np.arange(offs[4][1], offs[4][1] + 5 * (8 if width == 64 else 16))])
    checked = 0
    for p in idx:
        h = np.float32(1e-4)  # small enough not to cross ReLU kinks
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        d = float(Pp[p]) - float(Pm[p])
        lp, _ = oracle.mlp_loss_grad(Pp, S, bf16=False, width=width)
        lm, _ = oracle.mlp_loss_grad(Pm, S, bf16=False, width=width)
        fd = (lp - lm) / d
        if abs(g[p]) < 1e-7 and abs(fd) < 1e-7:
            continue
        assert abs(fd - g[p]) <= 1e-3 * max(abs(g[p]), abs(fd)) + 1e-9, (p, fd, g[p])
        checked += 1
    assert checked > 60


@pytest.mark.parametrize("seed,width", [(0, 64), (1, 64), (2, 64), (3, 128)])
def test_mixture_pdf_integrates_to_one(seed, width):
    rng = np.random.default_rng(seed)
    o = rng.standard_normal(48 if width == 64 else 80)
    o[1::5] = np.clip(o[1::5], -3, 3)      # kappa up to e^3 for a quadrature-friendly check
    n = 400_000  # Fibonacci sphere
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = math.pi * (1 + 5 ** 0.5) * i
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    q = oracle.mlp_mixture_pdf(o, d, width=width)
    assert abs(q.mean() * 4 * math.pi - 1.0) < 2e-3


def test_bf16_emulation_close_to_fp64():
    P = random_params(7)
    X = np.random.default_rng(8).random((32, 7)).astype(np.float32)
    a = oracle.mlp_forward(P, X, bf16=True)
    b = oracle.mlp_forward(P, X, bf16=False)
    assert np.abs(a - b).max() < 0.05 * np.abs(b).max()


def e4m3(x):
    """Round to the nearest OCP e4m3 value (ties to even), after clamping to
    +-448: 3 mantissa bits, exponent bias 7, subnormal step 2^-9."""
    x = np.clip(np.asarray(x, np.float64), -448.0, 448.0)
    a = np.abs(x)
    e = np.maximum(np.floor(np.log2(np.maximum(a, 2.0 ** -10))), -6.0)
    step = 2.0 ** (e - 3)
    return np.sign(x) * np.round(a / step) * step


def act_scale(M):
    """Power-of-two scale 2^k of a group of activations with maximum M: the largest
    k with M 2^k <= 448, capped at 126 (M = m 2^e, m in [0.5, 1): k = 8 - e when
    m > 0.875, else 9 - e); 1 for M = 0."""
    M = np.asarray(M, np.float64)
    m, e = np.frexp(np.where(M > 0, M, 1.0))
    k = np.minimum(np.where(m > 0.875, 8, 9) - e, 126)
    return np.where(M > 0, np.ldexp(1.0, k), 1.0)


def test_act_scale_is_the_largest_power_of_two_within_range():
    rng = np.random.default_rng(3)
    M = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), [448.0, 448.0 * (1 + 2.0 ** -52), 224.0, 1.0, 0.875, 1e-38]])
    s = act_scale(M)
    assert (np.log2(s) == np.round(np.log2(s))).all()       # powers of two
    assert (M * s <= 448.0).all()                          # within range, no clamp needed
    big = np.log2(s) < 126
    assert (M[big] * s[big] * 2 > 448.0).all()             # and the largest such
    assert act_scale(448.0) == 1.0 and act_scale(449.0) == 0.5 and act_scale(1.0) == 256.0
    assert act_scale(0.0) == 1.0 and act_scale(1e-40) == 2.0 ** 126


def mlp_forward_fp8(params, inputs, quant=e4m3, tile=64, width=64, group=16):
    """FP8 inference pipeline of the guide network restated in numpy: weights
    scaled per layer by 448 / max|W|, one-blob features by 448, each hidden
    layer's ReLU outputs per group of 16 rows of a 64-sample tile by the largest
    power of two 2^k with (group maximum) 2^k <= 448 (k capped at 126), the padding
    rows of a partial tile taking part in that maximum as zero-feature samples.
    quant=identity gives the exact network."""
    K, N, _, _ = net_shape(width)
    P = np.asarray(params, np.float64)
    Ws, bs, o = [], [], 0
    for l in range(5):
        Ws.append(P[o:o + K[l] * N[l]].reshape(K[l], N[l]))
        o += K[l] * N[l]
        bs.append(P[o:o + N[l]])
        o += N[l]
    Wq, inv_w = [], []
    for W in Ws:
        m = np.abs(W).max()
        Wq.append(quant(W * (448.0 / m)))
        inv_w.append(m / 448.0)
    X = np.asarray(inputs, np.float32).reshape(-1, 7)
    n = X.shape[0]
    out = np.zeros((n, N[4]))
    centres = np.arange(32, dtype=np.float32) + np.float32(0.5)
    for t0 in range(0, n, tile):
        xs = X[t0:t0 + tile]
        d = xs[:, :, None] * np.float32(32) - centres[None, None, :]
        f = np.zeros((tile, 224))
        f[:len(xs)] = np.exp(np.float32(-0.5) * d * d).reshape(len(xs), 224)
        h = quant(f * 448.0)
        inv_a = np.full((tile, 1), 1.0 / 448.0)
        for l in range(4):
            z = np.maximum(h @ Wq[l] * (inv_a * inv_w[l]) + bs[l], 0.0)
            M = np.repeat(z.reshape(tile // group, -1).max(axis=1), group)[:, None]
            s = act_scale(M)
            inv_a = 1.0 / s
            h = quant(z * s)
        out[t0:t0 + len(xs)] = (h @ Wq[4] * (inv_a * inv_w[4]) + bs[4])[:len(xs)]
    return out


def test_e4m3_rounding_known_values():
    # largest normal, smallest subnormal, ties to even in both directions,
    # the binade edge, clamping
    assert e4m3(448.0) == 448.0 and e4m3(1e6) == 448.0 and e4m3(-500.0) == -448.0
    assert e4m3(2.0 ** -9) == 2.0 ** -9 and e4m3(2.0 ** -11) == 0.0 and e4m3(0.75 * 2.0 ** -9) == 2.0 ** -9
    assert e4m3(1.0625) == 1.0 and e4m3(1.1875) == 1.25 and e4m3(15.5) == 16.0
    assert e4m3(2.0 ** -6 * 1.0625) == 2.0 ** -6   # smallest normal binade, step 2^-9
    assert e4m3(240.0) == 240.0 and e4m3(232.0) == 224.0 and e4m3(432.0) == 448.0
    # every representable positive value maps to itself: 7 subnormals + 15 binades x 8
    vals = [k * 2.0 ** -9 for k in range(1, 8)] + [(1 + m / 8) * 2.0 ** e for e in range(-6, 9) for m in range(8)]
    vals = [v for v in vals if v <= 448.0]
    assert len(vals) == 7 + 15 * 8 - 1
    assert np.array_equal(e4m3(np.array(vals)), np.array(vals))


@pytest.mark.parametrize("width", [64, 128])
def test_fp8_pipeline_without_rounding_is_the_network(width):
    """Also pins the oracle's forward pass and parameter layout: numpy matrix
    products over the documented layout give the same outputs."""
    P = random_params(11, width=width)
    X = np.random.default_rng(12).random((150, 7)).astype(np.float32)
    exact = mlp_forward_fp8(P, X, quant=lambda v: v, width=width)
    r = oracle.mlp_forward(P, X, bf16=False, width=width)
    assert np.abs(exact - r).max() < 1e-5 * np.abs(r).max()
    q = mlp_forward_fp8(P, X, width=width)
    rel = np.linalg.norm(q - r) / np.linalg.norm(r)
    assert 1e-3 < rel < 0.1, rel


def test_blob_ratio_recurrence():
    """The FP8 encoding's recurrence: blob(x, b+k+1) = blob(x, b+k) exp(d_k - 1/2),
    d_k = 32 x - (b + k + 1/2), with the ratio shrinking by e^-1 per bin; restated in
    float32 over runs of 8 bins, it stays within 1e-5 of the direct exponentials for
    every bin that is not 0 in FP8 (within 5.2 of the peak)."""
    rng = np.random.default_rng(3)
    x = rng.random(4000).astype(np.float32)
    worst = 0.0
    for b in range(0, 32, 8):
        d0 = x * np.float32(32) - np.float32(b + 0.5)
        val = np.exp(np.float32(-0.5) * d0 * d0).astype(np.float32)
        r = np.exp(d0 - np.float32(0.5)).astype(np.float32)
        for k in range(8):
            d = x.astype(np.float64) * 32 - (b + k + 0.5)
            direct = np.exp(-0.5 * d * d)
            near = np.abs(d) <= 5.2
            worst = max(worst, float((np.abs(val - direct) / direct)[near].max(initial=0.0)))
            val = (val * r).astype(np.float32)
            r = (r * np.float32(np.exp(-1.0))).astype(np.float32)
    assert worst < 1e-5, worst
