"""Pin of the oracle's Adam step (orc_mlp_adam; no reference code exists for the
learned guide) by known answers worked out by hand from Kingma and Ba (2015),
Algorithm 1: m_t = b1 m + (1 - b1) g, v_t = b2 v + (1 - b2) g^2,
m^ = m_t / (1 - b1^t), v^ = v_t / (1 - b2^t), theta -= lr m^ / (sqrt(v^) + eps)."""
import numpy as np

import oracle


def test_adam_first_step_is_normalised_gradient():
    """With zero moments the bias corrections give m^ = g, v^ = g^2 at step 1, so
    theta_1 = theta_0 - lr g / (|g| + eps)."""
    n = oracle.mlp_num_params()
    rng = np.random.default_rng(2)
    p0 = rng.normal(size=n).astype(np.float32)
    g = rng.normal(size=n) * 10.0 ** rng.uniform(-6, 2, size=n)
    p, _, _ = oracle.mlp_adam(p0, np.zeros(n), np.zeros(n), g, 1, lr=0.01, b1=0.9, b2=0.99, eps=1e-8)
    want = p0.astype(np.float64) - 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p, want, rtol=0, atol=2e-7 * (1 + np.abs(want)))


def test_adam_three_steps_known_answer():
    """theta_0 = 1, lr 0.1, b1 0.9, b2 0.99, eps 1e-8, gradients 2, -1, 0.5:
    t=1: m = 0.2,   v = 0.04,     m^ = 2,               v^ = 4,               theta = 1 - 0.1*2/2         = 0.9000000005
    t=2: m = 0.08,  v = 0.0496,   m^ = 0.08/0.19,        v^ = 0.0496/0.0199,   theta = 0.9 - 0.0266699427  = 0.8733300578
    t=3: m = 0.122, v = 0.051604, m^ = 0.122/0.271,      v^ = 0.051604/0.029701, theta = 0.87333 - 0.0341534 = 0.8391766228
    (a second parameter sees the negated gradients and mirrors the first about 1; a
    third sees zero gradients and stays put)."""
    n = oracle.mlp_num_params()
    p = np.zeros(n, np.float32)
    p[:3] = 1.0
    m, v = np.zeros(n), np.zeros(n)
    want = [0.9000000005, 0.8733300578, 0.8391766228]
    for t, gt in enumerate([2.0, -1.0, 0.5], 1):
        g = np.zeros(n)
        g[0], g[1] = gt, -gt
        p, m, v = oracle.mlp_adam(p, m, v, g, t, lr=0.1, b1=0.9, b2=0.99, eps=1e-8)
        assert abs(p[0] - want[t - 1]) < 2e-7, (t, p[0])  # float parameters
        assert abs(p[1] - (2.0 - want[t - 1])) < 2e-7, (t, p[1])
        assert p[2] == 1.0
