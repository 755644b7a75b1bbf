"""The warm-start network's CPU restatement (oracle/nn_oracle.py), pinned by
closed forms and finite differences (float64, no BF16 rounding)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nn_oracle as NO  # noqa: E402


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    G, K, N = 4, 3, 2
    x = rng.standard_normal((G, G, G, K))
    W = rng.standard_normal((27, N, K))
    b = rng.standard_normal(N)
    y = NO.conv(x, W, b)
    # direct 27-tap sum at a few voxels, taps outside the grid are zero
    for (z, yy, xx) in [(0, 0, 0), (1, 2, 3), (3, 3, 3), (2, 0, 1)]:
        ref = b.copy()
        for t in range(27):
            dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
            a, c, e = xx + dx, yy + dy, z + dz
            if 0 <= a < G and 0 <= c < G and 0 <= e < G:
                ref += W[t] @ x[e, c, a]
        np.testing.assert_allclose(y[z, yy, xx], ref, rtol=1e-12, atol=1e-12)


def test_conv_adjoints():
    """<conv(x), dy> = <x, dgrad(dy)> and d<conv(x), dy>/dW = wgrad(dy, x)."""
    rng = np.random.default_rng(1)
    G, K, N = 4, 3, 5
    x = rng.standard_normal((G, G, G, K))
    dy = rng.standard_normal((G, G, G, N))
    W = rng.standard_normal((27, N, K))
    lhs = np.sum(NO.conv(x, W, np.zeros(N)) * dy)
    np.testing.assert_allclose(lhs, np.sum(x * NO.conv_dgrad(dy, W)), rtol=1e-12)
    np.testing.assert_allclose(lhs, np.sum(W * NO.conv_wgrad(dy, x)), rtol=1e-12)


def test_pool_up_adjoints():
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal((4, 4, 4, 2)), rng.standard_normal((2, 2, 2, 2))
    np.testing.assert_allclose(np.sum(NO.pool(a) * b), np.sum(a * NO.pool_bwd(b)), rtol=1e-12)
    np.testing.assert_allclose(np.sum(NO.up(b) * a), np.sum(b * NO.up_bwd(a)), rtol=1e-12)


def test_gradient_finite_differences():
    rng = np.random.default_rng(3)
    c, G = (2, 3, 2, 2), 8
    p = rng.standard_normal(NO.nparams(c)) * 0.3
    x = rng.standard_normal((G, G, G, NO.NIN))
    t = rng.standard_normal((G, G, G, 3))
    loss, g = NO.loss_grad(p, c, x, t)
    idx = rng.choice(len(p), 40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h; pm[i] -= h
        fd = (NO.loss_grad(pp, c, x, t)[0] - NO.loss_grad(pm, c, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.14159, 0.0])
    r = NO.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2 ** -7                 # nearest of the 2^-7 spaced values
    assert r[3] == -3.140625 and r[4] == 0.0


def test_adam_closed_forms():
    """Adam's bias correction makes the moment estimates of a constant
    gradient exact at every step (m_hat = g, v_hat = g^2), so each step moves a
    parameter by lr g / (|g| + eps); a dropped correction (a first step of
    lr (1-b1)/sqrt(1-b2) = 3.16 lr) or swapped betas fail.  A gradient that
    changes sign mixes the moments as Kingma & Ba's Algorithm 1 states."""
    rng = np.random.default_rng(4)
    g = rng.standard_normal(64)
    p0 = rng.standard_normal(64)
    m, v, p = np.zeros(64), np.zeros(64), p0.copy()
    lr, eps = 1e-3, 1e-8
    for t in range(1, 6):
        p = NO.adam(p, g, m, v, t, lr)
        np.testing.assert_allclose(p, p0 - t * lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15 * t)
    # second step with the gradient reversed: m = 0.09 g - 0.1 g, v = (0.000999 + 0.001) g^2
    m, v, p = np.zeros(64), np.zeros(64), p0.copy()
    p1 = NO.adam(p, g, m, v, 1, lr)
    p2 = NO.adam(p1, -g, m, v, 2, lr)
    mh = (0.9 * 0.1 - 0.1) * g / (1 - 0.9 ** 2)
    vh = (0.999 * 0.001 + 0.001) * g * g / (1 - 0.999 ** 2)
    np.testing.assert_allclose(p2, p1 - lr * mh / (np.sqrt(vh) + eps), rtol=1e-13, atol=1e-16)
