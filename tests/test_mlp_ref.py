"""Pins of the CPU reference network (tests/mlp_ref.py) that the device kernels are checked against:
analytic gradients against central finite differences, Adam's first step against its closed form."""
import numpy as np

import mlp_ref as R


def setup(width=128, hidden=4, n=64, seed=0):
    rng = np.random.default_rng(seed)
    layers = [(width, 4)] + [(width, width)] * (hidden - 1) + [(4, width)]
    npar = sum(o * i + o for o, i in layers)
    p = rng.normal(0, 0.3, npar)
    X = rng.uniform(-1, 1, (n, 4))
    T = rng.normal(0, 1, (n, 4))
    sc = (rng.uniform(-0.5, 0.5, 4), rng.uniform(0.5, 2, 4), rng.uniform(-0.5, 0.5, 4), rng.uniform(0.5, 2, 4))
    return layers, p, X, T, sc


def test_reference_gradient_matches_finite_differences():
    for width, hidden in ((8, 2), (128, 4)):
        layers, p, X, T, sc = setup(width, hidden)
        loss, g = R.loss_and_grad(p, layers, X, T, sc)
        rng = np.random.default_rng(1)
        idx = rng.choice(p.size, 60, replace=False)
        h = 1e-6
        for i in idx:
            e = np.zeros_like(p)
            e[i] = h
            fd = (R.loss_and_grad(p + e, layers, X, T, sc)[0] - R.loss_and_grad(p - e, layers, X, T, sc)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()), (width, i, fd, g[i])


def test_reference_adam_first_step_is_a_sign_step():
    rng = np.random.default_rng(2)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m, v = R.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=1e-3)
    # first step: m_hat = g, v_hat = g^2, so the update is lr * g / (|g| + eps)
    assert np.allclose(p - p1, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
