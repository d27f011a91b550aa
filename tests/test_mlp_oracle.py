"""Pins of the CPU reference MLP (oracle/mlp_oracle.py): its analytic
gradient against central finite differences of the FP64 network, the
closed form of a one-layer network, and the BF16 rounding rule."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def rand_params(widths, rng, scale=0.5):
    n = sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(len(widths) - 1))
    return rng.normal(0.0, scale, n)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    widths = [5, 7, 6, 3]
    p = rand_params(widths, rng)
    x = rng.normal(size=(4, 5))
    t = rng.normal(size=(4, 3))
    _, g = mo.loss_and_grad(p, widths, x, t, emulate_bf16=False)
    h = 1e-6
    for k in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = mo.loss_and_grad(p + e, widths, x, t, emulate_bf16=False)
        lm, _ = mo.loss_and_grad(p - e, widths, x, t, emulate_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(fd)), (k, fd, g[k])


def test_one_layer_closed_form():
    # y = x W^T + b, loss = mean((y-t)^2): dW = 2/n (y-t)^T x, db = 2/n sum(y-t)
    rng = np.random.default_rng(1)
    widths = [3, 2]
    p = rand_params(widths, rng)
    x = rng.normal(size=(5, 3))
    t = rng.normal(size=(5, 2))
    W = p[:6].reshape(2, 3)
    b = p[6:]
    y = x @ W.T + b
    loss, g = mo.loss_and_grad(p, widths, x, t, emulate_bf16=False)
    assert np.isclose(loss, ((y - t) ** 2).mean())
    assert np.allclose(g[:6].reshape(2, 3), 2.0 / y.size * (y - t).T @ x)
    assert np.allclose(g[6:], 2.0 / y.size * (y - t).sum(axis=0))


def test_gelu_derivative():
    x = np.linspace(-6, 6, 101)
    h = 1e-6
    assert np.allclose(mo.gelu_d(x), (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h), atol=1e-8)
    assert mo.gelu(0.0) == 0.0 and np.isclose(mo.gelu(10.0), 10.0)


def test_bf16_rounding():
    assert mo.bf16(1.0) == 1.0
    assert mo.bf16(1.0 + 2.0 ** -8) == 1.0            # tie, to even
    assert mo.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie, to even (up)
    assert mo.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7  # representable
    x = np.random.default_rng(2).normal(size=1000)
    assert np.all(np.abs(mo.bf16(x) - x) <= np.abs(x) * 2.0 ** -8)
