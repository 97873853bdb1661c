"""The FP64 MLP oracle: its analytic backward against central differences,
closed forms of degenerate networks, and the BF16 rounding it emulates."""
import numpy as np
import pytest

from oracle import mlp


def test_param_count_matches_library_header():
    assert mlp.NPARAM == 40856


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7                       # BF16 spacing in [1, 2)
    assert mlp.bf16(one + 0.4 * ulp) == one
    assert mlp.bf16(one + 0.6 * ulp) == one + ulp
    assert mlp.bf16(one + 0.5 * ulp) == one          # tie -> even (mantissa 0)
    assert mlp.bf16(one + 1.5 * ulp) == one + 2 * ulp  # tie -> even
    assert mlp.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).normal(size=1000)
    assert np.all(np.abs(mlp.bf16(x) - x) <= np.abs(x) * 2.0 ** -8 * 1.0001)


def test_analytic_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = mlp.glorot(3) + 0.05 * rng.normal(size=mlp.NPARAM)   # nonzero biases too
    x = rng.normal(size=(5, mlp.IN))
    dy = rng.normal(size=(5, mlp.OUT))
    g = mlp.backward(p, x, dy, emulate=False)
    loss = lambda q: float(np.sum(dy * mlp.forward(q, x, emulate=False)[0]))
    idx = rng.choice(mlp.NPARAM, size=200, replace=False)
    eps = 1e-6
    fd = np.array([(loss(p + eps * np.eye(1, mlp.NPARAM, i)[0]) -
                    loss(p - eps * np.eye(1, mlp.NPARAM, i)[0])) / (2 * eps) for i in idx])
    assert np.allclose(g[idx], fd, rtol=1e-6, atol=1e-8)


def test_degenerate_networks():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(7, mlp.IN))
    d = mlp.unflatten(np.zeros(mlp.NPARAM))
    d = {k: v.copy() for k, v in d.items()}
    d["b4"][:] = rng.normal(size=mlp.OUT)
    p = mlp.flatten(d)
    y, _ = mlp.forward(p, x, emulate=False)
    assert np.allclose(y, d["b4"][None, :])
    dy = rng.normal(size=(7, mlp.OUT))
    g = mlp.unflatten(mlp.backward(p, x, dy, emulate=False))
    assert np.allclose(g["b4"], dy.sum(axis=0))
    for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4"):
        assert np.all(g[k] == 0.0), k      # hidden units are tanh(0) = 0


def test_emulated_forward_is_close_to_exact():
    rng = np.random.default_rng(4)
    p = mlp.glorot(5)
    x = rng.normal(size=(64, mlp.IN))
    ye, _ = mlp.forward(p, x, emulate=True)
    yx, _ = mlp.forward(p, x, emulate=False)
    rel = np.linalg.norm(ye - yx) / np.linalg.norm(yx)
    assert rel < 2e-2
