"""Pins of the MLP oracle: backpropagation against float64 central finite
differences, and the bfloat16 rounding emulation against exact cases."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding_exact_cases():
    # values representable in bf16 are unchanged; halfway cases round to even
    assert M.bf16(np.array([1.0, -2.5, 0.15625]))[0] == 1.0
    assert np.all(M.bf16(np.array([1.0, -2.5, 0.15625])) == np.array([1.0, -2.5, 0.15625]))
    one_plus_half_ulp = 1.0 + 2.0 ** -8          # halfway between 1 and 1 + 2^-7
    assert M.bf16(np.array([one_plus_half_ulp]))[0] == 1.0
    assert M.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    r = np.random.default_rng(0).normal(size=1000)
    assert np.max(np.abs(M.bf16(r) - r) / np.abs(r)) <= 2.0 ** -8


def test_gradient_against_finite_differences():
    g = np.random.default_rng(1)
    p = M.init_params(3) * 0.5
    p[M.OFF["b0"][0]:M.OFF["b0"][0] + 128] = g.normal(0, 0.1, 128)
    cfg = np.array([12.0, 0.1, np.arcsinh(30.0), 1, 1, 1, 1, 0, 0, 0, 0])
    n = 7
    x, y, z = g.normal(0, 3, n), g.normal(0, 3, n), g.normal(0, 0.5, n)
    t = g.normal(size=(4, n))
    w = g.uniform(0.5, 2.0, n)
    grad, loss = M.gradient(p, cfg, x, y, z, t, w, bf=False)
    L = lambda q: M.gradient(q, cfg, x, y, z, t, w, bf=False)[1] / n
    idx = []
    for name in ("W0", "b0", "W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"):
        o, s = M.OFF[name]
        idx += list(g.integers(o, o + s, 3))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (L(p + e) - L(p - e)) / (2 * h)
        assert fd == pytest.approx(grad[i], rel=1e-5, abs=1e-9), i
