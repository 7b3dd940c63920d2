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


def _torch_net(p, cfg, x, y, z):
    """The same network written with torch.nn.functional in float64, parameters
    as leaf tensors for autograd."""
    import torch
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True)
         for k, v in M.unpack(p).items()}
    sxy, hz, wz = (float(c) for c in cfg[:3])
    f = torch.stack([torch.tensor(x) / sxy, torch.tensor(y) / sxy,
                     torch.asinh(torch.tensor(z) / hz) / wz], dim=1)          # (n, 3)
    h = f
    for k in range(4):
        h = torch.nn.functional.silu(torch.nn.functional.linear(h, P[f"W{k}"], P[f"b{k}"]))
    out = torch.nn.functional.linear(h, P["W4"], P["b4"])                     # (n, 4)
    return P, out


def test_forward_and_gradient_against_torch_autograd():
    """Forward and backpropagation of the oracle (no bf16 rounding) against
    torch float64 autograd of an independently written network."""
    torch = pytest.importorskip("torch")
    g = np.random.default_rng(5)
    p = M.init_params(7) * 0.7
    for k in ("b0", "b1", "b2", "b3", "b4"):
        o, s = M.OFF[k]
        p[o:o + s] = g.normal(0, 0.2, s)
    cfg = np.array([12.0, 0.1, np.arcsinh(30.0), 1, 1, 1, 1, 0, 0, 0, 0])
    n = 33
    x, y, z = g.normal(0, 3, n), g.normal(0, 3, n), g.normal(0, 0.5, n)
    t = g.normal(size=(4, n))
    w = g.uniform(0.5, 2.0, n)
    Y, _ = M.forward(p, cfg, x, y, z, bf=False)
    P, out = _torch_net(p, cfg, x, y, z)
    assert np.max(np.abs(Y.T - out.detach().numpy())) < 1e-12 * max(1.0, np.max(np.abs(Y)))
    loss = (torch.tensor(w)[:, None] * (out - torch.tensor(t.T)) ** 2).sum() / n
    loss.backward()
    grad, nloss = M.gradient(p, cfg, x, y, z, t, w, bf=False)
    assert nloss / n == pytest.approx(loss.item(), rel=1e-13)
    ref = M.pack({k: v.grad.numpy() for k, v in P.items()})
    for k, (o, s) in M.OFF.items():
        a, b = grad[o:o + s], ref[o:o + s]
        assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(b)), 1e-30), k


def test_bf16_rounding_against_torch():
    """bf16() against torch's float32 -> bfloat16 conversion (nearest-even)."""
    torch = pytest.importorskip("torch")
    r = np.random.default_rng(2)
    v = np.concatenate([r.normal(size=20000) * 10.0 ** r.uniform(-6, 6, 20000),
                        1.0 + np.arange(-64, 64) * 2.0 ** -9])
    ref = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(v), ref)
