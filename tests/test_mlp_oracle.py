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


def test_base_field_closed_forms():
    """The analytic base: forces are -grad Phi (central differences), a
    Miyamoto-Nagai disc with a = 0 is a Plummer sphere, and the Hernquist
    sphere's force is -M / (r + a)^2."""
    g = np.random.default_rng(3)
    base = np.array([2, 1.3, 0.6, 0.2, -0.4, 1.7, 0.3, 9.0, 4.0, -0.05])
    x, y, z = g.normal(0, 3, 50), g.normal(0, 3, 50), g.normal(0, 0.5, 50)
    out = M.base_field(base, x, y, z)
    h = 1e-6
    R = np.hypot(x, y)
    pot = lambda R_, z_: M.base_field(base, R_, 0 * R_, z_)[0]
    dR = (pot(R + h, z) - pot(R - h, z)) / (2 * h)
    dz = (pot(R, z + h) - pot(R, z - h)) / (2 * h)
    assert np.allclose(out[1], -dR, rtol=1e-6, atol=1e-9)
    assert np.allclose(out[3], -dz, rtol=1e-6, atol=1e-9)
    assert np.all(out[2] == 0)
    r = np.sqrt(x * x + y * y + z * z)
    pl = M.base_field(np.array([1, 2.0, 0.0, 0.7, 0.0, 1.0, 0.0]), x, y, z)
    assert np.allclose(pl[0], -2.0 / np.sqrt(r * r + 0.49), rtol=1e-14)
    he = M.base_field(np.array([0, 3.0, 2.0, 0.0]), x, y, z)
    assert np.allclose(he[0], -3.0 / (r + 2.0), rtol=1e-14)
    assert np.allclose(np.hypot(he[1], he[3]), 3.0 / (r + 2.0) ** 2, rtol=1e-12)


def test_forward_grad_is_the_potential_gradient():
    """forward_grad (no bf16 rounding): the tangents are the derivatives of the
    potential output along the features (central differences), agree with
    torch autograd, and conservative() turns them into F = -grad Phi."""
    torch = pytest.importorskip("torch")
    g = np.random.default_rng(8)
    p = M.init_params(9) * 0.6
    cfg = np.array([12.0, 0.1, np.arcsinh(30.0), 1.7, 1, 1, 1, -0.4, 0, 0, 0])
    n = 40
    x, y, z = g.normal(0, 3, n), g.normal(0, 3, n), g.normal(0, 0.4, n)
    y0, dy = M.forward_grad(p, cfg, x, y, z, bf=False)
    Y, _ = M.forward(p, cfg, x, y, z, bf=False)
    assert np.allclose(y0, Y[0], rtol=1e-13, atol=1e-13)
    P, out = _torch_net(p, cfg, x, y, z)
    f = torch.stack([torch.tensor(x) / cfg[0], torch.tensor(y) / cfg[0],
                     torch.asinh(torch.tensor(z) / cfg[1]) / cfg[2]], dim=1).requires_grad_(True)
    h = f
    for k in range(4):
        h = torch.nn.functional.silu(torch.nn.functional.linear(h, P[f"W{k}"], P[f"b{k}"]))
    o = torch.nn.functional.linear(h, P["W4"], P["b4"])[:, 0]
    ga = torch.autograd.grad(o.sum(), f)[0].detach().numpy().T
    assert np.max(np.abs(dy - ga)) <= 1e-12 * max(1.0, np.max(np.abs(ga)))
    base = np.array([1, 1.0, 0.5, 0.2, 5.0, 3.0, 0.1])
    pot = lambda xx, yy, zz: M.conservative(cfg, M.forward_grad(p, cfg, xx, yy, zz, bf=False)[0],
                                            np.zeros((3, xx.size)), xx, yy, zz, base)[0]
    c = M.conservative(cfg, y0, dy, x, y, z, base)
    h = 1e-6
    gx = (pot(x + h, y, z) - pot(x - h, y, z)) / (2 * h)
    gy = (pot(x, y + h, z) - pot(x, y - h, z)) / (2 * h)
    gz = (pot(x, y, z + h) - pot(x, y, z - h)) / (2 * h)
    R = np.hypot(x, y)
    assert np.allclose(c[1], -(gx * x + gy * y) / R, rtol=1e-6, atol=1e-8)
    assert np.allclose(c[2], -(gy * x - gx * y) / R, rtol=1e-6, atol=1e-8)
    assert np.allclose(c[3], -gz, rtol=1e-6, atol=1e-8)
