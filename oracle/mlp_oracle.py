"""FP64 CPU oracle for the MLP neural field (no reference counterpart).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the GPU forward pass and
gradients (libexptool_bfe.so bfe_mlp_forward / bfe_mlp_gradient).

Network: features f = (x/sxy, y/sxy, asinh(z/hz)/wz) -> 128 -> 128 -> 128 ->
128 -> 4 with SiLU, parameter vector [W0(128x3), b0, W1(128x128), b1, W2, b2,
W3, b3, W4(4x128), b4].  Loss L = sum_p w_p sum_k (y_pk - t_pk)^2 / n.

With bf16=True the values the GPU rounds to bfloat16 are rounded here too
(round-to-nearest-even): W1..W3, the inputs H1..H3 of the three 128 x 128
layers, the backward operands gZ2..gZ4 of those layers, and the operands of
the weight-gradient products.  Everything else is float64.  The backward pass
is pinned by central finite differences in float64 (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

H = 128
NPARAMS = 50564
OFF = {}
_o = 0
for name, size in (("W0", H * 3), ("b0", H), ("W1", H * H), ("b1", H), ("W2", H * H), ("b2", H),
                   ("W3", H * H), ("b3", H), ("W4", 4 * H), ("b4", 4)):
    OFF[name] = (_o, size)
    _o += size
assert _o == NPARAMS


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = lambda k: p[OFF[k][0]:OFF[k][0] + OFF[k][1]]
    return dict(W0=g("W0").reshape(H, 3), b0=g("b0"), W1=g("W1").reshape(H, H), b1=g("b1"),
                W2=g("W2").reshape(H, H), b2=g("b2"), W3=g("W3").reshape(H, H), b3=g("b3"),
                W4=g("W4").reshape(4, H), b4=g("b4"))


def pack(d):
    return np.concatenate([np.asarray(d[k], dtype=np.float64).ravel() for k in
                           ("W0", "b0", "W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")])


def bf16(a):
    """Round float64 values to bfloat16 (through float32, nearest-even)."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def features(cfg, x, y, z):
    sxy, hz, wz = cfg[0], cfg[1], cfg[2]
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    return np.stack([x / sxy, y / sxy, np.arcsinh(z / hz) / wz])     # (3, n)


def forward(params, cfg, x, y, z, bf=True):
    """Normalised outputs y (4, n) and the intermediates."""
    P = unpack(params)
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    f = features(cfg, x, y, z)
    Z1 = P["W0"] @ f + P["b0"][:, None]
    H1 = silu(Z1)
    Z2 = q(P["W1"]) @ q(H1) + P["b1"][:, None]
    H2 = silu(Z2)
    Z3 = q(P["W2"]) @ q(H2) + P["b2"][:, None]
    H3 = silu(Z3)
    Z4 = q(P["W3"]) @ q(H3) + P["b3"][:, None]
    H4 = silu(Z4)
    Y = P["W4"] @ H4 + P["b4"][:, None]
    return Y, dict(f=f, Z1=Z1, Z2=Z2, Z3=Z3, Z4=Z4, H1=H1, H2=H2, H3=H3, H4=H4)


def base_field(base, x, y, z):
    """Analytic base the network's outputs are added to: base = [nmn,
    (M, a, b) x nmn, Mh, ah, phi0]; Miyamoto-Nagai discs
    Phi = -M / sqrt(R^2 + (a + sqrt(z^2 + b^2))^2), a Hernquist sphere
    Phi = -Mh / (r + ah), a constant.  Returns (4, n): Phi, F_R, 0, F_z
    with F = -grad Phi (closed forms)."""
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    R = np.sqrt(x * x + y * y)
    out = np.zeros((4, R.size))
    if base is None:
        return out
    b = np.asarray(base, dtype=np.float64)
    nmn = int(b[0])
    for k in range(nmn):
        M, a, bb = b[1 + 3 * k:4 + 3 * k]
        B = np.sqrt(z * z + bb * bb)
        D = np.sqrt(R * R + (a + B) ** 2)
        out[0] -= M / D
        out[1] -= M * R / D ** 3
        out[3] -= M * z * (a + B) / (B * D ** 3)
    mh, ah, phi0 = b[1 + 3 * nmn:4 + 3 * nmn]
    r = np.sqrt(R * R + z * z)
    out[0] += phi0
    if mh != 0.0:
        out[0] -= mh / (r + ah)
        with np.errstate(invalid="ignore", divide="ignore"):
            g = np.where(r > 0, mh / (r * (r + ah) ** 2), 0.0)
        out[1] -= g * R
        out[3] -= g * z
    return out


def forward_grad(params, cfg, x, y, z, bf=True):
    """Potential output y0 (n,) and its derivatives along the three features
    dy0/df (3, n) by forward-mode tangents, rounded to bf16 where the
    conservative-field kernels round: the activations and tangents fed to
    the 128 x 128 layers, and W1..W3."""
    P = unpack(params)
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    f = features(cfg, x, y, z)
    Z = P["W0"] @ f + P["b0"][:, None]
    s_ = 1.0 / (1.0 + np.exp(-Z))
    H = Z * s_
    D = s_ * (1.0 + Z * (1.0 - s_))
    T = [D * P["W0"][:, k:k + 1] for k in range(3)]
    for Wk, bk in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        W = q(P[Wk])
        Z = W @ q(H) + P[bk][:, None]
        s_ = 1.0 / (1.0 + np.exp(-Z))
        Tn = [W @ q(t) for t in T]
        H = Z * s_
        D = s_ * (1.0 + Z * (1.0 - s_))
        T = [D * t for t in Tn]
    y0 = P["W4"][0] @ H + P["b4"][0]
    dy = np.stack([P["W4"][0] @ t for t in T])
    return y0, dy


def conservative(cfg, y0, dy, x, y, z, base=None):
    """(Phi, F_R, F_phi, F_z) of the conservative field: Phi = os0 y0 + ob0
    + base, F = -grad Phi through the features f = (x/sxy, y/sxy,
    asinh(z/hz)/wz)."""
    sxy, hz, wz = cfg[0], cfg[1], cfg[2]
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    R = np.sqrt(x * x + y * y)
    b = base_field(base, x, y, z)
    gx = cfg[3] * dy[0] / sxy
    gy = cfg[3] * dy[1] / sxy
    gz = cfg[3] * dy[2] / (wz * hz * np.sqrt(1.0 + (z / hz) ** 2))
    with np.errstate(invalid="ignore", divide="ignore"):
        cp = np.where(R > 0, x / R, 1.0)
        sp = np.where(R > 0, y / R, 0.0)
    fx, fy = -gx + b[1] * cp, -gy + b[1] * sp
    return np.stack([cfg[3] * y0 + cfg[7] + b[0], fx * cp + fy * sp, fy * cp - fx * sp,
                     -gz + b[3]])


def physical(cfg, Y, base=None, x=None, y=None, z=None):
    """Physical outputs os * Y + ob, plus the base at (x, y, z) when given."""
    os_ = np.asarray(cfg[3:7], dtype=np.float64)[:, None]
    ob = np.asarray(cfg[7:11], dtype=np.float64)[:, None]
    out = os_ * Y + ob
    if base is not None:
        out = out + base_field(base, x, y, z)
    return out


def gradient(params, cfg, x, y, z, target, weight=None, bf=True, lw=None):
    """(grad vector, n * loss) of L = sum_p w_p sum_k lw_k (y_pk - t_pk)^2 / n
    (lw: per-output loss weights, default 1)."""
    P = unpack(params)
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    Y, I = forward(params, cfg, x, y, z, bf)
    n = Y.shape[1]
    t = np.asarray(target, dtype=np.float64).reshape(4, n)
    w = np.ones(n) if weight is None else np.asarray(weight, dtype=np.float64)
    d = Y - t
    lwv = np.ones(4) if lw is None else np.asarray(lw, dtype=np.float64)
    loss = float(np.sum(lwv[:, None] * w[None, :] * d * d))
    gY = 2.0 * lwv[:, None] * w[None, :] * d / n
    G = {}
    G["W4"] = gY @ I["H4"].T
    G["b4"] = gY.sum(axis=1)
    gZ4 = (P["W4"].T @ gY) * dsilu(I["Z4"])
    gZ3 = (q(P["W3"]).T @ q(gZ4)) * dsilu(I["Z3"])
    gZ2 = (q(P["W2"]).T @ q(gZ3)) * dsilu(I["Z2"])
    gZ1 = (q(P["W1"]).T @ q(gZ2)) * dsilu(I["Z1"])
    G["W3"] = q(gZ4) @ q(I["H3"]).T
    G["b3"] = gZ4.sum(axis=1)
    G["W2"] = q(gZ3) @ q(I["H2"]).T
    G["b2"] = gZ3.sum(axis=1)
    G["W1"] = q(gZ2) @ q(I["H1"]).T
    G["b1"] = gZ2.sum(axis=1)
    G["W0"] = gZ1 @ I["f"].T
    G["b0"] = gZ1.sum(axis=1)
    return pack(G), loss


def init_params(seed=0):
    """He-uniform weights (fan-in scaled), zero biases."""
    g = np.random.Generator(np.random.PCG64(seed))
    d = {}
    for k, (fo, fi) in (("W0", (H, 3)), ("W1", (H, H)), ("W2", (H, H)), ("W3", (H, H)), ("W4", (4, H))):
        lim = np.sqrt(6.0 / fi)
        d[k] = g.uniform(-lim, lim, (fo, fi))
    for k, n in (("b0", H), ("b1", H), ("b2", H), ("b3", H), ("b4", 4)):
        d[k] = np.zeros(n)
    return pack(d)
