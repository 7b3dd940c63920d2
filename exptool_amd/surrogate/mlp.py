"""MLP neural field of a frozen BFE field: (x, y, z) -> (Phi, F_R, F_phi, F_z).

The network (3 -> 128 -> 128 -> 128 -> 128 -> 4, SiLU) is trained only on
evaluations of the plain BFE field of a BFESystem, by the library's own
kernels (bfe_mlp_train_step: bf16 MFMA forward/backward, float32 Adam), and
evaluated by bfe_mlp_forward.  In the conservative mode the forces are minus
the gradient of the potential output, taken in the kernels by forward-mode
tangents through the same bf16 MFMA layers, so the surrogate field is exactly
a gradient field and orbits integrated in it conserve its energy; only the
potential output is trained.  F_phi is the physical azimuthal component
-(1/R) dPhi/dphi.  Inputs are mapped to features (x/sxy, y/sxy,
asinh(z/hz)/asinh(zmax/hz)) so that the thin disc is resolved.  The network
learns the field minus an analytic base (Miyamoto-Nagai discs and a Hernquist
sphere fitted to the same BFE evaluations), which the kernels add back in
float64; its outputs are standardised (one scale for the three forces).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from .. import _lib
from .._lib import ptr, stream_ptr

H = 128
NPARAMS = _lib.MLP_NPARAMS


def init_params(seed=0):
    """He-uniform weights (limit sqrt(6 / fan_in)), zero biases, in the
    parameter order W0, b0, W1, b1, W2, b2, W3, b3, W4, b4."""
    g = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for fo, fi in ((H, 3), (H, H), (H, H), (H, H), (4, H)):
        lim = math.sqrt(6.0 / fi)
        parts.append(g.uniform(-lim, lim, fo * fi))
        parts.append(np.zeros(fo))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAMS
    return p


def bfe_labels(system, x, y, z):
    """(4, n) float32 device tensor Phi, F_R, F_phi, F_z from the frozen BFE
    field of `system` (Fields.return_forces_cart semantics)."""
    fx, fy, fz, pot = system.field_forces(x, y, z)
    x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, np.float32), device=system.device)
    y = y if isinstance(y, torch.Tensor) else torch.as_tensor(np.asarray(y, np.float32), device=system.device)
    R = torch.sqrt(x * x + y * y).clamp_min(1e-20)
    fR = (fx * x + fy * y) / R
    fphi = (fy * x - fx * y) / R
    return torch.stack([pot, fR, fphi, fz])


def base_values(base, x, y, z):
    """(4, n) float32 device tensor Phi, F_R, 0, F_z of the analytic base
    (bfe_mlp_set_base layout: nmn, (M, a, b) x nmn, Mh, ah, phi0), computed in
    float64 by the formulas the kernels use."""
    x, y, z = (t.double() for t in (x, y, z))
    R2, z2 = x * x + y * y, z * z
    R = torch.sqrt(R2)
    pot = torch.full_like(R, float(base[-1]))
    fR = torch.zeros_like(R)
    fz = torch.zeros_like(R)
    nmn = int(base[0])
    for k in range(nmn):
        M, a, b = (float(v) for v in base[1 + 3 * k:4 + 3 * k])
        bz = torch.sqrt(z2 + b * b)
        s = a + bz
        idd = 1.0 / torch.sqrt(R2 + s * s)
        m3 = M * idd ** 3
        pot = pot - M * idd
        fR = fR - m3 * R
        fz = fz - m3 * z * s / bz
    mh, ah = float(base[1 + 3 * nmn]), float(base[2 + 3 * nmn])
    if mh != 0.0:
        r = torch.sqrt(R2 + z2)
        ir = 1.0 / (r + ah)
        pot = pot - mh * ir
        g = torch.where(r > 0, mh * ir * ir / r.clamp_min(1e-300), torch.zeros_like(r))
        fR = fR - g * R
        fz = fz - g * z
    return torch.stack([pot, fR, torch.zeros_like(R), fz]).float()


def fit_base(R, z, labels, nmn=3, npts=12000, seed=0):
    """Least-squares fit of the analytic base (nmn Miyamoto-Nagai discs and a
    Hernquist sphere) to BFE evaluations: labels (4, n) Phi, F_R, F_phi, F_z
    at planar radius R and height z (host float64 arrays).  The force
    residuals are taken relative to |F| and the potential's relative to
    |Phi|.  Returns the bfe_mlp_set_base vector."""
    from scipy.optimize import least_squares
    g = np.random.default_rng(seed)
    idx = g.choice(R.size, min(npts, R.size), replace=False)
    R, z = R[idx], z[idx]
    pot, fR, fp, fz = (labels[k][idx] for k in range(4))
    Fn = np.sqrt(fR * fR + fp * fp + fz * fz) + 1e-300

    def unpack(p):
        b = [nmn]
        for k in range(nmn):
            b += [p[3 * k], np.exp(p[3 * k + 1]), np.exp(p[3 * k + 2])]
        return np.array(b + [np.exp(p[3 * nmn]), np.exp(p[3 * nmn + 1]), p[3 * nmn + 2]])

    def model(b):
        P = np.full_like(R, b[-1]); FR = np.zeros_like(R); FZ = np.zeros_like(R)
        for k in range(nmn):
            M, a, bb = b[1 + 3 * k:4 + 3 * k]
            B = np.sqrt(z * z + bb * bb)
            D2 = R * R + (a + B) ** 2
            D = np.sqrt(D2)
            P -= M / D
            FR -= M * R / (D2 * D)
            FZ -= M * z * (a + B) / (B * D2 * D)
        mh, ah = b[1 + 3 * nmn], b[2 + 3 * nmn]
        r = np.sqrt(R * R + z * z)
        P -= mh / (r + ah)
        gg = mh / (np.maximum(r, 1e-300) * (r + ah) ** 2)
        return P, FR - gg * R, FZ - gg * z

    def res(p):
        P, FR, FZ = model(unpack(p))
        return np.concatenate([(FR - fR) / Fn, (FZ - fz) / Fn, 0.3 * (P - pot) / np.abs(pot)])

    # discs of growing scale length, alternating in sign after the first
    # (a sum of Miyamoto-Nagai discs approximates an exponential disc with
    # one negative member), and the halo
    p0 = []
    for k in range(nmn):
        p0 += [(1.0, 0.5, -0.3, 0.2)[k % 4], np.log((0.5, 1.5, 3.0, 6.0)[k % 4]),
               np.log((0.1, 0.1, 0.2, 0.3)[k % 4])]
    p0 += [np.log(10.0), np.log(5.0), 0.0]
    r = least_squares(res, np.array(p0), method="trf", x_scale="jac", max_nfev=400)
    return unpack(r.x)


def sample_domain(n, rmax=12.0, zmax=3.0, seed=0, device="cuda"):
    """n points uniform in the cylinder R < rmax, |z| < zmax (seeded)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    u = torch.rand(3, n, generator=gen, device=device, dtype=torch.float32)
    R = rmax * torch.sqrt(u[0])
    ph = 2.0 * math.pi * u[1]
    return R * torch.cos(ph), R * torch.sin(ph), zmax * (2.0 * u[2] - 1.0)


class MLPField:
    """Network parameters resident in the context of a BFESystem."""

    def __init__(self, system, sxy=12.0, hz=0.1, zmax=3.0, rmax=12.0, seed=0,
                 conservative=False):
        self.sys = system
        self.ctx = system.ctx
        self.rmax, self.zmax = float(rmax), float(zmax)
        self.cfg = np.array([sxy, hz, math.asinh(zmax / hz), 1, 1, 1, 1, 0, 0, 0, 0],
                            dtype=np.float64)
        self.params = init_params(seed)
        self.base = None
        self.set_mode(conservative)
        self.set_base(None)
        self.upload()

    def set_mode(self, conservative, loss_weights=None):
        """conservative: forces are -grad of the network's potential (plus the
        base), evaluated by forward-mode tangents in the kernels; otherwise
        the four outputs are read directly.  loss_weights: per-output training
        weights (default: potential only when conservative, else all 1)."""
        self.conservative = bool(conservative)
        if loss_weights is None:
            loss_weights = (1.0, 0.0, 0.0, 0.0) if self.conservative else (1.0, 1.0, 1.0, 1.0)
        self.loss_weights = np.asarray(loss_weights, dtype=np.float64)
        rc = self.ctx.lib.bfe_mlp_set_mode(self.ctx.h, int(self.conservative),
                                           _lib.host_ptr(self.loss_weights))
        self.ctx.check(rc, "bfe_mlp_set_mode")

    # ---------------------------------------------------------------- state
    def set_base(self, base):
        """Analytic base the network's outputs are added to (None: none)."""
        self.base = None if base is None else np.asarray(base, dtype=np.float64)
        b = np.array([0, 0, 0, 0], dtype=np.float64) if base is None else self.base
        rc = self.ctx.lib.bfe_mlp_set_base(self.ctx.h, _lib.host_ptr(b), int(b.size))
        self.ctx.check(rc, "bfe_mlp_set_base")

    def set_scaling(self, os_, ob):
        self.cfg[3:7] = np.asarray(os_, dtype=np.float64)
        self.cfg[7:11] = np.asarray(ob, dtype=np.float64)

    def upload(self, params=None):
        if params is not None:
            self.params = np.ascontiguousarray(params, dtype=np.float32)
        rc = self.ctx.lib.bfe_mlp_set_params(self.ctx.h, _lib.host_ptr(self.params), NPARAMS,
                                             _lib.host_ptr(self.cfg))
        self.ctx.check(rc, "bfe_mlp_set_params")

    def download(self):
        out = np.empty(NPARAMS, dtype=np.float32)
        rc = self.ctx.lib.bfe_mlp_get_params(self.ctx.h, _lib.host_ptr(out), NPARAMS)
        self.ctx.check(rc, "bfe_mlp_get_params")
        self.params = out
        return out

    # ------------------------------------------------------------- compute
    def forward(self, x, y, z, stream=None):
        """(4, n) float32 device tensor: Phi, F_R, F_phi, F_z."""
        x, y, z, n = self.sys._xyz(x, y, z)
        out = torch.empty((4, n), dtype=torch.float32, device=self.sys.device)
        rc = self.ctx.lib.bfe_mlp_forward(self.ctx.h, ptr(x), ptr(y), ptr(z), n, ptr(out),
                                          stream_ptr(stream))
        self.ctx.check(rc, "bfe_mlp_forward")
        return out

    def normalise(self, labels):
        os_ = torch.as_tensor(self.cfg[3:7], dtype=torch.float32, device=labels.device)[:, None]
        ob = torch.as_tensor(self.cfg[7:11], dtype=torch.float32, device=labels.device)[:, None]
        return ((labels - ob) / os_).contiguous()

    def gradient(self, x, y, z, target, weight=None, stream=None):
        """(grad (NPARAMS,) float32 tensor, n*loss) for normalised targets (4, n)."""
        x, y, z, n = self.sys._xyz(x, y, z)
        target = target.to(torch.float32).contiguous()
        g = torch.empty(NPARAMS, dtype=torch.float32, device=self.sys.device)
        loss = torch.empty(1, dtype=torch.float64, device=self.sys.device)
        w = None if weight is None else weight.to(torch.float32).contiguous()
        rc = self.ctx.lib.bfe_mlp_gradient(self.ctx.h, ptr(x), ptr(y), ptr(z), ptr(target),
                                           ptr(w), n, ptr(g), ptr(loss), stream_ptr(stream))
        self.ctx.check(rc, "bfe_mlp_gradient")
        return g, loss

    def train_step(self, x, y, z, target, weight, lr, betas=(0.9, 0.999), eps=1e-8, loss=None,
                   stream=None):
        n = x.numel()
        rc = self.ctx.lib.bfe_mlp_train_step(self.ctx.h, ptr(x), ptr(y), ptr(z), ptr(target),
                                             ptr(weight), n, C.c_float(lr), C.c_float(betas[0]),
                                             C.c_float(betas[1]), C.c_float(eps), ptr(loss),
                                             stream_ptr(stream))
        self.ctx.check(rc, "bfe_mlp_train_step")

    # --------------------------------------------------------------- training
    def fit(self, npoints=16_000_000, steps=20000, batch=65536, lr=2e-3, lr_min=1e-5,
            seed=0, log_every=0, time_budget=None, base_mn=3, point_weight=None, extra=None):
        """Train on `npoints` BFE evaluations drawn uniformly in R < rmax,
        |z| < zmax.  With base_mn >= 0, an analytic base (base_mn
        Miyamoto-Nagai discs and a Hernquist sphere) is first fitted to a
        subsample of the same evaluations and the network learns the
        remainder (base_mn = None: the network learns the whole field).
        point_weight(x, y, z), when given, multiplies each point's loss
        weight (device tensors in, one weight per point out).  extra: more
        training points (x, y, z) (device tensors, e.g. drawn where the
        orbits to be integrated live), also labelled by the BFE.
        Cosine learning-rate decay.  Returns a training log."""
        dev = self.sys.device
        t0 = time.perf_counter()
        X = []
        T = []
        chunk = 4_000_000
        for k in range(0, npoints, chunk):
            m = min(chunk, npoints - k)
            x, y, z = sample_domain(m, self.rmax, self.zmax, seed=seed + k, device=dev)
            X.append(torch.stack([x, y, z]))
            T.append(bfe_labels(self.sys, x, y, z))
        if extra is not None:
            ex = torch.stack([torch.as_tensor(v, dtype=torch.float32, device=dev).reshape(-1)
                              for v in extra])
            X.append(ex)
            T.append(bfe_labels(self.sys, ex[0], ex[1], ex[2]))
        X = torch.cat(X, dim=1)
        T = torch.cat(T, dim=1)
        npoints = X.shape[1]
        F2 = (T[1] ** 2 + T[2] ** 2 + T[3] ** 2)
        t_base = time.perf_counter()
        if base_mn is not None:
            sub = torch.randint(0, npoints, (20000,), generator=torch.Generator(device=dev).manual_seed(seed + 3),
                                device=dev)
            Xs = X[:, sub].double().cpu().numpy()
            base = fit_base(np.hypot(Xs[0], Xs[1]), Xs[2], T[:, sub].double().cpu().numpy(),
                            nmn=int(base_mn), seed=seed)
            self.set_base(base)
            Res = T - base_values(base, X[0], X[1], X[2])
        else:
            self.set_base(None)
            Res = T
        t_base = time.perf_counter() - t_base
        ob = Res.mean(dim=1).double().cpu().numpy()
        os_ = Res.std(dim=1).double().cpu().numpy()
        ob[1:] = 0.0          # forces: scale only, one scale for the three
        os_[1:] = np.sqrt(np.mean(os_[1:] ** 2))
        self.set_scaling(os_, ob)
        self.upload()
        Tn = self.normalise(Res)
        del Res
        if self.conservative:
            # potential only (the forces are its gradient): equal weights
            Tn[1:] = 0.0
            w = torch.ones_like(F2)
        else:
            # relative-force weighting: w = <|F|^2> / |F|^2 (clipped)
            w = (F2.mean() / F2.clamp_min(1e-30)).clamp(0.05, 20.0)
        if point_weight is not None:
            w = w * point_weight(X[0], X[1], X[2])
        w = (w / w.mean()).contiguous()
        t_data = time.perf_counter() - t0
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed + 7)
        loss = torch.zeros(1, dtype=torch.float64, device=dev)
        log = []
        t1 = time.perf_counter()
        for s in range(steps):
            idx = torch.randint(0, npoints, (batch,), generator=gen, device=dev)
            xb = X[0, idx].contiguous()
            yb = X[1, idx].contiguous()
            zb = X[2, idx].contiguous()
            tb = Tn[:, idx].contiguous()
            wb = w[idx].contiguous()
            lr_s = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * s / max(steps - 1, 1)))
            self.train_step(xb, yb, zb, tb, wb, lr_s, loss=loss)
            if log_every and (s % log_every == 0 or s == steps - 1):
                log.append((s, lr_s, float(loss.item()) / batch))
            if time_budget is not None and (s % 200 == 0) and time.perf_counter() - t1 > time_budget:
                torch.cuda.synchronize()
                log.append((s, lr_s, float(loss.item()) / batch))
                break
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        self.download()
        return dict(log=log, data_seconds=t_data, base_seconds=t_base, train_seconds=t_train,
                    steps=s + 1, batch=batch, samples=(s + 1) * batch, npoints=npoints,
                    base=None if self.base is None else self.base.tolist())

    def force_error(self, x, y, z):
        """Per-point |F_mlp - F_bfe| / |F_bfe| and |Phi_mlp - Phi|/|Phi| (float64 numpy)."""
        ref = bfe_labels(self.sys, x, y, z).double()
        got = self.forward(x, y, z).double()
        dF = torch.sqrt(((got[1:] - ref[1:]) ** 2).sum(0))
        nF = torch.sqrt((ref[1:] ** 2).sum(0))
        return (dF / nF).cpu().numpy(), (torch.abs(got[0] - ref[0]) / torch.abs(ref[0])).cpu().numpy()

    # ------------------------------------------------------------ integration
    def leapfrog(self, state, dt, nsteps, t0=0.0, rotfreq=0.0, energy_every=100,
                 check_period=100, stream=None):
        """Advance state = (x, y, z, vx, vy, vz) (float32 device tensors) in
        place with the surrogate field; points outside R < rmax, |z| < zmax
        use the BFE.  Returns dict(drift (n,) f64, maxerr (n,) f32 largest
        checked force error, fallback (n,) int, checked (n,) int,
        energy (checks,) f64 or None)."""
        x = state[0]
        n = x.numel()
        for t in state:
            if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n:
                raise ValueError("state tensors must be contiguous float32 of equal length")
        dev = self.sys.device
        drift = torch.empty(n, dtype=torch.float64, device=dev)
        maxerr = torch.empty(n, dtype=torch.float32, device=dev)
        counts = torch.empty((n, 2), dtype=torch.int32, device=dev)
        elog = (torch.empty(nsteps // energy_every + 1, dtype=torch.float64, device=dev)
                if energy_every > 0 else None)
        rc = self.ctx.lib.bfe_mlp_leapfrog(self.ctx.h, *[ptr(t) for t in state], n, float(dt),
                                           int(nsteps), float(t0), float(rotfreq),
                                           int(energy_every), C.c_float(self.rmax),
                                           C.c_float(self.zmax), int(check_period), ptr(drift),
                                           ptr(maxerr), ptr(counts), ptr(elog),
                                           stream_ptr(stream))
        self.ctx.check(rc, "bfe_mlp_leapfrog")
        return dict(drift=drift, maxerr=maxerr, fallback=counts[:, 0], checked=counts[:, 1],
                    energy=elog)
