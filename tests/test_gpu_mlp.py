"""GPU MLP kernels (bfe_mlp_forward / bfe_mlp_gradient / bfe_mlp_train_step)
against the FP64 MLP oracle with the same bfloat16 rounding points.

Tolerances: outputs max|d| / max|ref| <= 5e-3; gradient blocks
|g - g_ref| / |g_ref| <= 2e-2 (bf16 operands: a value near a rounding
boundary may round the other way when its float32 and float64 values differ).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def field():
    import torch
    from exptool_amd.bfe import BFESystem
    from exptool_amd.utils import workloads as W
    sl, eof = W.disc_halo_tables(seed=1, n_eof_sample=20000, lmax_s=12, nmax_s=10,
                                 eof=dict(mmax=4, norder=8, numx=128, numy=64),
                                 sl=dict(lmax=4, nmax=10, numr=1000))
    s = BFESystem(sl=sl, eof=eof)
    (d, h) = W.disc_halo_particles(200000, 200000, seed=3)
    c, sn = s.eof_accumulate(*d)
    hc = s.sl_accumulate(*h)
    s.set_field(c, sn, hc)
    torch.cuda.synchronize()
    return s


def _pts(n, seed):
    g = np.random.default_rng(seed)
    R = 12 * np.sqrt(g.random(n))
    ph = 2 * np.pi * g.random(n)
    z = g.uniform(-3, 3, n)
    return (R * np.cos(ph)).astype(np.float32), (R * np.sin(ph)).astype(np.float32), z.astype(np.float32)


def test_forward_matches_oracle(field):
    from exptool_amd.surrogate.mlp import MLPField
    m = MLPField(field, seed=4)
    m.set_scaling([2.0, 0.5, 0.1, 0.3], [-1.0, 0.0, 0.0, 0.0])
    m.upload()
    x, y, z = _pts(1000, 1)     # not a multiple of the 128-point tile
    got = m.forward(x, y, z).cpu().numpy().astype(np.float64)
    Y, _ = MO.forward(m.params, m.cfg, x, y, z, bf=True)
    ref = MO.physical(m.cfg, Y)
    for k in range(4):
        assert np.max(np.abs(got[k] - ref[k])) / np.max(np.abs(ref[k])) <= 5e-3, k


def test_forward_adds_analytic_base(field):
    """Outputs are os y + ob + base: with a zero network the forward pass is
    the analytic base (float64 in the kernel, float32 out); with a network,
    the oracle's network outputs plus the same base."""
    from exptool_amd.surrogate.mlp import MLPField
    base = np.array([2, 1.3, 0.6, 0.2, -0.4, 1.7, 0.3, 9.0, 4.0, -0.05])
    x, y, z = _pts(3000, 7)
    m = MLPField(field, seed=4)
    m.set_base(base)
    m.params = np.zeros_like(m.params)
    m.set_scaling([1, 1, 1, 1], [0, 0, 0, 0])
    m.upload()
    got = m.forward(x, y, z).cpu().numpy().astype(np.float64)
    ref = MO.base_field(base, x, y, z)
    for k in (0, 1, 3):
        assert np.max(np.abs(got[k] - ref[k]) / np.maximum(np.abs(ref[k]), 1e-3)) <= 1e-6, k
    assert np.all(got[2] == 0)
    m2 = MLPField(field, seed=4)
    m2.set_base(base)
    m2.set_scaling([2.0, 0.5, 0.1, 0.3], [-1.0, 0.0, 0.0, 0.0])
    m2.upload()
    got2 = m2.forward(x, y, z).cpu().numpy().astype(np.float64) - ref
    Y, _ = MO.forward(m2.params, m2.cfg, x, y, z, bf=True)
    net = MO.physical(m2.cfg, Y)
    for k in range(4):
        assert np.max(np.abs(got2[k] - net[k])) / np.max(np.abs(net[k])) <= 5e-3, k
    m2.set_base(None)


def test_conservative_forward_matches_oracle(field):
    """The conservative field (forces = -grad of the potential output, by
    forward-mode tangents on the bf16 MFMA layers) against the oracle's
    forward_grad with the same bf16 rounding points: potential and forces
    max|d| / max|ref| <= 1e-2 (the tangents are bf16 operands too)."""
    from exptool_amd.surrogate.mlp import MLPField
    base = np.array([2, 1.3, 0.6, 0.2, -0.4, 1.7, 0.3, 9.0, 4.0, -0.05])
    m = MLPField(field, seed=4, conservative=True)
    m.set_base(base)
    m.set_scaling([0.3, 1, 1, 1], [-1.0, 0.0, 0.0, 0.0])
    m.upload()
    x, y, z = _pts(1000, 11)
    got = m.forward(x, y, z).cpu().numpy().astype(np.float64)
    y0, dy = MO.forward_grad(m.params, m.cfg, x, y, z, bf=True)
    ref = MO.conservative(m.cfg, y0, dy, x, y, z, base)
    nb = MO.base_field(base, x, y, z)
    for k in range(4):
        scale = np.max(np.abs(ref[k] - nb[k])) if k else np.max(np.abs(ref[k]))
        assert np.max(np.abs(got[k] - ref[k])) / scale <= 1e-2, k
    m.set_mode(False)
    m.set_base(None)


def test_loss_weights_match_oracle(field):
    import torch
    from exptool_amd.surrogate.mlp import MLPField
    m = MLPField(field, seed=5)
    lw = (1.0, 0.0, 0.5, 2.0)
    m.set_mode(False, lw)
    x, y, z = _pts(555, 12)
    g = np.random.default_rng(13)
    t = g.normal(size=(4, 555)).astype(np.float32)
    gg, loss = m.gradient(x, y, z, torch.tensor(t, device="cuda"))
    ref, rloss = MO.gradient(m.params, m.cfg, x, y, z, t, None, bf=True, lw=lw)
    assert float(loss.item()) == pytest.approx(rloss, rel=2e-3)
    got = gg.cpu().numpy().astype(np.float64)
    for name, (o, s_) in MO.OFF.items():
        a, b = got[o:o + s_], ref[o:o + s_]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 2e-2, name
    m.set_mode(False)


def test_gradient_matches_oracle(field):
    import torch
    from exptool_amd.surrogate.mlp import MLPField
    m = MLPField(field, seed=5)
    x, y, z = _pts(777, 2)
    g = np.random.default_rng(3)
    t = g.normal(size=(4, 777)).astype(np.float32)
    w = g.uniform(0.5, 2, 777).astype(np.float32)
    gg, loss = m.gradient(x, y, z, torch.tensor(t, device="cuda"), torch.tensor(w, device="cuda"))
    got = gg.cpu().numpy().astype(np.float64)
    ref, rloss = MO.gradient(m.params, m.cfg, x, y, z, t, w, bf=True)
    assert float(loss.item()) == pytest.approx(rloss, rel=2e-3)
    for name, (o, s) in MO.OFF.items():
        a, b = got[o:o + s], ref[o:o + s]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 2e-2, name


def test_training_reduces_loss_and_force_error(field):
    from exptool_amd.surrogate.mlp import MLPField
    m = MLPField(field, seed=6)
    x, y, z = _pts(4000, 9)
    e0, _ = m.force_error(x, y, z)       # untrained (scaling set by fit below)
    rep = m.fit(npoints=400_000, steps=600, batch=16384, lr=3e-3, seed=1, log_every=100)
    losses = [l for _, _, l in rep["log"]]
    assert losses[-1] < 0.2 * losses[0], losses
    e1, p1 = m.force_error(x, y, z)
    assert np.median(e1) < 0.1, np.median(e1)
    assert np.median(e1) < 0.5 * np.median(e0)


def test_surrogate_leapfrog(field):
    """Orbits that stay outside the training domain follow the BFE exactly;
    orbits inside are integrated by the surrogate, checked against the BFE
    and keep their energy close to the BFE integrator's."""
    import torch
    from exptool_amd.surrogate.mlp import MLPField
    m = MLPField(field, seed=8)
    m.fit(npoints=400_000, steps=800, batch=16384, lr=3e-3, seed=2)
    g = np.random.default_rng(4)
    n_in, n_out = 3000, 200
    R = np.concatenate([g.uniform(1.0, 8.0, n_in), g.uniform(40.0, 60.0, n_out)])
    ph = g.uniform(0, 2 * np.pi, R.size)
    x, y = R * np.cos(ph), R * np.sin(ph)
    z = np.concatenate([g.normal(0, 0.05, n_in), g.uniform(10.0, 20.0, n_out)])
    fx, fy, fz, _ = field.field_forces(R, 0 * R, 0 * R)
    from exptool_amd.utils.samplers import circular_orbits
    vx, vy, vz = circular_orbits(x, y, z, fx.cpu().numpy().astype(np.float64), 0.05, seed=5)
    st0 = [np.asarray(v, np.float32) for v in (x, y, z, vx, vy, vz)]
    mk = lambda: [torch.tensor(v, device="cuda") for v in st0]
    a = mk()
    ra = m.leapfrog(a, dt=2e-3, nsteps=300, energy_every=100, check_period=10)
    b = mk()
    rb = field.leapfrog(b, dt=2e-3, nsteps=300, energy_every=100)
    torch.cuda.synchronize()
    A = torch.stack(a).cpu().numpy()
    Bv = torch.stack(b).cpu().numpy()
    fb = ra["fallback"].cpu().numpy()
    ck = ra["checked"].cpu().numpy()
    assert np.all(fb[n_in:] == 301)                  # every evaluation of the outer orbits fell back
    assert np.allclose(A[:, n_in:], Bv[:, n_in:], rtol=1e-5, atol=1e-6)  # and matches the BFE integrator
    assert ck[:n_in].sum() > 0
    err = ra["maxerr"].cpu().numpy()[:n_in][ck[:n_in] > 0]
    assert np.all(np.isfinite(err)) and np.median(err) < 0.1
    dr = ra["drift"].cpu().numpy()[:n_in]
    assert np.median(dr) < 1e-2
    assert np.all(np.isfinite(A))
