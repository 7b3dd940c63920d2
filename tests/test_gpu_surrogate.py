"""The MLP surrogate on the benchmark's frozen field (disc EOF mmax 6 norder
18 on 256 x 128 and halo SL lmax 6 nmax 20 from 1M + 1M particles), held to
the bounds it is built for, and what the network adds over the method
without it.  The surrogate runs in its conservative mode: the forces are
minus the gradient of the learned potential (plus the analytic base).

* force error against the plain BFE on held-out points of the training
  domain (R < 12, |z| < 3): median <= 1e-3 (p99 and max printed);
* orbits drawn as the benchmark draws them, 10k leapfrog steps of dt = 1e-3
  with the surrogate: median relative energy drift (measured with the BFE
  potential) <= 1e-4;
* the network against its analytic base alone (the same kernels with zero
  network output): it must cut the median force error at least fourfold;
  the surrogate's orbit-steps/s against the plain BFE integrator's is
  measured and printed.

Training as in the benchmark: 16M points uniform in the domain plus 4M
drawn from the disc the orbits start in, 60 000 Adam steps of 65 536.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    import torch
    from exptool_amd.bfe import BFESystem
    from exptool_amd.surrogate.mlp import MLPField
    from exptool_amd.utils import samplers as S
    from exptool_amd.utils import workloads as W
    print("surrogate: workload field and training", flush=True)
    sl_t, eof_t = W.disc_halo_tables(seed=1)
    s = BFESystem(sl=sl_t, eof=eof_t)
    d, h = W.disc_halo_particles(1_000_000, 1_000_000, seed=0)
    c, sn, hc = s.accumulate_disc_halo(d, h)
    s.set_field(c, sn, hc)
    m = MLPField(s, seed=0, conservative=True)
    ex, ey, ez, _ = S.exponential_disc(4_000_000, seed=123, **W.DISC)
    rep = m.fit(npoints=16_000_000, steps=60000, batch=65536, lr=2e-3, seed=11, extra=(ex, ey, ez))
    print(f"trained {rep['steps']} steps in {rep['train_seconds']:.1f} s, base fit "
          f"{rep['base_seconds']:.1f} s", flush=True)
    ox, oy, oz, _ = S.exponential_disc(20000, seed=5, **W.DISC)
    R = np.sqrt(ox * ox + oy * oy)
    fx = s.field_forces(R, 0 * R, 0 * R)[0].cpu().numpy().astype(np.float64)
    vx, vy, vz = S.circular_orbits(ox, oy, oz, fx, 0.1, seed=6)
    st0 = [np.asarray(v, np.float32) for v in (ox, oy, oz, vx, vy, vz)]
    torch.cuda.synchronize()
    return s, m, st0


def test_surrogate_force_error_bound(trained):
    from exptool_amd.surrogate.mlp import sample_domain
    s, m, _ = trained
    x, y, z = sample_domain(500_000, m.rmax, m.zmax, seed=99, device=s.device)
    ferr, perr = m.force_error(x, y, z)
    med, p99, mx = np.median(ferr), np.quantile(ferr, 0.99), np.max(ferr)
    print(f"held-out force error median {med:.2e} p99 {p99:.2e} max {mx:.2e}; "
          f"potential median {np.median(perr):.2e}", flush=True)
    assert med <= 1e-3


def test_surrogate_energy_drift_bound(trained):
    import torch
    s, m, st0 = trained
    st = [torch.tensor(v, device="cuda") for v in st0]
    out = m.leapfrog(st, dt=1e-3, nsteps=10_000, energy_every=100, check_period=100)
    dr = out["drift"].cpu().numpy()
    ref = [torch.tensor(v, device="cuda") for v in st0]
    bd = s.leapfrog(ref, dt=1e-3, nsteps=10_000, energy_every=100)["drift"].cpu().numpy()
    me = out["maxerr"].cpu().numpy()
    ck = out["checked"].cpu().numpy() > 0
    print(f"energy drift median {np.median(dr):.2e} p99 {np.quantile(dr, 0.99):.2e} (plain BFE "
          f"median {np.median(bd):.2e}); in-flight checked force error median "
          f"{np.median(me[ck]):.2e} max {np.max(me[ck]):.2e}", flush=True)
    assert np.all(np.isfinite(dr))
    assert np.median(dr) <= 1e-4


def test_network_gain_over_its_base(trained):
    """What the network adds: accuracy over the analytic base it corrects,
    and the speed against the plain BFE integrator (printed)."""
    import torch
    from exptool_amd.surrogate.mlp import MLPField, sample_domain
    s, m, st0 = trained
    x, y, z = sample_domain(300_000, m.rmax, m.zmax, seed=98, device=s.device)
    e_net, _ = m.force_error(x, y, z)
    b = MLPField(s, seed=0, conservative=True)
    b.params = np.zeros_like(b.params)        # zero network: the base alone
    b.set_base(m.base)
    b.set_scaling([1, 1, 1, 1], [0, 0, 0, 0])
    b.upload()
    e_base, _ = b.force_error(x, y, z)
    print(f"median force error: base alone {np.median(e_base):.2e}, base + network "
          f"{np.median(e_net):.2e}", flush=True)
    assert np.median(e_net) * 4.0 <= np.median(e_base)
    m.set_mode(True)                          # the trained network back in the context
    m.set_base(m.base)
    m.upload()
    rates = {}
    for name, run in (("surrogate", lambda st: m.leapfrog(st, dt=1e-3, nsteps=2000, energy_every=0,
                                                          check_period=0)),
                      ("bfe", lambda st: s.leapfrog(st, dt=1e-3, nsteps=2000, energy_every=0))):
        st = [torch.tensor(np.tile(v, 50), device="cuda") for v in st0]    # 1M orbits
        run([t[:4096].clone() for t in st])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(st)
        torch.cuda.synchronize()
        rates[name] = st[0].numel() * 2000 / (time.perf_counter() - t0)
    print(f"orbit-steps/s: surrogate {rates['surrogate']:.3e}, plain BFE {rates['bfe']:.3e} "
          f"({rates['surrogate'] / rates['bfe']:.2f}x)", flush=True)
    assert rates["surrogate"] > 0 and rates["bfe"] > 0
