"""Synthetic disc + halo workloads (tables and particles) for tests and benchmarks.

disc: exponential R_d = 1, sech^2 z_d = 0.1, R < 10, total mass 1.
halo: Hernquist a = 5, truncated r < 100, total mass 10.
EOF basis of the disc: mmax = 6, norder = 18 on a 256 x 128 (xi, eta) grid
(cmap 1, ASCALE 1, HSCALE 0.1, RMIN 0.001, RMAX 20 so R_table = 14.14).
SL basis of the halo: lmax = 6, nmax = 20, numr = 2000 (cmap 1, scale 5,
r in [1e-3, 200]), the exact SL basis of the Hernquist background.
"""
from __future__ import annotations

import numpy as np

from ..basis import tablegen as tg
from . import samplers as S

DISC = dict(rd=1.0, zd=0.1, rmax=10.0, mass=1.0)
HALO = dict(a=5.0, rtrunc=100.0, mass=10.0)
EOF_BASIS = dict(mmax=6, norder=18, numx=256, numy=128, rmin=0.001, rmax=20.0, ascale=1.0,
                 hscale=0.1, cmap=1)
SL_BASIS = dict(lmax=6, nmax=20, numr=2000, rmin=1e-3, rmax=200.0, cmap=1, scale=5.0)


def disc_halo_tables(seed=1, n_eof_sample=60000, lmax_s=24, nmax_s=16, eof=None, sl=None):
    """(sl_tables, eof_tables) dicts for BFESystem; keyword dicts override the basis sizes."""
    eb = dict(EOF_BASIS, **(eof or {}))
    sb = dict(SL_BASIS, **(sl or {}))
    d = S.exponential_disc(n_eof_sample, seed=seed, **DISC)
    et = tg.make_eof_tables(d, lmax_s=lmax_s, nmax_s=nmax_s, **eb)
    st = tg.make_sl_tables(sb["lmax"], sb["nmax"], sb["numr"], sb["rmin"], sb["rmax"], sb["cmap"],
                           sb["scale"], model_mass=HALO["mass"])
    return st, et


def disc_halo_particles(n_disc, n_halo, seed=0):
    """((x, y, z, m) disc, (x, y, z, m) halo) float64 arrays."""
    return (S.exponential_disc(n_disc, seed=seed, **DISC),
            S.hernquist_sphere(n_halo, seed=seed + 1000, **HALO))


def batch_snapshots(nsnap=64, lo=200_000, hi=4_000_000, seed=0, halo_ratio=1.0, device="cuda"):
    """Batched-snapshot workload, sampled on the device: nsnap disc + halo
    snapshots with disc counts log-uniform in [lo, hi] (numpy PCG64, seed) and
    halo counts round(halo_ratio * count).  Each disc is the DISC model with
    an m=2 perturbation 1 + A cos 2(phi - phi_b), A ~ U[0, 0.3], phi_b ~ U[0, pi);
    each halo the HALO model.  Particle masses are model mass / count.

    Returns dict(disc=(x, y, z, m), halo=(x, y, z, m)) float32 device tensors
    (snapshots concatenated), disc_counts, halo_counts (int64 numpy), amp,
    phib (numpy).  Sampling uses a seeded torch generator on the device; the
    inverse CDFs are those of exptool_amd.utils.samplers."""
    import torch
    g = S.rng(seed)
    dcounts = S.log_uniform_counts(nsnap, lo, hi, seed=seed)
    hcounts = np.round(halo_ratio * dcounts).astype(np.int64)
    amp = g.uniform(0.0, 0.3, nsnap)
    phib = g.uniform(0.0, np.pi, nsnap)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed + 17)
    f64 = dict(dtype=torch.float64, device=device)

    def u(n):
        return torch.rand(n, generator=gen, **f64)

    nd, nh = int(dcounts.sum()), int(hcounts.sum())
    # disc: R by Newton on the truncated exponential CDF, sech^2 z, barred phi
    rmax, rd, zd = DISC["rmax"] / DISC["rd"], DISC["rd"], DISC["zd"]
    norm = 1.0 - (1.0 + rmax) * np.exp(-rmax)
    target = u(nd) * norm
    R = (-torch.log1p(-target.clamp(0, 1 - 1e-16)) + 1.0).clamp(1e-6, rmax)
    for _ in range(60):
        F = 1.0 - (1.0 + R) * torch.exp(-R) - target
        R = (R - F / (R * torch.exp(-R)).clamp_min(1e-300)).clamp(0.0, rmax)
    R = R * rd
    z = zd * torch.atanh((2.0 * u(nd) - 1.0).clamp(-1 + 1e-15, 1 - 1e-15))
    snap = torch.repeat_interleave(torch.arange(nsnap, device=device),
                                   torch.as_tensor(dcounts, device=device))
    A = torch.as_tensor(amp, **f64)[snap]
    pb = torch.as_tensor(phib, **f64)[snap]
    tgt = 2.0 * np.pi * u(nd)
    psi = tgt.clone()
    for _ in range(50):
        psi = (psi - (psi + 0.5 * A * torch.sin(2.0 * psi) - tgt) / (1.0 + A * torch.cos(2.0 * psi))
               ).clamp(0.0, 2.0 * np.pi)
    ph = torch.remainder(psi + pb, 2.0 * np.pi)
    dm = (DISC["mass"] / torch.as_tensor(dcounts, **f64))[snap]
    disc = tuple(v.float().contiguous() for v in (R * torch.cos(ph), R * torch.sin(ph), z, dm))
    del R, z, A, pb, tgt, psi, ph, dm, snap
    # halo: Hernquist inverse CDF, isotropic
    a, rt = HALO["a"], HALO["rtrunc"]
    s = torch.sqrt(u(nh) * (rt / (rt + a)) ** 2)
    r = a * s / (1.0 - s)
    cth = 2.0 * u(nh) - 1.0
    ph = 2.0 * np.pi * u(nh)
    sth = torch.sqrt(1.0 - cth * cth)
    hsnap = torch.repeat_interleave(torch.arange(nsnap, device=device),
                                    torch.as_tensor(hcounts, device=device))
    hm = (HALO["mass"] / torch.as_tensor(np.maximum(hcounts, 1), **f64))[hsnap]
    halo = tuple(v.float().contiguous() for v in (r * sth * torch.cos(ph), r * sth * torch.sin(ph),
                                                  r * cth, hm))
    return dict(disc=disc, halo=halo, disc_counts=dcounts, halo_counts=hcounts, amp=amp, phib=phib)
