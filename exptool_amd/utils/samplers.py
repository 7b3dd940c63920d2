"""Seeded particle samplers for synthetic BFE workloads.

The reference has no usable sampler (its DF sampler references undefined
names, exptool/models/distributionfunction.py:237-247), so these are written
from the closed-form cumulative distributions:

* Hernquist sphere, M(<r)/M = r^2/(r+a)^2 (exptool/models/hernquist.py:30-110):
  s = r/(r+a) = sqrt(u f) with f = (rt/(rt+a))^2 samples r < rt.
* Exponential disc, M(<R) = 1 - (1 + R/Rd) exp(-R/Rd), truncated at Rmax,
  inverted by Newton iteration; sech^2 vertical profile z = zd atanh(2u - 1).
* Optional m=2 perturbation of the azimuthal density 1 + A cos 2(phi - phi_b),
  sampled exactly by inverting its CDF with Newton iterations.
All draws use numpy.random.Generator(PCG64(seed)).
"""
from __future__ import annotations

import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def hernquist_sphere(n, a=1.0, rtrunc=20.0, mass=1.0, seed=0):
    """n equal-mass particles of total mass `mass` inside r < rtrunc.
    Returns (x, y, z, m) float64 arrays."""
    g = rng(seed)
    f = (rtrunc / (rtrunc + a)) ** 2
    s = np.sqrt(g.random(n) * f)
    r = a * s / (1.0 - s)
    cth = 2.0 * g.random(n) - 1.0
    ph = 2.0 * np.pi * g.random(n)
    sth = np.sqrt(1.0 - cth * cth)
    m = np.full(n, mass / n)
    return r * sth * np.cos(ph), r * sth * np.sin(ph), r * cth, m


def _invert_expdisc(u, rmax):
    """R with [1-(1+R)e^-R] / [1-(1+rmax)e^-rmax] = u (scale length 1)."""
    norm = 1.0 - (1.0 + rmax) * np.exp(-rmax)
    target = u * norm
    R = np.clip(-np.log1p(-np.clip(target, 0, 1 - 1e-16)) + 1.0, 1e-6, rmax)
    for _ in range(60):
        F = 1.0 - (1.0 + R) * np.exp(-R) - target
        dF = R * np.exp(-R)
        step = F / np.maximum(dF, 1e-300)
        R = np.clip(R - step, 0.0, rmax)
    return R


def _sample_bar_phi(g, n, amp, phib):
    """phi on [0, 2 pi) with density (1 + amp cos 2(phi - phib)) / (2 pi)."""
    u = g.random(n)
    if amp == 0.0:
        return 2.0 * np.pi * u
    # CDF of psi = phi - phib on [0, 2pi): (psi + amp/2 sin 2 psi)/(2 pi)
    target = 2.0 * np.pi * u
    psi = target.copy()
    for _ in range(50):
        F = psi + 0.5 * amp * np.sin(2.0 * psi) - target
        dF = 1.0 + amp * np.cos(2.0 * psi)
        psi = np.clip(psi - F / dF, 0.0, 2.0 * np.pi)
    return np.mod(psi + phib, 2.0 * np.pi)


def exponential_disc(n, rd=1.0, zd=0.1, rmax=10.0, mass=1.0, bar_amp=0.0, bar_angle=0.0, seed=0):
    """Exponential (scale rd, truncated R < rmax) sech^2 (scale zd) disc of total
    mass `mass`; optional m=2 perturbation.  Returns (x, y, z, m)."""
    g = rng(seed)
    R = rd * _invert_expdisc(g.random(n), rmax / rd)
    z = zd * np.arctanh(np.clip(2.0 * g.random(n) - 1.0, -1 + 1e-15, 1 - 1e-15))
    ph = _sample_bar_phi(g, n, bar_amp, bar_angle)
    m = np.full(n, mass / n)
    return R * np.cos(ph), R * np.sin(ph), z, m


def circular_orbits(x, y, z, fr_plane, disp=0.1, seed=0):
    """Velocities v_c(R) phi_hat + N(0, (disp v_c)^2 I3) where
    v_c^2 = -R F_R(R, z=0) (fr_plane: planar radial force at (R, 0)).
    Returns (vx, vy, vz)."""
    g = rng(seed)
    R = np.sqrt(x * x + y * y)
    vc = np.sqrt(np.maximum(-R * fr_plane, 0.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(R > 0, x / R, 1.0)
        s = np.where(R > 0, y / R, 0.0)
    noise = g.standard_normal((3, x.size)) * (disp * vc)[None, :]
    return -vc * s + noise[0], vc * c + noise[1], noise[2]


def log_uniform_counts(nsnap, lo, hi, seed=0):
    """Particle counts log-uniform in [lo, hi]."""
    g = rng(seed)
    return np.round(np.exp(g.uniform(np.log(lo), np.log(hi), nsnap))).astype(np.int64)
