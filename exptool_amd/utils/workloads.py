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
