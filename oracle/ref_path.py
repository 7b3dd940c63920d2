"""The reference's own CPU decomposition of the hot path, restated for timing.

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it on the host
cores as the "reference CPU path" beside the GPU.  It computes what the NumPy
oracle computes (tests/ check that), but with the reference's granularity,
so that its cost per particle is the reference's:

  * a process pool of `ncores` workers over contiguous particle blocks
    (spheresl.py:383-431 redistribute_particles / multi_compute_coefficients,
    eof.py:1333-1411), tables loaded once per worker;
  * SL accumulate: a Python loop over particles, per particle the Legendre
    table by recurrences, the radial matrix of the particle's cell, and the
    l, m loop of row updates over n (spheresl.py:567-656);
  * EOF accumulate: the vectorised form over the block (eof.py:492-640);
  * field evaluation: per point, eof.force_eval with its (m, n) slices of the
    six tables at the four cell corners (eof.py:756-872) and
    spheresl.force_eval with get_halo_dens_pot_force's (l, n) loop and the
    l, m loop (spheresl.py:106-160, :1107-1234), combined as
    Fields.return_forces_cart (potential.py:445-497).

The cost model is the point: the loops are those of the reference, scalar
where the reference is scalar and vectorised where it is vectorised.
"""
from __future__ import annotations

import math
import multiprocessing as mp
import os
import tempfile
import time

import numpy as np

from . import bfe_oracle as O

FOURPI = 4.0 * math.pi


# ------------------------------------------------------------ scalar kernels
def _legendre(lmax, x, deriv=False):
    """P_l^m(x) (and dP/dx with x clamped to +-(1 - 1e-8)) as (lmax+1)^2
    lists by recurrences, one point (spheresl.py:664-770)."""
    W = lmax + 1
    p = [[0.0] * W for _ in range(W)]
    p[0][0] = 1.0
    if lmax > 0:
        somx2 = math.sqrt(max((1.0 - x) * (1.0 + x), 0.0))
        fact, pll = 1.0, 1.0
        for m in range(1, W):
            pll *= -fact * somx2
            p[m][m] = pll
            fact += 2.0
    for m in range(lmax):
        pl2 = p[m][m]
        pl1 = x * (2 * m + 1) * pl2
        p[m + 1][m] = pl1
        for l in range(m + 2, W):
            v = (x * (2 * l - 1) * pl1 - (l + m - 1) * pl2) / (l - m)
            p[l][m] = v
            pl2, pl1 = pl1, v
    if not deriv:
        return p, None
    xc = x
    if 1.0 - abs(xc) < 1e-8:
        xc = 1.0 - 1e-8 if xc > 0 else -(1.0 - 1e-8)
    s = 1.0 / (xc * xc - 1.0)
    dp = [[0.0] * W for _ in range(W)]
    for l in range(1, W):
        for m in range(l):
            dp[l][m] = s * (xc * l * p[l][m] - (l + m) * p[l - 1][m])
        dp[l][l] = s * xc * l * p[l][l]
    return p, dp


def _sl_cell(t, r):
    x = float(O.r_to_xi(r, t.cmap, t.scale))
    if t.cmap == 1:
        x = -1.0 if x < -1.0 else (1.0 - 1e-8 if x >= 1.0 else x)
    elif t.cmap == 2:
        x = min(max(x, t.xi[0]), t.xi[-1])
    dxi = t.xi[1] - t.xi[0]
    i = int(math.floor((x - np.min(t.xi)) / dxi))
    i = 0 if i < 0 else (t.numr - 2 if i > t.numr - 2 else i)
    return x, i, (t.xi[i + 1] - x) / dxi, (x - t.xi[i]) / dxi, dxi


def sl_accumulate_loop(t, x, y, z, mass, fac, sev):
    """compute_coefficients_solitary's particle loop (spheresl.py:598-654)."""
    out = np.zeros(((t.lmax + 1) ** 2, t.nmax))
    for p in range(len(x)):
        r = max(math.sqrt(x[p] * x[p] + y[p] * y[p] + z[p] * z[p]), 1e-10)
        costh, phi = z[p] / r, math.atan2(y[p], x[p])
        legs, _ = _legendre(t.lmax, costh)
        _, i, x1, x2, _ = _sl_cell(t, r)
        potd = (x1 * t.eftable[:, :, i] + x2 * t.eftable[:, :, i + 1]) / sev * (
            x1 * t.p0[i] + x2 * t.p0[i + 1])
        for l in range(t.lmax + 1):
            for m in range(l + 1):
                fac4 = potd[l] * (fac[l, m] * legs[l][m] * -FOURPI)
                if m == 0:
                    out[l * l] += fac4 * mass[p]
                else:
                    out[l * l + 2 * m - 1] += math.cos(phi * m) * fac4 * mass[p]
                    out[l * l + 2 * m] += math.sin(phi * m) * fac4 * mass[p]
    return out


def _sl_force_point(t, r, costh, phi, coef, fac, sev):
    """spheresl.force_eval at one point with get_halo_dens_pot_force's (l, n)
    loop (quadratic stencil taken at the particle in the first cell)."""
    xv, i, x1, x2, dxi = _sl_cell(t, r)
    dfac = float(O.d_xi_to_r(xv, t.cmap, t.scale)) / dxi
    c = 1 if i == 0 else i
    ps = x2 - 1.0 if i == 0 else x2
    L, N = t.lmax + 1, t.nmax
    pot = np.zeros((L, N))
    dpot = np.zeros((L, N))
    for l in range(L):
        for n in range(N):
            ef = t.eftable[l, n]
            pot[l][n] = (x1 * ef[i] + x2 * ef[i + 1]) / sev[l, n] * (x1 * t.p0[i] + x2 * t.p0[i + 1])
            dpot[l][n] = dfac * ((ps - 0.5) * ef[c - 1] * t.p0[c - 1] - 2.0 * ps * ef[c] * t.p0[c]
                                 + (ps + 0.5) * ef[c + 1] * t.p0[c + 1]) / sev[l, n]
    legs, dlegs = _legendre(t.lmax, costh, True)
    pot0 = np.sum(fac[0, 0] * coef[0] * pot[0])
    potr = np.sum(fac[0, 0] * coef[0] * dpot[0])
    pot1 = pott = potp = 0.0
    for l in range(1, L):
        for m in range(l + 1):
            f = fac[l, m]
            if m == 0:
                a = coef[l * l]
                pot1 += np.sum(f * legs[l][0] * a * pot[l])
                potr += np.sum(f * legs[l][0] * a * dpot[l])
                pott += np.sum(f * dlegs[l][0] * a * pot[l])
            else:
                a, b = coef[l * l + 2 * m - 1], coef[l * l + 2 * m]
                cm, sm = math.cos(phi * m), math.sin(phi * m)
                pot1 += np.sum(f * legs[l][m] * (a * pot[l] * cm + b * pot[l] * sm))
                potr += np.sum(f * legs[l][m] * (a * dpot[l] * cm + b * dpot[l] * sm))
                pott += np.sum(f * dlegs[l][m] * (a * pot[l] * cm + b * pot[l] * sm))
                potp += np.sum(f * legs[l][m] * m * (-a * pot[l] * sm + b * pot[l] * cm))
    return potr, pott, potp, pot1, pot0


def _eof_force_point(e, R, z, phi, ac, as_):
    """eof.force_eval at one point: (m, n) slices of the tables at the cell
    corners, summed with the cos/sin(m phi) factors (eof.py:791-870)."""
    X, Y, ix, iy = O.return_bins(R, z, e.XMIN, e.dX, e.YMIN, e.dY, e.numx, e.numy, e.ascale,
                                 e.hscale, e.cmap)
    X, Y, ix, iy = float(X[0]), float(Y[0]), int(ix[0]), int(iy[0])
    c00 = (ix + 1.0 - X) * (iy + 1.0 - Y)
    c10 = (X - ix) * (iy + 1.0 - Y)
    c01 = (ix + 1.0 - X) * (Y - iy)
    c11 = (X - ix) * (Y - iy)
    mo = np.arange(1.0, e.mmax + 1.0)[:, None]
    ccos, ssin = np.cos(phi * mo), np.sin(phi * mo)
    g = lambda T, s: (T[s, :, ix, iy] * c00 + T[s, :, ix + 1, iy] * c10 + T[s, :, ix, iy + 1] * c01
                      + T[s, :, ix + 1, iy + 1] * c11)
    hi = slice(1, None)
    fc = ac[1:] * ccos
    p = np.sum(fc * g(e.potC, hi))
    fr = np.sum(fc * g(e.rforceC, hi))
    fz = np.sum(fc * g(e.zforceC, hi))
    p0 = np.sum(ac[0] * g(e.potC, 0))
    fr0 = np.sum(ac[0] * g(e.rforceC, 0))
    fz0 = np.sum(ac[0] * g(e.zforceC, 0))
    fp = np.sum(ac[1:] * ssin * mo * g(e.potC, hi))
    fs = as_[1:] * ssin
    p += np.sum(fs * g(e.potS, hi))
    fr += np.sum(fs * g(e.rforceS, hi))
    fz += np.sum(fs * g(e.zforceS, hi))
    fp += np.sum(-as_[1:] * ccos * mo * g(e.potS, hi))
    return fr + fr0, fp, fz + fz0, p + p0


def field_points(F, x, y, z, fac, sev):
    """Fields.return_forces_cart point by point (potential.py:445-497):
    (fx, fy, fz, pot) arrays."""
    out = np.zeros((4, len(x)))
    ximin = np.min(F.sl.xi)
    for k in range(len(x)):
        r2 = math.sqrt(x[k] * x[k] + y[k] * y[k]) + 1e-15
        r3 = math.sqrt(r2 * r2 + z[k] * z[k]) + 1e-15
        phi = math.atan2(y[k], x[k])
        dfr, dfp, dfz, dp = _eof_force_point(F.eof, r2, z[k], phi, F.eof_cos, F.eof_sin)
        hfr, hft, hfp, hp1, hp0 = _sl_force_point(F.sl, r3, z[k] / r3, phi, F.sl_coef, fac, sev)
        if r3 < ximin:
            hfp = dfp = 0.0
        out[0, k] = dfr * x[k] / r2 - dfp * y[k] / (r2 * r2) - (
            hfr * x[k] / r3 - hft * x[k] * z[k] / r3 ** 3) + hfp * y[k] / (r2 * r2)
        out[1, k] = dfr * y[k] / r2 + dfp * x[k] / (r2 * r2) - (
            hfr * y[k] / r3 - hft * y[k] * z[k] / r3 ** 3) - hfp * x[k] / (r2 * r2)
        out[2, k] = dfz - (hfr * z[k] / r3 + hft * (r2 * r2) / r3 ** 3)
        out[3, k] = dp + hp1 + hp0
    return out


# ------------------------------------------------------------ process pool
_W = {}


def _load(d):
    """Worker initializer: tables from .npy files (memory-mapped), once."""
    a = lambda k: np.load(os.path.join(d, k + ".npy"), mmap_mode="r")
    sl = O.SLTable(*[int(v) for v in a("sl_int")], *a("sl_flt"), a("ev"), a("ef"), a("xi"),
                   a("p0"), a("d0"))
    ei = [int(v) for v in a("eof_int")]
    ef = [float(v) for v in a("eof_flt")]
    eo = O.EOFTable(*ei, *ef, *(np.asarray(a(k)) for k in
                                ("potC", "rforceC", "zforceC", "potS", "rforceS", "zforceS")))
    _W["sl"], _W["eof"] = sl, eo
    _W["fac"] = O.factorial_return(sl.lmax)
    _W["sev"] = np.sqrt(sl.evtable)


def _job(args):
    kind, data = args[0], args[1:]
    t0 = time.perf_counter()
    if kind == "eof":
        O.eof_accumulate(_W["eof"], *data)
    elif kind == "sl":
        sl_accumulate_loop(_W["sl"], *data, _W["fac"], _W["sev"])
    else:
        c, s, h, x, y, z = data
        F = O.FieldSpec(_W["eof"], c, s, _W["sl"], h)
        field_points(F, x, y, z, _W["fac"], _W["sev"])
    return time.perf_counter() - t0


def _save_tables(d, slt, eot):
    np.save(os.path.join(d, "sl_int.npy"), np.array([slt.lmax, slt.nmax, slt.numr, slt.cmap]))
    np.save(os.path.join(d, "sl_flt.npy"), np.array([slt.rmin, slt.rmax, slt.scale]))
    for k, v in (("ev", slt.evtable), ("ef", slt.eftable), ("xi", slt.xi), ("p0", slt.p0),
                 ("d0", slt.d0)):
        np.save(os.path.join(d, k + ".npy"), v)
    np.save(os.path.join(d, "eof_int.npy"), np.array([eot.mmax, eot.norder, eot.numx, eot.numy,
                                                      eot.cmap]))
    np.save(os.path.join(d, "eof_flt.npy"), np.array([eot.rmin, eot.rmax, eot.ascale, eot.hscale]))
    for k in ("potC", "rforceC", "zforceC", "potS", "rforceS", "zforceS"):
        np.save(os.path.join(d, k + ".npy"), getattr(eot, k))


def _blocks(arrs, nb):
    """Contiguous blocks, the first taking the remainder (spheresl.py:388-392)."""
    n = len(arrs[0])
    per = n // nb
    first = n - per * (nb - 1)
    cuts = [0, first] + [first + per * k for k in range(1, nb)]
    return [[a[cuts[k]:cuts[k + 1]] for a in arrs] for k in range(nb)]


def time_reference_path(slt, eot, disc, halo, coefs, ncores, n_eof, n_sl, n_field):
    """Wall seconds of each stage over `ncores` pool workers on samples of
    n_eof / n_sl disc / halo particles and n_field points (disc and halo
    halves).  Returns dict of stage seconds and the sample sizes."""
    ctx = mp.get_context("spawn")   # the parent may hold a GPU context
    with tempfile.TemporaryDirectory(prefix="refpath_") as d:
        _save_tables(d, slt, eot)
        with ctx.Pool(ncores, initializer=_load, initargs=(d,)) as pool:
            # a worker that cannot start is restarted by the pool forever:
            # every map has a time limit
            run = lambda jobs: pool.map_async(_job, jobs).get(timeout=600)
            run([("eof",) + tuple(v[:8] for v in disc)] * ncores)   # warm-up
            out = {}
            stages = (("eof", [("eof",) + tuple(b) for b in _blocks([v[:n_eof] for v in disc], ncores)]),
                      ("sl", [("sl",) + tuple(b) for b in _blocks([v[:n_sl] for v in halo], ncores)]))
            for name, jobs in stages:
                t0 = time.perf_counter()
                run(jobs)
                out[name] = time.perf_counter() - t0
            h = n_field // 2
            pts = [np.concatenate([disc[k][:h], halo[k][:n_field - h]]) for k in range(3)]
            jobs = [("field",) + tuple(coefs) + tuple(b) for b in _blocks(pts, ncores)]
            t0 = time.perf_counter()
            run(jobs)
            out["field"] = time.perf_counter() - t0
    out.update(n_eof=n_eof, n_sl=n_sl, n_field=n_field, ncores=ncores)
    return out
