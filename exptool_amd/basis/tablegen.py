"""Basis-table generation in the EXP cache formats.

The reference consumes SLGridSph and EmpCylSL cache files but cannot make them
(exptool/basis/eof.py:94-95), and none ship with it.  This module makes both:

* ``make_sl_tables``: the spherical Sturm-Liouville basis of a Hernquist
  background model.  For a Hernquist model of scale ``a`` the
  Hernquist-Ostriker (1992) functions are exact eigenfunctions of the SL
  problem  lap_l(phi) = 4 pi ev (rho0/Phi0) phi  with eigenvalue ev = K_nl
  (rho_nl/Phi_nl = K_nl * rho0/Phi0 for every r), so the tables are written
  with the EXP convention phi_ln = ef*p0/sqrt(ev), rho_ln = ef*d0*sqrt(ev)/(4 pi)
  (exptool/basis/spheresl.py:148,157-158,332) and are exact.
  Normalisation: int phi_ln rho_ln' r^2 dr = -delta/(4 pi), which together with
  real orthonormal harmonics makes the -4 pi accumulate factor
  (spheresl.py:586) reproduce the potential of the particles.

* ``make_eof_tables``: an EmpCylSL-style empirical basis.  For each m the
  covariance of per-particle projections of a disc sample onto a spherical HO
  basis (l = m..lmax_s, n < nmax_s) is diagonalised; the leading ``norder``
  eigenvectors define meridional functions Psi_mk(R, z) that are tabulated with
  -dPsi/dR, -dPsi/dz and the density on the (xi, eta) grid of
  set_table_params (eof.py:316-347).  Forces are analytic derivatives of the
  same functions, so pot/force tables are mutually consistent (F = -grad Phi
  up to bilinear-interpolation error).

Closed forms used (Hernquist & Ostriker 1992, ApJ 386, 375; eqs. 2.19-2.31):
  Phi_nl(s) = -s^l/(1+s)^(2l+1) C_n^(2l+3/2)(xi),  xi = (s-1)/(s+1)
  rho_nl(s) = K_nl/(2 pi) s^(l-1)/(1+s)^(2l+3) C_n^(2l+3/2)(xi)
  K_nl = n(n+4l+3)/2 + (l+1)(2l+1)
  I_nl = -K_nl 4 pi/2^(8l+6) Gamma(n+4l+3)/(n!(n+2l+3/2) Gamma(2l+3/2)^2)
with s = r/a and G = 1.
"""
from __future__ import annotations

import math
import numpy as np

from .compatibility import r_to_xi, xi_to_r, z_to_y, y_to_z


# ----------------------------------------------------------------------------
# Hernquist-Ostriker basis
# ----------------------------------------------------------------------------

def gegenbauer(nmax, alpha, x):
    """C_n^(alpha)(x) for n = 0..nmax-1 by the three-term recurrence
    n C_n = 2x(n+alpha-1) C_n-1 - (n+2 alpha-2) C_n-2.  Shape (nmax, ...)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((max(nmax, 1),) + x.shape)
    out[0] = 1.0
    if nmax > 1:
        out[1] = 2.0 * alpha * x
    for n in range(2, nmax):
        out[n] = (2.0 * x * (n + alpha - 1.0) * out[n - 1] - (n + 2.0 * alpha - 2.0) * out[n - 2]) / n
    return out[:nmax]


def Knl(n, l):
    return 0.5 * n * (n + 4 * l + 3) + (l + 1) * (2 * l + 1)


def log_abs_Inl(n, l):
    """log|I_nl| (HO92 eq. 2.31), computed in log space."""
    return (math.log(Knl(n, l)) + math.log(4.0 * math.pi) - (8 * l + 6) * math.log(2.0)
            + math.lgamma(n + 4 * l + 3) - math.lgamma(n + 1) - math.log(n + 2 * l + 1.5)
            - 2.0 * math.lgamma(2 * l + 1.5))


def ho_norm(n, l, a):
    """Factor alpha_nl with  int (alpha Phi_nl/a)(alpha rho_nl/a^3) r^2 dr = -1/(4 pi):
    int Phi_raw rho_raw s^2 ds = I_nl/(4 pi)  =>  alpha^2 = a/|I_nl|."""
    return math.exp(0.5 * (math.log(a) - log_abs_Inl(n, l)))


def ho_functions(lmax, nmax, r, a, derivs=False):
    """Normalised HO potential, density (and dPhi/dr) for l<=lmax, n<nmax at
    radii r (any shape).  Returns arrays shaped (lmax+1, nmax) + r.shape."""
    r = np.asarray(r, dtype=np.float64)
    s = r / a
    xi = (s - 1.0) / (s + 1.0)
    pot = np.zeros((lmax + 1, nmax) + r.shape)
    den = np.zeros_like(pot)
    dpot = np.zeros_like(pot) if derivs else None
    with np.errstate(divide="ignore", invalid="ignore"):
        for l in range(lmax + 1):
            alpha = 2.0 * l + 1.5
            C = gegenbauer(nmax, alpha, xi)
            sl = s ** l
            base = sl / (1.0 + s) ** (2 * l + 1)
            if derivs:
                Cd = np.zeros_like(C)
                if nmax > 1:
                    Cd[1:] = 2.0 * alpha * gegenbauer(nmax - 1, alpha + 1.0, xi)
                dbase = (l * (s ** (l - 1) if l > 0 else 0.0) / (1.0 + s) ** (2 * l + 1)
                         - (2 * l + 1) * sl / (1.0 + s) ** (2 * l + 2))
                dxi = 2.0 / (1.0 + s) ** 2
            for n in range(nmax):
                nrm = ho_norm(n, l, a)
                pot[l, n] = -nrm * base * C[n] / a
                den[l, n] = nrm * Knl(n, l) / (2.0 * math.pi) * (sl / (s * (1.0 + s) ** (2 * l + 3))) * C[n] / a ** 3
                if derivs:
                    dpot[l, n] = -nrm * (dbase * C[n] + base * Cd[n] * dxi) / (a * a)
    if derivs:
        return pot, den, dpot
    return pot, den


def hernquist_model(r, mass=1.0, a=1.0):
    """Hernquist (1990) profile: (density, enclosed mass, potential), G=1."""
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore"):
        dens = mass * a / (2.0 * np.pi * r * (r + a) ** 3)
    return dens, mass * r * r / (r + a) ** 2, -mass / (r + a)


# ----------------------------------------------------------------------------
# SLGridSph tables
# ----------------------------------------------------------------------------

def sl_grid(numr, rmin, rmax, cmap, scale):
    """xi grid of init_table (halo_methods.py:190-213; cmap=2 uses the
    intended xi = ln r)."""
    if cmap == 1:
        xmin = (rmin / scale - 1.0) / (rmin / scale + 1.0)
        xmax = (rmax / scale - 1.0) / (rmax / scale + 1.0)
    elif cmap == 2:
        xmin, xmax = math.log(rmin), math.log(rmax)
    else:
        xmin, xmax = rmin, rmax
    dxi = (xmax - xmin) / (numr - 1)
    xi = xmin + dxi * np.arange(numr, dtype=np.float64)
    return xi, xi_to_r(xi, cmap, scale)


def make_sl_tables(lmax, nmax, numr, rmin, rmax, cmap, scale, model_mass=1.0):
    """Return dict(lmax, nmax, numr, cmap, rmin, rmax, scale, ev (L+1,n),
    ef (L+1,n,numr), xi, r, p0, d0, model=(R, D, M, P)) for a Hernquist
    background of scale ``scale``.  p0 = model potential at the nodes,
    d0 = 4 pi model density (halo_methods.py:184,208)."""
    xi, r = sl_grid(numr, rmin, rmax, cmap, scale)
    dens, menc, pot0 = hernquist_model(r, model_mass, scale)
    phi, _ = ho_functions(lmax, nmax, r, scale)
    ev = np.array([[Knl(n, l) for n in range(nmax)] for l in range(lmax + 1)], dtype=np.float64)
    ef = phi * np.sqrt(ev)[:, :, None] / pot0[None, None, :]
    return dict(lmax=lmax, nmax=nmax, numr=numr, cmap=cmap, rmin=rmin, rmax=rmax,
                scale=scale, ev=ev, ef=ef, xi=xi, r=r, p0=pot0.copy(),
                d0=4.0 * np.pi * dens, model=(r.copy(), dens, menc, pot0.copy()))


# ----------------------------------------------------------------------------
# EmpCylSL-style EOF tables
# ----------------------------------------------------------------------------

def _legendre_all(lmax, x):
    """P_l^m(x) with Condon-Shortley phase, shape (lmax+1, lmax+1, ...)."""
    x = np.asarray(x, dtype=np.float64)
    p = np.zeros((lmax + 1, lmax + 1) + x.shape)
    p[0, 0] = 1.0
    somx2 = np.sqrt(np.clip((1.0 - x) * (1.0 + x), 0.0, None))
    fact = 1.0
    for m in range(1, lmax + 1):
        p[m, m] = p[m - 1, m - 1] * (-fact * somx2)
        fact += 2.0
    for m in range(lmax):
        p[m + 1, m] = x * (2 * m + 1) * p[m, m]
        for l in range(m + 2, lmax + 1):
            p[l, m] = (x * (2 * l - 1) * p[l - 1, m] - (l + m - 1) * p[l - 2, m]) / (l - m)
    return p


def _nlm(l, m):
    v = math.sqrt((0.5 * l + 0.25) / math.pi * math.factorial(l - m) / math.factorial(l + m))
    return v * math.sqrt(2.0) if m else v


def make_eof_tables(disc_xyzm, mmax=6, norder=18, numx=256, numy=128, rmin=0.001, rmax=20.0,
                    ascale=1.0, hscale=0.1, cmap=1, lmax_s=24, nmax_s=16, a_s=1.0,
                    with_density=True, shared_sincos=True):
    """Build EOF tables from a disc particle sample (x, y, z, m arrays).

    Returns dict with mmax, norder, numx, numy, cmap, rmin, rmax, ascale,
    hscale and potC, rforceC, zforceC, densC, potS, rforceS, zforceS, densS,
    each (mmax+1, norder, numx+1, numy+1) float64 (S arrays zero at m=0).

    shared_sincos=True diagonalises the summed cos and sin covariances, so
    potS = potC for m > 0 and the expansion is closed under rotations about
    z; False gives each parity its own eigenbasis (the spans then differ)."""
    x, y, z, ms = (np.asarray(v, dtype=np.float64) for v in disc_xyzm)
    if norder > (lmax_s - mmax + 1) * nmax_s:
        raise ValueError("norder exceeds the spherical basis size for m=mmax")
    # --- projections of the sample on the spherical basis -------------------
    rp = np.sqrt(x * x + y * y + z * z)
    rp = np.maximum(rp, 1e-10)
    cp = z / rp
    php = np.arctan2(y, x)
    Pp = _legendre_all(lmax_s, cp)
    phi_p, _ = ho_functions(lmax_s, nmax_s, rp, a_s)
    # --- table nodes -------------------------------------------------------
    Rtable = math.sqrt(0.5) * rmax
    XMIN = float(r_to_xi(rmin * ascale, cmap, ascale))
    XMAX = float(r_to_xi(Rtable * ascale, cmap, ascale))
    dX = (XMAX - XMIN) / numx
    YMIN = float(z_to_y(-Rtable * ascale, hscale))
    YMAX = float(z_to_y(Rtable * ascale, hscale))
    dY = (YMAX - YMIN) / numy
    Rn = xi_to_r(XMIN + dX * np.arange(numx + 1), cmap, ascale)
    zn = y_to_z(YMIN + dY * np.arange(numy + 1), hscale)
    RR, ZZ = np.meshgrid(Rn, zn, indexing="ij")
    rr = np.sqrt(RR * RR + ZZ * ZZ)
    cc = ZZ / rr
    Pn = _legendre_all(lmax_s, cc)
    phin, denn, dphin = ho_functions(lmax_s, nmax_s, rr, a_s, derivs=True)

    shape = (mmax + 1, norder, numx + 1, numy + 1)
    out = {k: np.zeros(shape) for k in ("potC", "rforceC", "zforceC", "densC",
                                         "potS", "rforceS", "zforceS", "densS")}
    for m in range(mmax + 1):
        ls = list(range(m, lmax_s + 1))
        nb = len(ls) * nmax_s
        # per-particle projections u_j(p)
        ang = [np.cos(m * php), np.sin(m * php)]
        # meridional basis functions at the nodes and their R, z derivatives
        B = np.zeros((nb,) + RR.shape)
        BR = np.zeros_like(B)
        BZ = np.zeros_like(B)
        BD = np.zeros_like(B)
        j = 0
        Up = np.zeros((nb, x.size))
        for l in ls:
            N = _nlm(l, m)
            P = N * Pn[l, m]
            Pm1 = N * Pn[l - 1, m] if l - 1 >= m else np.zeros_like(P)
            # l c P_l^m - (l+m) P_{l-1}^m, normalised by the same N_lm
            brk = l * cc * P - (l + m) * Pm1
            for n in range(nmax_s):
                f, df, rho = phin[l, n], dphin[l, n], denn[l, n]
                B[j] = f * P
                # dP/dc * dc/dR = brk * z/(R r);  dP/dc * dc/dz = -brk / r
                BR[j] = df * (RR / rr) * P + f * brk * ZZ / (RR * rr)
                BZ[j] = df * (ZZ / rr) * P - f * brk / rr
                BD[j] = rho * P
                Up[j] = ms * N * Pp[l, m] * phi_p[l, n]
                j += 1
        covs = [(Up * a[None, :]) @ (Up * a[None, :]).T for a in ang]
        for kind in (0, 1):
            if m == 0 and kind == 1:
                continue
            cov = covs[0] + covs[1] if (shared_sincos and m > 0) else covs[kind]
            w, V = np.linalg.eigh(cov)
            order = np.argsort(w)[::-1][:norder]
            V = V[:, order]
            # deterministic sign: largest-magnitude component positive
            idx = np.argmax(np.abs(V), axis=0)
            V = V * np.sign(V[idx, np.arange(V.shape[1])])[None, :]
            flat = lambda A: A.reshape(nb, -1)
            sfx = "C" if kind == 0 else "S"
            out["pot" + sfx][m] = (V.T @ flat(B)).reshape(norder, numx + 1, numy + 1)
            out["rforce" + sfx][m] = -(V.T @ flat(BR)).reshape(norder, numx + 1, numy + 1)
            out["zforce" + sfx][m] = -(V.T @ flat(BZ)).reshape(norder, numx + 1, numy + 1)
            out["dens" + sfx][m] = (V.T @ flat(BD)).reshape(norder, numx + 1, numy + 1)
    out.update(mmax=mmax, norder=norder, numx=numx, numy=numy, cmap=cmap, rmin=rmin,
               rmax=rmax, ascale=ascale, hscale=hscale, dens=1 if with_density else 0,
               nmax=nmax_s, cylmass=float(np.sum(ms)), time=0.0)
    if not with_density:
        out["densC"][:] = 0.0
        out["densS"][:] = 0.0
    return out
