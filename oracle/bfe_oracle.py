"""FP64 CPU oracle for exptool's basis-function-expansion (BFE) hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``exptool_amd`` imports this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline).

Every function restates the reference algorithm in vectorised NumPy float64
and cites the reference file:line it follows (paths relative to the exptool
source tree).  Where the reference has an arithmetic quirk, the function takes
``compat`` (reproduce the quirk) and otherwise computes the intended value:

  * SL ``force_eval`` multiplies the (l,m) term by cos/sin(l*phi)
    (spheresl.py:1222-1225)                                   -> compat
  * SL radial-derivative stencil at the first table cell is evaluated one cell
    outward (spheresl.py:150-153)                             -> compat
  * ``all_eval_particles`` uses P_1^m for den1 and densfac 0.25*pi
    (spheresl.py:1323,1351)                                   -> compat
  * EOF ``force_eval`` no_odd mask is |cos(pi m/2)| (~6e-17, not 0) for odd m
    (eof.py:833)                                              -> compat
  * EOF upper-edge linear extrapolation (eof.py:414-422): always reproduced.
  * ``compute_timestep`` uses dfz*hfz as the vertical force (integrate.py:249,
    251)                                                      -> compat_dt
  * cmap=2 (xi = ln r) raises NameError in the reference
    (halo_methods.py:195, spheresl.py:128-130): the intended map is used and
    xi is clamped to the table range.

How each function is pinned (golden vectors generated from the reference,
closed forms, invariants) is listed in tests/test_oracle_*.py.  Functions whose
only pin is a golden vector say so in their docstring.
"""
from __future__ import annotations

import math
import numpy as np

FOURPI = 4.0 * np.pi


# ----------------------------------------------------------------------------
# R1/R2: coordinate maps  (basis/compatibility.py:16-99, accumulate_c/accumulate.c:7-87)
# ----------------------------------------------------------------------------

def r_to_xi(r, cmap, scale):
    """compatibility.py:16-47.  cmap 1: (r/s-1)/(r/s+1); cmap 2: ln r; else r.
    r < 0 maps to 0."""
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        if cmap == 1:
            out = (r / scale - 1.0) / (r / scale + 1.0)
        elif cmap == 2:
            out = np.log(r)
        else:
            out = r.copy()
    return np.where(r < 0.0, 0.0, out)


def xi_to_r(xi, cmap, scale):
    """compatibility.py:54-62."""
    xi = np.asarray(xi, dtype=np.float64)
    if cmap == 1:
        return (1.0 + xi) / (1.0 - xi) * scale
    if cmap == 2:
        return np.exp(xi)
    return xi


def d_xi_to_r(xi, cmap, scale):
    """compatibility.py:65-80.  Returns d(xi)/dr despite the name."""
    xi = np.asarray(xi, dtype=np.float64)
    if cmap == 1:
        return 0.5 * (1.0 - xi) * (1.0 - xi) / scale
    if cmap == 2:
        return np.exp(-xi)
    return np.ones_like(xi)


def z_to_y(z, hscale, eps=1.0e-8):
    """compatibility.py:83-91 (eps=1e-8; the dead C twin uses 1e-10,
    accumulate.c:81)."""
    z = np.asarray(z, dtype=np.float64)
    return (z / (np.abs(z) + eps)) * np.arcsinh(np.abs(z / hscale))


def y_to_z(y, hscale):
    """compatibility.py:94-99."""
    return hscale * np.sinh(np.asarray(y, dtype=np.float64))


# ----------------------------------------------------------------------------
# R5/R6: associated Legendre functions and normalisation
# ----------------------------------------------------------------------------

def legendre_R(lmax, x):
    """spheresl.py:664-700, vectorised over x: P_l^m(x) with Condon-Shortley
    phase (Numerical-Recipes plgndr recurrences); non-finite entries -> 0.
    Returns array (lmax+1, lmax+1, npts)."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    p = np.zeros((lmax + 1, lmax + 1, x.size))
    with np.errstate(invalid="ignore", over="ignore"):
        p[0, 0] = 1.0
        pll = np.ones_like(x)
        if lmax > 0:
            somx2 = np.sqrt((1.0 - x) * (1.0 + x))
            fact = 1.0
            for m in range(1, lmax + 1):
                pll = pll * (-fact * somx2)
                p[m, m] = pll
                fact += 2.0
        for m in range(0, lmax):
            pl2 = p[m, m].copy()
            pl1 = x * (2.0 * m + 1) * pl2
            p[m + 1, m] = pl1
            for l in range(m + 2, lmax + 1):
                pll = (x * (2 * l - 1) * pl1 - (l + m - 1) * pl2) / (l - m)
                p[l, m] = pll
                pl2 = pl1
                pl1 = pll
    p[~np.isfinite(p)] = 0.0
    return p


def dlegendre_R(lmax, x):
    """spheresl.py:706-770, vectorised: (P, dP/dx).  x is clamped to
    +-(1-1e-8) for the derivative only (spheresl.py:752-755); P itself uses
    the unclamped x."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    p = legendre_R(lmax, x)
    dp = np.zeros_like(p)
    mineps = 1.0e-8
    with np.errstate(invalid="ignore"):
        xc = np.where(1.0 - np.abs(x) < mineps,
                      np.where(x > 0, 1.0 - mineps, -(1.0 - mineps)), x)
        with np.errstate(divide="ignore", over="ignore"):
            somx2 = 1.0 / (xc * xc - 1.0)
            for l in range(1, lmax + 1):
                for m in range(0, l):
                    dp[l, m] = somx2 * (xc * l * p[l, m] - (l + m) * p[l - 1, m])
                dp[l, l] = somx2 * xc * l * p[l, l]
    dp[~np.isfinite(dp)] = 0.0
    return p, dp


def factorial_return(lmax):
    """spheresl.py:823-863: N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), times
    sqrt(2) for m>0.  The reference uses gammaln; exact integer factorials are
    used here (agreement ~1e-16)."""
    f = np.zeros((lmax + 1, lmax + 1))
    for l in range(lmax + 1):
        for m in range(l + 1):
            v = math.sqrt((0.5 * l + 0.25) / math.pi
                          * math.factorial(l - m) / math.factorial(l + m))
            if m != 0:
                v *= math.sqrt(2.0)
            f[l, m] = v
    return f


# ----------------------------------------------------------------------------
# SL radial tables (R3/R4/R7/R10)
# ----------------------------------------------------------------------------

class SLTable:
    """In-memory form of an SLGridSph cache + model table
    (halo_methods.py:96-169 read_cached_table, :178-220 init_table)."""

    def __init__(self, lmax, nmax, numr, cmap, rmin, rmax, scale,
                 evtable, eftable, xi, p0, d0):
        self.lmax, self.nmax, self.numr = int(lmax), int(nmax), int(numr)
        self.cmap, self.rmin, self.rmax, self.scale = int(cmap), float(rmin), float(rmax), float(scale)
        self.evtable = np.asarray(evtable, dtype=np.float64)
        self.eftable = np.asarray(eftable, dtype=np.float64)
        self.xi = np.asarray(xi, dtype=np.float64)
        self.p0 = np.asarray(p0, dtype=np.float64)
        self.d0 = np.asarray(d0, dtype=np.float64)


def _sl_locate(t: SLTable, r):
    """Shared cell search of get_halo_pot_matrix / get_halo_dens_pot_force
    (spheresl.py:308-328 and :123-140).  Returns clamped xi, cell index, x1, x2."""
    x = r_to_xi(r, t.cmap, t.scale)
    if t.cmap == 1:
        x = np.where(x < -1.0, -1.0, x)
        x = np.where(x >= 1.0, 1.0 - 1.0e-8, x)
    elif t.cmap == 2:
        # intended semantics of spheresl.py:128-130 (NameError as written)
        x = np.clip(x, t.xi[0], t.xi[-1])
    dxi = t.xi[1] - t.xi[0]
    with np.errstate(invalid="ignore"):
        fl = np.floor((x - np.min(t.xi)) / dxi)
    fl = np.where(np.isfinite(fl), fl, 0.0)
    indx = np.clip(fl, 0, t.numr - 2).astype(np.int64)
    x1 = (t.xi[indx + 1] - x) / dxi
    x2 = (x - t.xi[indx]) / dxi
    return x, indx, x1, x2, dxi


def get_halo_pot_matrix(t: SLTable, r):
    """spheresl.py:301-335, vectorised: phi_ln(r) = (x1 ef_i + x2 ef_i+1)/sqrt(ev)
    * (x1 p0_i + x2 p0_i+1).  Returns (npts, lmax+1, nmax)."""
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    _, i, x1, x2, _ = _sl_locate(t, r)
    ef_lo = t.eftable[:, :, i].transpose(2, 0, 1)
    ef_hi = t.eftable[:, :, i + 1].transpose(2, 0, 1)
    pp = (x1 * t.p0[i] + x2 * t.p0[i + 1])[:, None, None]
    return (x1[:, None, None] * ef_lo + x2[:, None, None] * ef_hi) / np.sqrt(t.evtable)[None] * pp


def get_halo_dens_pot_force(t: SLTable, r, compat=False):
    """spheresl.py:106-160, vectorised: returns (dens, dpot, pot) each
    (npts, lmax+1, nmax).  dpot = d(phi_ln)/dr from the 3-point quadratic
    stencil times d(xi)/dr / dxi.  compat=True reproduces the first-cell
    stencil quirk (spheresl.py:150-153: nodes 0,1,2 with x2 unshifted);
    otherwise the quadratic through nodes 0,1,2 is evaluated at the particle."""
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    x, i, x1, x2, dxi = _sl_locate(t, r)
    sev = np.sqrt(t.evtable)[None]
    ef = t.eftable
    w1, w2 = x1[:, None, None], x2[:, None, None]
    efi = ef[:, :, i].transpose(2, 0, 1)
    efi1 = ef[:, :, i + 1].transpose(2, 0, 1)
    dens = (w1 * efi + w2 * efi1) * sev * (x1 * t.d0[i] + x2 * t.d0[i + 1])[:, None, None]
    pot = (w1 * efi + w2 * efi1) / sev * (x1 * t.p0[i] + x2 * t.p0[i + 1])[:, None, None]
    fac = (d_xi_to_r(x, t.cmap, t.scale) / dxi)[:, None, None]
    c = np.where(i == 0, 1, i)                      # stencil centre node
    pstar = np.where(i == 0, x2 if compat else x2 - 1.0, x2)[:, None, None]
    g = lambda k: ef[:, :, k].transpose(2, 0, 1) * t.p0[k][:, None, None]
    dpot = fac * ((pstar - 0.5) * g(c - 1) - 2.0 * pstar * g(c) + (pstar + 0.5) * g(c + 1)) / sev
    return dens, dpot, pot


# ----------------------------------------------------------------------------
# R8: SL accumulate  (spheresl.py:567-656)
# ----------------------------------------------------------------------------

def sl_row_index(lmax):
    """(row, l, m, kind) for the ((lmax+1)^2, nmax) expcoef layout:
    row l^2 is m=0; rows l^2+2m-1 / l^2+2m are cos/sin (spheresl.py:626-654)."""
    rows = []
    for l in range(lmax + 1):
        rows.append((l * l, l, 0, 0))
        for m in range(1, l + 1):
            rows.append((l * l + 2 * m - 1, l, m, 0))
            rows.append((l * l + 2 * m, l, m, 1))
    return rows


def sl_accumulate(t: SLTable, x, y, z, mass, no_odd=False):
    """compute_coefficients_solitary, spheresl.py:567-656, vectorised over
    particles.  r = max(|x|, 1e-10) (:611), cos(theta) = z/r, phi = atan2(y,x);
    row += (-4 pi) N_lm P_l^m phi_ln(r) {1, cos m phi, sin m phi} m_p.
    Odd l rows are skipped (left zero) when no_odd (:630-632).
    Returns expcoef ((lmax+1)^2, nmax)."""
    x, y, z, mass = (np.asarray(a, dtype=np.float64) for a in (x, y, z, mass))
    with np.errstate(invalid="ignore"):
        r = np.sqrt(x * x + y * y + z * z)
    r = np.where(np.isnan(r), 1.0e-10, np.maximum(r, 1.0e-10))   # np.nanmax semantics
    costh = z / r
    phi = np.arctan2(y, x)
    legs = legendre_R(t.lmax, costh)
    potd = get_halo_pot_matrix(t, r)                 # (P, L+1, n)
    fac = factorial_return(t.lmax)
    out = np.zeros(((t.lmax + 1) ** 2, t.nmax))
    for row, l, m, kind in sl_row_index(t.lmax):
        if no_odd and (l % 2) != 0:
            continue
        w = fac[l, m] * legs[l, m] * (-FOURPI) * mass
        if m > 0:
            w = w * (np.cos(phi * m) if kind == 0 else np.sin(phi * m))
        out[row] = w @ potd[:, l, :]
    return out


# ----------------------------------------------------------------------------
# R11/R12: SL evaluation
# ----------------------------------------------------------------------------

def _sl_eval_core(t, r, costh, phi, expcoef, lmax_use, nmax_use, lrange, no_odd,
                  cos_l_quirk, den_quirk, edge_compat, want_density):
    fac = factorial_return(t.lmax)
    dend, dpot, potd = get_halo_dens_pot_force(t, r, compat=edge_compat)
    legs, dlegs = dlegendre_R(t.lmax, costh)
    n = nmax_use
    c = expcoef
    f00 = fac[0, 0]
    den0 = f00 * (dend[:, 0, :n] @ c[0, :n])
    pot0 = f00 * (potd[:, 0, :n] @ c[0, :n])
    potr = f00 * (dpot[:, 0, :n] @ c[0, :n])
    den1 = np.zeros_like(pot0)
    pot1 = np.zeros_like(pot0)
    pott = np.zeros_like(pot0)
    potp = np.zeros_like(pot0)
    for l in range(1, lmax_use + 1):
        if lrange is not None and (l > lrange[1] or l < lrange[0]):
            continue
        if no_odd and (l % 2) != 0:
            continue
        for m in range(0, l + 1):
            f = fac[l, m]
            if m == 0:
                row = l * l
                pl = potd[:, l, :n] @ c[row, :n]
                dl = dpot[:, l, :n] @ c[row, :n]
                el = dend[:, l, :n] @ c[row, :n]
                lden = legs[1, 0] if den_quirk else legs[l, 0]
                den1 += f * lden * el
                pot1 += f * legs[l, 0] * pl
                potr += f * legs[l, 0] * dl
                pott += f * dlegs[l, 0] * pl
            else:
                rc, rs = l * l + 2 * m - 1, l * l + 2 * m
                ang = l if cos_l_quirk else m
                cm, sm = np.cos(phi * ang), np.sin(phi * ang)
                pc, ps = potd[:, l, :n] @ c[rc, :n], potd[:, l, :n] @ c[rs, :n]
                dc, ds = dpot[:, l, :n] @ c[rc, :n], dpot[:, l, :n] @ c[rs, :n]
                ec, es = dend[:, l, :n] @ c[rc, :n], dend[:, l, :n] @ c[rs, :n]
                den1 += f * legs[l, m] * (ec * cm + es * sm)
                pot1 += f * legs[l, m] * (pc * cm + ps * sm)
                potr += f * legs[l, m] * (dc * cm + ds * sm)
                pott += f * dlegs[l, m] * (pc * cm + ps * sm)
                potp += f * legs[l, m] * m * (-pc * sm + ps * cm)
    return den0, den1, pot0, pot1, potr, pott, potp


def sl_all_eval_particles(t: SLTable, x, y, z, expcoef, L1=-1000, L2=1000,
                          no_odd=False, compat=False, edge_compat=None):
    """all_eval_particles, spheresl.py:1240-1362, vectorised.  r = |x| with no
    floor (:1257), rr = planar radius, cos(theta) = z/r, phi = atan2(y,x).
    compat=True: den1 uses P_1^m for m=0 (:1323) and densfac = 0.25*pi
    (:1351); default uses P_l^m and 0.25/pi (as all_eval, :1092).
    Returns (den0, den1, pot0, pot1, potr, pott, potp, rr)."""
    x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
    if edge_compat is None:
        edge_compat = compat
    r = np.sqrt(x * x + y * y + z * z)
    rr = np.sqrt(x * x + y * y)
    phi = np.arctan2(y, x)
    with np.errstate(invalid="ignore", divide="ignore"):
        costh = z / r
    den0, den1, pot0, pot1, potr, pott, potp = _sl_eval_core(
        t, r, costh, phi, np.asarray(expcoef, dtype=np.float64), t.lmax, t.nmax,
        (L1, L2), no_odd, False, compat, edge_compat, True)
    densfac = 0.25 * np.pi if compat else 0.25 / np.pi
    return den0 * densfac, den1 * densfac, pot0, pot1, potr, pott, potp, rr


def sl_truncation(t: SLTable, lmax, nmax):
    """Orders used when force_eval / all_eval are asked for lmax, nmax
    (spheresl.py:1134-1147, :1009-1022): l is cut when lmax < the table's
    lmax, n when nmax < the table's nmax - 1.  The reference fails on the
    other requests (nmax = table nmax - 1 leaves expcoef and the radial
    matrices of different lengths, and all_eval keeps lmax + 1 instead of
    (lmax + 1)^2 rows, :1017); here they mean the full table.
    Returns (lmax_use, nmax_use)."""
    lmax = t.lmax if lmax is None else int(lmax)
    nmax = t.nmax if nmax is None else int(nmax)
    return min(lmax, t.lmax), (nmax if nmax < t.nmax - 1 else t.nmax)


def sl_all_eval(t: SLTable, r, costh, phi, expcoef, lmax=None, nmax=None, no_odd=False,
                edge_compat=True):
    """all_eval, spheresl.py:987-1102, vectorised over points: the expansion
    at (r, cos theta, phi) with density.  den1 is the total density (seeded
    with the l = 0 term, :1046), the azimuthal factor is cos/sin(m phi) and
    densfac = 0.25/pi (:1092).  get_halo_dens_pot_force's first-cell stencil
    (:150-153) is what the reference evaluates (edge_compat=True).
    Returns (den0, den1, pot0, pot1, potr, pott, potp)."""
    lmax_use, nmax_use = sl_truncation(t, lmax, nmax)
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    costh = np.atleast_1d(np.asarray(costh, dtype=np.float64))
    phi = np.atleast_1d(np.asarray(phi, dtype=np.float64))
    den0, den1, pot0, pot1, potr, pott, potp = _sl_eval_core(
        t, r, costh, phi, np.asarray(expcoef, dtype=np.float64), lmax_use, nmax_use,
        None, no_odd, False, False, edge_compat, True)
    densfac = 0.25 / np.pi
    return den0 * densfac, (den0 + den1) * densfac, pot0, pot1, potr, pott, potp


def sl_force_eval(t: SLTable, r, costh, phi, expcoef, lmax=None, nmax=None,
                  no_odd=False, compat=False, edge_compat=None):
    """force_eval, spheresl.py:1107-1234, vectorised over points.  Truncates to
    lmax/nmax (:1138-1148).  compat=True reproduces the cos/sin(l*phi)
    azimuthal factor (:1222-1225).  Returns (potr, pott, potp, pot1, pot0):
    d(Phi)/dr, d(Phi)/d(cos theta), d(Phi)/d(phi), the l>=1 and l=0 potential."""
    lmax = t.lmax if lmax is None else int(lmax)
    nmax = t.nmax if nmax is None else int(nmax)
    if edge_compat is None:
        edge_compat = compat
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    costh = np.atleast_1d(np.asarray(costh, dtype=np.float64))
    phi = np.atleast_1d(np.asarray(phi, dtype=np.float64))
    _, _, pot0, pot1, potr, pott, potp = _sl_eval_core(
        t, r, costh, phi, np.asarray(expcoef, dtype=np.float64), lmax, nmax,
        None, no_odd, compat, False, edge_compat, False)
    return potr, pott, potp, pot1, pot0


# ----------------------------------------------------------------------------
# EOF (R15-R21)
# ----------------------------------------------------------------------------

class EOFTable:
    """In-memory EmpCylSL cache (eof.py:98-313) plus derived grid parameters
    (set_table_params, eof.py:316-347)."""

    def __init__(self, mmax, norder, numx, numy, cmap, rmin, rmax, ascale, hscale,
                 potC, rforceC, zforceC, potS, rforceS, zforceS, densC=None, densS=None):
        self.mmax, self.norder, self.numx, self.numy = int(mmax), int(norder), int(numx), int(numy)
        self.cmap, self.rmin, self.rmax = int(cmap), float(rmin), float(rmax)
        self.ascale, self.hscale = float(ascale), float(hscale)
        self.potC, self.rforceC, self.zforceC = potC, rforceC, zforceC
        self.potS, self.rforceS, self.zforceS = potS, rforceS, zforceS
        self.densC, self.densS = densC, densS
        (self.XMIN, self.XMAX, self.dX, self.YMIN, self.YMAX, self.dY) = set_table_params(
            self.rmax, self.rmin, self.ascale, self.hscale, self.numx, self.numy, self.cmap)


def set_table_params(RMAX, RMIN, ASCALE, HSCALE, NUMX, NUMY, CMAP):
    """eof.py:316-347."""
    Rtable = math.sqrt(0.5) * RMAX
    XMIN = float(r_to_xi(RMIN * ASCALE, CMAP, ASCALE))
    XMAX = float(r_to_xi(Rtable * ASCALE, CMAP, ASCALE))
    dX = (XMAX - XMIN) / NUMX
    YMIN = float(z_to_y(-Rtable * ASCALE, HSCALE))
    YMAX = float(z_to_y(Rtable * ASCALE, HSCALE))
    dY = (YMAX - YMIN) / NUMY
    return XMIN, XMAX, dX, YMIN, YMAX, dY


def return_bins(r, z, rmin, dR, zmin, dZ, numx, numy, ASCALE, HSCALE, CMAP):
    """eof.py:354-427.  X = (xi(r)-rmin)/dR, ix = trunc(X); negative X -> 0;
    ix clamped to numx-1 but X is NOT clamped above (upper-edge linear
    extrapolation, :414-415).  Same for Y."""
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    X = (r_to_xi(r, CMAP, ASCALE) - rmin) / dR
    Y = (z_to_y(z, HSCALE) - zmin) / dZ
    X = np.where(X < 0, 0.0, X)
    Y = np.where(Y < 0, 0.0, Y)
    ix = np.minimum(np.trunc(np.minimum(X, numx)), numx - 1).astype(np.int64)
    iy = np.minimum(np.trunc(np.minimum(Y, numy)), numy - 1).astype(np.int64)
    return X, Y, ix, iy


def _bilinear_weights(X, Y, ix, iy):
    """eof.py:443-451."""
    delx0 = ix + 1.0 - X
    dely0 = iy + 1.0 - Y
    delx1 = X - ix
    dely1 = Y - iy
    return delx0 * dely0, delx1 * dely0, delx0 * dely1, delx1 * dely1


def _interp(T, ix, iy, c):
    """Bilinear gather of table T (M, N, numx+1, numy+1) -> (npts, M, N)."""
    c00, c10, c01, c11 = c
    g = lambda a, b: T[:, :, a, b].transpose(2, 0, 1)
    return (g(ix, iy) * c00[:, None, None] + g(ix + 1, iy) * c10[:, None, None]
            + g(ix, iy + 1) * c01[:, None, None] + g(ix + 1, iy + 1) * c11[:, None, None])


def eof_accumulate(t: EOFTable, x, y, z, mass, mmax=None, norder=None):
    """accumulate, eof.py:492-640 (the no_odd mask is computed but never
    applied there, :545-551, so it has no argument here).  R = sqrt(x^2+y^2+1e-10),
    phi = atan2(y,x); accum_cos[m,n] = sum_p -4 pi cos(m phi) m_p Vc[m,n],
    accum_sin likewise with sin and potS.  Returns (accum_cos, accum_sin),
    each (mmax+1, norder)."""
    mmax = t.mmax if mmax is None else int(mmax)
    norder = t.norder if norder is None else int(norder)
    x, y, z, mass = (np.asarray(a, dtype=np.float64) for a in (x, y, z, mass))
    r = (x * x + y * y + 1.0e-10) ** 0.5
    phi = np.arctan2(y, x)
    X, Y, ix, iy = return_bins(r, z, t.XMIN, t.dX, t.YMIN, t.dY, t.numx, t.numy,
                               t.ascale, t.hscale, t.cmap)
    c = _bilinear_weights(X, Y, ix, iy)
    vc = _interp(t.potC[:mmax + 1, :norder], ix, iy, c)
    vs = _interp(t.potS[:mmax + 1, :norder], ix, iy, c)
    m = np.arange(mmax + 1, dtype=np.float64)
    wc = -FOURPI * np.cos(phi[:, None] * m[None]) * mass[:, None]
    ws = -FOURPI * np.sin(phi[:, None] * m[None]) * mass[:, None]
    accum_cos = np.einsum("pm,pmn->mn", wc, vc)
    accum_sin = np.einsum("pm,pmn->mn", ws, vs)
    return accum_cos, accum_sin


def eof_force_eval(t: EOFTable, r, z, phi, accum_cos, accum_sin, MMAX=None, NMAX=None,
                   no_odd=False, compat=False):
    """force_eval, eof.py:756-872, vectorised over points (r is the planar
    radius as passed by the caller).  Returns (fr+fr0, fp, fz+fz0, p+p0, p0),
    where fp = -dPhi/dphi.  no_odd masks odd m: compat=True uses the
    reference's |cos(pi m/2)| (~6e-17 for odd m, :833); default uses 0."""
    MMAX = t.mmax if MMAX is None else int(MMAX)
    NMAX = t.norder if NMAX is None else int(NMAX)
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    phi = np.atleast_1d(np.asarray(phi, dtype=np.float64))
    ac = np.asarray(accum_cos, dtype=np.float64)[:MMAX + 1, :NMAX]
    as_ = np.asarray(accum_sin, dtype=np.float64)[:MMAX + 1, :NMAX]
    X, Y, ix, iy = return_bins(r, z, t.XMIN, t.dX, t.YMIN, t.dY, t.numx, t.numy,
                               t.ascale, t.hscale, t.cmap)
    c = _bilinear_weights(X, Y, ix, iy)
    sl = (slice(0, MMAX + 1), slice(0, NMAX))
    pc = _interp(t.potC[sl], ix, iy, c)
    rc = _interp(t.rforceC[sl], ix, iy, c)
    zc = _interp(t.zforceC[sl], ix, iy, c)
    ps = _interp(t.potS[sl], ix, iy, c)
    rs = _interp(t.rforceS[sl], ix, iy, c)
    zs = _interp(t.zforceS[sl], ix, iy, c)
    m = np.arange(1, MMAX + 1, dtype=np.float64)
    if no_odd:
        mask = np.abs(np.cos((np.pi / 2.0) * m)) if compat else (1.0 - (np.arange(1, MMAX + 1) % 2))
    else:
        mask = np.ones_like(m)
    ccos = np.cos(phi[:, None] * m[None])
    ssin = np.sin(phi[:, None] * m[None])
    fc = ac[None, 1:, :] * (ccos * mask)[:, :, None]
    p = np.sum(fc * pc[:, 1:], axis=(1, 2))
    fr = np.sum(fc * rc[:, 1:], axis=(1, 2))
    fz = np.sum(fc * zc[:, 1:], axis=(1, 2))
    p0 = pc[:, 0] @ ac[0]
    fr0 = rc[:, 0] @ ac[0]
    fz0 = zc[:, 0] @ ac[0]
    fpc = ac[None, 1:, :] * (ssin * mask * m)[:, :, None]
    fp = np.sum(fpc * pc[:, 1:], axis=(1, 2))
    fs = as_[None, 1:, :] * (ssin * mask)[:, :, None]
    p += np.sum(fs * ps[:, 1:], axis=(1, 2))
    fr += np.sum(fs * rs[:, 1:], axis=(1, 2))
    fz += np.sum(fs * zs[:, 1:], axis=(1, 2))
    fps = -as_[None, 1:, :] * (ccos * mask * m)[:, :, None]
    fp += np.sum(fps * ps[:, 1:], axis=(1, 2))
    return fr + fr0, fp, fz + fz0, p + p0, p0


def _eof_eval_core(t: EOFTable, R, z, phi, accum_cos, accum_sin, mlist, density):
    """Per-m sums shared by accumulated_eval (eof.py:874-929) and
    accumulated_eval_particles (eof.py:1092-1138) at planar radius R, height
    z, angle phi, over the orders in mlist: for each m, the bilinear table
    values times the coefficients summed over all norder orders; p, fr, fz
    (and d) take cos(m phi) C + sin(m phi) S, fp = m (sin(m phi) C potC -
    cos(m phi) S potS).  Returns (p0, p, fr, fp, fz, d0, d) with p0/d0 the
    m = 0 potential/density and p/d the m >= 1 part only."""
    R = np.atleast_1d(np.asarray(R, dtype=np.float64))
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    phi = np.atleast_1d(np.asarray(phi, dtype=np.float64))
    X, Y, ix, iy = return_bins(R, z, t.XMIN, t.dX, t.YMIN, t.dY, t.numx, t.numy,
                               t.ascale, t.hscale, t.cmap)
    c = _bilinear_weights(X, Y, ix, iy)
    npt = R.size
    p0 = np.zeros(npt); p = np.zeros(npt); fr = np.zeros(npt); fz = np.zeros(npt)
    fp = np.zeros(npt); d0 = np.zeros(npt); d = np.zeros(npt)
    ac = np.asarray(accum_cos, dtype=np.float64)
    as_ = np.asarray(accum_sin, dtype=np.float64)
    N = t.norder
    for mm in mlist:
        cm, sm = np.cos(phi * mm), np.sin(phi * mm)
        sl = (slice(mm, mm + 1), slice(0, N))
        vc = _interp(t.potC[sl], ix, iy, c)[:, 0] @ ac[mm, :N]
        rcv = _interp(t.rforceC[sl], ix, iy, c)[:, 0] @ ac[mm, :N]
        zcv = _interp(t.zforceC[sl], ix, iy, c)[:, 0] @ ac[mm, :N]
        pm = cm * vc
        fr += cm * rcv
        fz += cm * zcv
        fp += sm * mm * vc
        dm = np.zeros(npt)
        if density:
            dm = cm * (_interp(t.densC[sl], ix, iy, c)[:, 0] @ ac[mm, :N])
        if mm > 0:
            vs = _interp(t.potS[sl], ix, iy, c)[:, 0] @ as_[mm, :N]
            pm += sm * vs
            fr += sm * (_interp(t.rforceS[sl], ix, iy, c)[:, 0] @ as_[mm, :N])
            fz += sm * (_interp(t.zforceS[sl], ix, iy, c)[:, 0] @ as_[mm, :N])
            fp += -cm * mm * vs
            if density:
                dm += sm * (_interp(t.densS[sl], ix, iy, c)[:, 0] @ as_[mm, :N])
        if mm == 0:
            p0 += pm
            d0 += dm
        else:
            p += pm
            d += dm
    return p0, p, fr, fp, fz, d0, d


def eof_accumulated_eval(t: EOFTable, r, z, phi, accum_cos, accum_sin, MMAX=None,
                         no_odd=False):
    """accumulated_eval, eof.py:874-929, vectorised over points (r is the
    planar radius as passed by the caller).  Orders m = 0..MMAX, odd m skipped
    with no_odd (:891-892); every order n of the tables is used (NMAX has no
    effect there).  Returns (p0, p, fr, fp, fz, d0, d) with p and d the
    totals over m (p0, d0: the m = 0 terms, :926-928)."""
    MMAX = t.mmax if MMAX is None else int(MMAX)
    ms = [m for m in range(0, MMAX + 1) if not (no_odd and m % 2)]
    p0, p, fr, fp, fz, d0, d = _eof_eval_core(t, r, z, phi, accum_cos, accum_sin, ms, True)
    return p0, p0 + p, fr, fp, fz, d0, d0 + d


def eof_eval_particles(t: EOFTable, x, y, z, accum_cos, accum_sin, m1=0, m2=1000,
                       density=False):
    """accumulated_eval_particles, eof.py:989-1144, intended semantics (the
    reference reads Particles.data and Particles.zpos together, :1055 vs
    :1079, so it cannot run).  R = sqrt(x^2+y^2+1e-10) (:1070), phi =
    atan2(y,x); for m in [m1, m2]: the per-m sums of accumulated_eval
    (_eof_eval_core); the m=0 potential is returned as p0 and removed from p
    (:1129-1138).
    Returns (p0, p, fr, fp, fz, R) or, with density, (p0, p, d0, d, fr, fp, fz, R)."""
    x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
    R = (x * x + y * y + 1.0e-10) ** 0.5
    phi = np.arctan2(y, x)
    ms = [m for m in range(0, t.mmax + 1) if m1 <= m <= m2]
    p0, p, fr, fp, fz, d0, d = _eof_eval_core(t, R, z, phi, accum_cos, accum_sin, ms, density)
    if density:
        return p0, p, d0, d, fr, fp, fz, R
    return p0, p, fr, fp, fz, R


# ----------------------------------------------------------------------------
# R22: Fields.return_forces_cart  (potential.py:445-497)
# ----------------------------------------------------------------------------

class FieldSpec:
    """Frozen disc (EOF) + halo (SL) field, the state Fields carries after
    total_coefficients + prep_tables (potential.py:98-294)."""

    def __init__(self, eof: EOFTable, eof_cos, eof_sin, sl: SLTable, sl_coef,
                 halofac=1.0, no_odd=False, disk_m=None, disk_n=None,
                 halo_l=None, halo_n=None):
        self.eof, self.eof_cos, self.eof_sin = eof, np.asarray(eof_cos), np.asarray(eof_sin)
        self.sl, self.sl_coef = sl, np.asarray(sl_coef)
        self.halofac = float(halofac)
        self.no_odd = bool(no_odd)
        self.disk_m = eof.mmax if disk_m is None else int(disk_m)
        self.disk_n = eof.norder if disk_n is None else int(disk_n)
        self.halo_l = sl.lmax if halo_l is None else int(halo_l)
        self.halo_n = sl.nmax if halo_n is None else int(halo_n)


def return_forces_cart(F: FieldSpec, x, y, z, rotpos=0.0, compat=False):
    """potential.py:445-497, vectorised: R = sqrt(x^2+y^2)+1e-15,
    r = sqrt(R^2+z^2)+1e-15, disc eof.force_eval(R, z, phi+rotpos), halo
    spheresl.force_eval(r, z/r, phi+rotpos, halofac*expcoef); both azimuthal
    forces are zeroed where r < min(xi_halo) (:483-485).
    Returns (fx_d, fx_h, fy_d, fy_h, fz_d, fz_h, pot_d, pot_h)."""
    x, y, z = (np.atleast_1d(np.asarray(a, dtype=np.float64)) for a in (x, y, z))
    r2 = (x * x + y * y) ** 0.5 + 1.0e-15
    r3 = (r2 * r2 + z * z) ** 0.5 + 1.0e-15
    costh = z / r3
    phi = np.arctan2(y, x) + rotpos
    dfr, dfp, dfz, dp, _ = eof_force_eval(F.eof, r2, z, phi, F.eof_cos, F.eof_sin,
                                          MMAX=F.disk_m, NMAX=F.disk_n,
                                          no_odd=F.no_odd, compat=compat)
    hfr, hft, hfp, hp1, hp0 = sl_force_eval(F.sl, r3, costh, phi, F.halofac * F.sl_coef,
                                            lmax=F.halo_l, nmax=F.halo_n,
                                            no_odd=F.no_odd, compat=compat)
    guard = r3 < np.min(F.sl.xi)
    hfp = np.where(guard, 0.0, hfp)
    dfp = np.where(guard, 0.0, dfp)
    fxd = dfr * (x / r2) - dfp * (y / (r2 * r2))
    fxh = -1.0 * (hfr * (x / r3) - hft * (x * z / (r3 * r3 * r3))) + hfp * (y / (r2 * r2))
    fyd = dfr * (y / r2) + dfp * (x / (r2 * r2))
    fyh = -1.0 * (hfr * (y / r3) - hft * (y * z / (r3 * r3 * r3))) - hfp * (x / (r2 * r2))
    fzd = dfz
    fzh = -1.0 * (hfr * (z / r3) + hft * ((r2 * r2) / (r3 * r3 * r3)))
    return fxd, fxh, fyd, fyh, fzd, fzh, dp, hp1 + hp0


def total_force(F: FieldSpec, x, y, z, rotpos=0.0, compat=False):
    """Sum of disc and halo parts of return_forces_cart: (fx, fy, fz, pot)."""
    fxd, fxh, fyd, fyh, fzd, fzh, pd, ph = return_forces_cart(F, x, y, z, rotpos, compat)
    return fxd + fxh, fyd + fyh, fzd + fzh, pd + ph


# ----------------------------------------------------------------------------
# R23: leapfrog_integrate  (utils/integrate.py:53-190), vectorised over orbits
# ----------------------------------------------------------------------------

def leapfrog(F: FieldSpec, nint, dt, pos, vel, rotfreq=0.0, compat=False):
    """integrate.py:117-155 for many orbits at once (no apse stop):
    x_k+1 = x_k + v_k dt + a_k dt^2/2;  a_k+1 = F(x_k+1, 2 pi rotfreq t_k+1);
    v_k+1 = v_k + (a_k + a_k+1) dt/2.  Returns dict of (norb, nint) arrays
    T, X, Y, Z, VX, VY, VZ, P, FX, FY, FZ.  dt may be an (norb,) array of
    per-orbit step sizes (then T is (norb, nint) too)."""
    pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
    vel = np.atleast_2d(np.asarray(vel, dtype=np.float64))
    norb = pos.shape[0]
    if np.ndim(dt) == 0:
        times = np.arange(0, nint) * dt
        barpos = np.broadcast_to(2.0 * np.pi * rotfreq * times, (norb, nint))
    else:
        dt = np.asarray(dt, dtype=np.float64)
        times = dt[:, None] * np.arange(0, nint)[None, :]
        barpos = 2.0 * np.pi * rotfreq * times
    out = {k: np.zeros((norb, nint)) for k in ("X", "Y", "Z", "VX", "VY", "VZ", "P", "FX", "FY", "FZ")}
    x, y, z = pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()
    vx, vy, vz = vel[:, 0].copy(), vel[:, 1].copy(), vel[:, 2].copy()
    fx, fy, fz, p = total_force(F, x, y, z, barpos[:, 0], compat)
    for k, v in zip(("X", "Y", "Z", "VX", "VY", "VZ", "P", "FX", "FY", "FZ"),
                    (x, y, z, vx, vy, vz, p, fx, fy, fz)):
        out[k][:, 0] = v
    for step in range(1, nint):
        x = x + vx * dt + 0.5 * fx * dt ** 2.0
        y = y + vy * dt + 0.5 * fy * dt ** 2.0
        z = z + vz * dt + 0.5 * fz * dt ** 2.0
        nfx, nfy, nfz, p = total_force(F, x, y, z, barpos[:, step], compat)
        vx = vx + 0.5 * (fx + nfx) * dt
        vy = vy + 0.5 * (fy + nfy) * dt
        vz = vz + 0.5 * (fz + nfz) * dt
        fx, fy, fz = nfx, nfy, nfz
        for k, v in zip(("X", "Y", "Z", "VX", "VY", "VZ", "P", "FX", "FY", "FZ"),
                        (x, y, z, vx, vy, vz, p, fx, fy, fz)):
            out[k][:, step] = v
    out["T"] = times
    return out


# ----------------------------------------------------------------------------
# R24: compute_timestep and the orbit-grid drivers (utils/integrate.py:217-272,
# 760-967)
# ----------------------------------------------------------------------------

def compute_timestep(F: FieldSpec, pos, vel, dyn_res=100.0, compat_dt=True, compat=False):
    """integrate.py:217-272 for many start points (pos, vel (norb, 3)):
    eps = 1e-10 (:241); vtot = |v| + eps; forces from return_forces_cart at
    the unrotated start (:247); dtr = x.(fx, fy, fz') and
    atot = |(fx, fy, fz')| + eps with fz' = dfz*hfz when compat_dt (the
    reference's product, :249, :251), else dfz+hfz; ptot = |Phi_d + Phi_h|;
    criteria dtd = 1/sqrt(vtot+eps), dtv = sqrt(vtot/(atot+eps)),
    dta = ptot/(|dtr|+eps), dtA = sqrt(ptot/(atot^2+eps)) (:256-259); the
    smallest over dyn_res (:270).  Returns (norb,) dt and (4, norb) criteria."""
    pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
    vel = np.atleast_2d(np.asarray(vel, dtype=np.float64))
    eps = 1.e-10
    vtot = np.sum(vel ** 2.0, axis=1) ** 0.5 + eps
    dfx, hfx, dfy, hfy, dfz, hfz, dp, hp = return_forces_cart(F, pos[:, 0], pos[:, 1], pos[:, 2],
                                                              0.0, compat)
    fz = dfz * hfz if compat_dt else dfz + hfz
    dtr = pos[:, 0] * (dfx + hfx) + pos[:, 1] * (dfy + hfy) + pos[:, 2] * fz
    atot = ((dfx + hfx) ** 2.0 + (dfy + hfy) ** 2.0 + fz ** 2.0) ** 0.5 + eps
    ptot = np.abs(dp + hp)
    crit = np.array([1.0 / np.sqrt(vtot + eps), np.sqrt(vtot / (atot + eps)),
                     ptot / (np.abs(dtr) + eps), np.sqrt(ptot / (atot * atot + eps))])
    return np.min(crit, axis=0) / dyn_res, crit


def integrate_grid(F: FieldSpec, rads, vels, nint, dt, rotfreq=0.0, dyn_res=100.0, zs=None,
                   vzs=None, launch="x", compat_dt=True, compat=False):
    """integrate_grid / _launchy / _3D / _3D_launchy (integrate.py:760-922):
    for each rad, vel (, z, vz) in loop order, start [rad, 0, z] (launch x)
    or [0, rad, z] (launch y) with velocity [0, vel, vz]; dtime =
    max(compute_timestep, dt) (:871); the leapfrog for nint steps; rows
    X, Y, TX, TY, VX, VY, T (2D) or X, Y, Z, TX, TY, VX, VY, VZ, T (3D), with
    (TX, TY) the counter-clockwise rotation by barpos when barpos has a
    negative value, else the clockwise one (:180-188).  Returns the grid
    array (nrad, nvel, [nz, nvz,] 7 | 9, nint) and the per-orbit dtime."""
    three = zs is not None
    rads, vels = np.atleast_1d(rads).astype(np.float64), np.atleast_1d(vels).astype(np.float64)
    zl = np.atleast_1d(zs).astype(np.float64) if three else np.zeros(1)
    vzl = np.atleast_1d(vzs).astype(np.float64) if three else np.zeros(1)
    starts = []
    for rad in rads:
        for vel in vels:
            for z in zl:
                for vz in vzl:
                    p = [rad, 0.0, z] if launch == "x" else [0.0, rad, z]
                    starts.append((p, [0.0, vel, vz]))
    pos = np.array([s[0] for s in starts])
    vel = np.array([s[1] for s in starts])
    dtc, _ = compute_timestep(F, pos, vel, dyn_res, compat_dt, compat)
    dtime = np.maximum(dtc, dt)
    o = leapfrog(F, nint, dtime, pos, vel, rotfreq, compat)
    barpos = 2.0 * np.pi * rotfreq * o["T"]
    c, s = np.cos(barpos), np.sin(barpos)
    ccw = np.min(barpos, axis=1, keepdims=True) < 0.0
    TX = np.where(ccw, c * o["X"] - s * o["Y"], c * o["X"] + s * o["Y"])
    TY = np.where(ccw, s * o["X"] + c * o["Y"], -1.0 * s * o["X"] + c * o["Y"])
    rows = dict(o, TX=TX, TY=TY)
    keys = ("X", "Y", "Z", "TX", "TY", "VX", "VY", "VZ", "T") if three else \
        ("X", "Y", "TX", "TY", "VX", "VY", "T")
    arr = np.stack([rows[k] for k in keys], 1)
    shape = (rads.size, vels.size, zl.size, vzl.size) if three else (rads.size, vels.size)
    return arr.reshape(shape + arr.shape[1:]), dtime


# ----------------------------------------------------------------------------
# f4: RK4 integrator (utils/rk4.py:26-135), vectorised over orbits
# ----------------------------------------------------------------------------

def rk4(F: FieldSpec, nint, dt, pos, vel, omegab=0.0, frame_sign=1.0, compat=False):
    """rk4 / rk4_rotation (rk4.py:26-63, 94-135) with the field unrotated:
    k1 = (v, a(x, v)), k2 = (v + k1v dt/2, a(x + k1r dt/2, .)), k3 likewise
    from k2, k4 = (v + k3v dt, a(x + k3r dt, .)); v += dt/6 (k1v + 2 k2v +
    2 k3v + k4v), x likewise.  a = total force plus, for omegab != 0, the
    planar pseudo-force frame_sign * (omegab^2 R - 2 omegab z^ x v):
    frame_sign = +1 is Binney & Tremaine 3.117, -1 the reference's
    add_rotation_acceleration (rk4.py:66-90).  Returns dict of (norb, nint)
    arrays X, Y, Z, VX, VY, VZ."""
    r = np.atleast_2d(np.asarray(pos, dtype=np.float64)).copy()
    v = np.atleast_2d(np.asarray(vel, dtype=np.float64)).copy()
    norb = r.shape[0]
    out = {k: np.zeros((norb, nint)) for k in ("X", "Y", "Z", "VX", "VY", "VZ")}

    def acc(q, u):
        fx, fy, fz, _ = total_force(F, q[:, 0], q[:, 1], q[:, 2], 0.0, compat)
        a = np.stack([fx, fy, fz], 1)
        if omegab != 0.0:
            a[:, 0] += frame_sign * (omegab * omegab * q[:, 0] + 2.0 * omegab * u[:, 1])
            a[:, 1] += frame_sign * (omegab * omegab * q[:, 1] - 2.0 * omegab * u[:, 0])
        return a

    def rec(k):
        for j, key in enumerate(("X", "Y", "Z")):
            out[key][:, k] = r[:, j]
            out["V" + key][:, k] = v[:, j]

    rec(0)
    for k in range(1, nint):
        k1r, k1v = v, acc(r, v)
        k2r = v + k1v * dt * 0.5
        k2v = acc(r + k1r * dt * 0.5, k2r)
        k3r = v + k2v * dt * 0.5
        k3v = acc(r + k2r * dt * 0.5, k3r)
        k4r = v + k3v * dt
        k4v = acc(r + k3r * dt, k4r)
        v = v + (1 / 6) * dt * (k1v + 2 * k2v + 2 * k3v + k4v)
        r = r + (1 / 6) * dt * (k1r + 2 * k2r + 2 * k3r + k4r)
        rec(k)
    return out
