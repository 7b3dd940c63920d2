"""Spherical Sturm-Liouville (SL) basis: drop-in for exptool.basis.spheresl on the GPU.

Device work (libexptool_bfe.so through exptool_amd.bfe.BFESystem):
  compute_coefficients_solitary   spheresl.py:567-656
  compute_coefficients            spheresl.py:439-475 (one launch replaces the Pool)
  all_eval_particles              spheresl.py:1240-1362
  eval_particles                  spheresl.py:502-528
  force_eval                      spheresl.py:1107-1234 (one point or many per call)
  all_eval                        spheresl.py:987-1102 (one point or many per call)
Coefficient files (host):
  save_sl_coefficients / restore_sl_coefficients   spheresl.py:1390-1488
Return types follow the reference: float64 NumPy arrays and SL_Object.

Arithmetic quirks of the reference are reproduced only with compat=True:
den1/densfac of all_eval_particles (:1323, :1351), the first-cell derivative
stencil (:150-153) and the cos/sin(l phi) factor of force_eval (:1222-1225).
"""
from __future__ import annotations

import os

import numpy as np

from ..io.particle import particle_arrays
from ..utils import halo_methods

_SYSTEMS = {}


class SL_Object(object):
    """spheresl.py:1377-1386."""
    time = None
    dump = None
    comp = None
    nbodies = None
    lmax = None
    nmax = None
    sph_file = None
    model_file = None
    expcoef = None
    filename = None


def system_for(sph_file, model_file):
    """BFESystem holding the SL tables of (sph_file, model_file), cached per
    file pair and modification time."""
    from ..bfe import BFESystem
    key = tuple((os.path.abspath(f), os.path.getmtime(f)) for f in (sph_file, model_file))
    s = _SYSTEMS.get(key)
    if s is None:
        s = BFESystem(sl=halo_methods.load_sl(sph_file, model_file))
        _SYSTEMS.clear()
        _SYSTEMS[key] = s
    return s


def _data_dict(P):
    x, y, z, m = particle_arrays(P)
    return x, y, z, m


def compute_coefficients_solitary(ParticleInstance, sph_file, model_file, verbose=0, no_odd=False):
    """spheresl.py:567-656: expcoef ((lmax+1)^2, nmax) float64.  Odd l rows
    are zero with no_odd."""
    x, y, z, m = particle_arrays(ParticleInstance)
    s = system_for(sph_file, model_file)
    return s.sl_accumulate(x, y, z, m, no_odd=no_odd).cpu().numpy()


def compute_coefficients(PSPInput, sph_file, mod_file, verbose=1, no_odd=False):
    """spheresl.py:439-475: SL_Object of the particles of PSPInput."""
    x, y, z, m = particle_arrays(PSPInput)
    s = system_for(sph_file, mod_file)
    out = SL_Object()
    out.time = getattr(PSPInput, "time", None)
    out.filename = getattr(PSPInput, "filename", None)
    out.comp = getattr(PSPInput, "comp", None)
    out.nbodies = np.asarray(m).size
    out.sph_file = sph_file
    out.model_file = mod_file
    out.lmax = s.sl["lmax"]
    out.nmax = s.sl["nmax"]
    out.expcoef = s.sl_accumulate(x, y, z, m, no_odd=no_odd).cpu().numpy()
    return out


def _eval(s, x, y, z, expcoef, L1, L2, no_odd, compat, density=True, cos_l=False):
    o = s.sl_eval(x, y, z, np.asarray(expcoef, dtype=np.float64), L1=L1, L2=L2, no_odd=no_odd,
                  compat=compat, density=density, cos_l=cos_l)
    return {k: v.double().cpu().numpy() for k, v in o.items()}


def all_eval_particles(Particles, expcoef, sph_file, mod_file, verbose=0, L1=-1000, L2=1000,
                       NO_ODD=False, compat=False):
    """spheresl.py:1240-1362: (den0, den1, pot0, pot1, potr, pott, potp, R)
    with R the planar radius.  potr = dPhi/dr, pott = dPhi/dcos(theta),
    potp = dPhi/dphi."""
    x, y, z, _ = particle_arrays(Particles)
    g = _eval(system_for(sph_file, mod_file), x, y, z, expcoef, L1, L2, NO_ODD, compat)
    x64, y64 = np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)
    rr = (x64 * x64 + y64 * y64) ** 0.5
    return g["den0"], g["den1"], g["pot0"], g["pot1"], g["potr"], g["pott"], g["potp"], rr


def eval_particles(ParticleInstance, expcoef, sph_file, mod_file, nprocs=-1, verbose=1, l1=0,
                   l2=1000, compat=False):
    """spheresl.py:502-528 (one launch for any nprocs)."""
    return all_eval_particles(ParticleInstance, expcoef, sph_file, mod_file, verbose, L1=l1, L2=l2,
                              compat=compat)


def truncation(lmax, nmax, lmax_t, nmax_t):
    """Orders force_eval / all_eval use when asked for lmax, nmax against
    tables of lmax_t, nmax_t (spheresl.py:1134-1147, :1009-1022): l is cut
    when lmax < lmax_t, n when nmax < nmax_t - 1.  The reference fails on
    the other requests (nmax = nmax_t - 1 leaves expcoef and the radial
    matrices of different lengths; all_eval keeps lmax + 1 instead of
    (lmax + 1)^2 rows, :1017); here they mean the full table."""
    lmax, nmax = int(lmax), int(nmax)
    return min(lmax, lmax_t), (nmax if nmax < nmax_t - 1 else nmax_t)


def _truncated(expcoef, lmax, nmax, lmax_t, nmax_t):
    """expcoef ((lmax_t+1)^2, nmax_t) with the orders truncation() drops set
    to zero."""
    lu, nu = truncation(lmax, nmax, lmax_t, nmax_t)
    c = np.array(expcoef, dtype=np.float64, copy=True).reshape((lmax_t + 1) ** 2, nmax_t)
    c[(max(lu, 0) + 1) ** 2:, :] = 0.0
    c[:, max(nu, 0):] = 0.0
    return c


def _point_system(xi, p0, d0, cmap, scale, evtable, eftable):
    from ..bfe import system_for_tables
    ev = np.asarray(evtable, dtype=np.float64)
    ef = np.asarray(eftable, dtype=np.float64)
    lmax_t, nmax_t = ev.shape[0] - 1, ev.shape[1]
    tables = dict(xi=np.asarray(xi, dtype=np.float64), ev=ev, ef=ef,
                  p0=np.asarray(p0, dtype=np.float64), d0=np.asarray(d0, dtype=np.float64))
    scal = dict(lmax=lmax_t, nmax=nmax_t, numr=ef.shape[2], cmap=int(cmap), scale=float(scale))
    return system_for_tables("sl", tables, scal), lmax_t, nmax_t


def _coef_rows(expcoef, lmax_t, nmax_t):
    """expcoef as ((lmax_t+1)^2, nmax_t), missing rows/orders zero."""
    c = np.asarray(expcoef, dtype=np.float64)
    c = c.reshape(-1, c.shape[-1]) if c.ndim > 1 else c.reshape(1, -1)
    out = np.zeros(((lmax_t + 1) ** 2, nmax_t))
    nr, nc = min(c.shape[0], out.shape[0]), min(c.shape[1], nmax_t)
    out[:nr, :nc] = c[:nr, :nc]
    return out


def _point_eval(r, costh, phi, expcoef, xi, p0, d0, cmap, scale, lmax, nmax, evtable, eftable,
                no_odd, density, edge, cos_l):
    s, lmax_t, nmax_t = _point_system(xi, p0, d0, cmap, scale, evtable, eftable)
    lu, nu = truncation(lmax, nmax, lmax_t, nmax_t)
    scalar = all(np.ndim(v) == 0 for v in (r, costh, phi))
    r, ct, ph = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in (r, costh, phi)))
    out = s.sl_eval_sph(r.reshape(-1), ct.reshape(-1), ph.reshape(-1),
                        _coef_rows(expcoef, lmax_t, nmax_t), lmax_use=lu, nmax_use=nu,
                        no_odd=no_odd, density=density, edge=edge, cos_l=cos_l).cpu().numpy()
    if scalar:
        return [float(v[0]) for v in out]
    return [v.reshape(r.shape) for v in out]


def force_eval(r, costh, phi, expcoef, xi, p0, d0, cmap, scale, lmax, nmax, evtable, eftable,
               no_odd=False, verbose=0, compat=False):
    """spheresl.py:1107-1234 at one point or arrays of points (r the 3D
    radius as given): (potr, pott, potp, pot1, pot0) = dPhi/dr,
    dPhi/dcos(theta), dPhi/dphi, the l >= 1 and the l = 0 potential.
    compat=True gives the reference's cos/sin(l phi) azimuthal factor
    (:1222-1225) and its first-cell stencil (:150-153).  The tables are kept
    on the device between calls with the same arrays."""
    den0, den1, pot0, pot1, potr, pott, potp = _point_eval(
        r, costh, phi, expcoef, xi, p0, d0, cmap, scale, lmax, nmax, evtable, eftable,
        no_odd, False, compat, compat)
    return potr, pott, potp, pot1, pot0


def all_eval(r, costh, phi, expcoef, xi, p0, d0, cmap, scale, lmax, nmax, evtable, eftable,
             no_odd=False, verbose=0, compat=False):
    """spheresl.py:987-1102 at one point or arrays of points: (den0, den1,
    pot0, pot1, potr, pott, potp) with den1 the total density (the l = 0 term
    included, :1046), densities times 0.25/pi (:1092) and the cos/sin(m phi)
    azimuthal factor.  compat=True gives the first-cell stencil of
    get_halo_dens_pot_force (:150-153), which the reference evaluates."""
    den0, den1, pot0, pot1, potr, pott, potp = _point_eval(
        r, costh, phi, expcoef, xi, p0, d0, cmap, scale, lmax, nmax, evtable, eftable,
        no_odd, True, compat, False)
    if isinstance(den0, float):
        return den0, den0 + den1, pot0, pot1, potr, pott, potp
    return den0, den0 + den1, pot0, pot1, potr, pott, potp


# ----------------------------------------------------------------------------
# coefficient files (spheresl.py:1390-1488)
# ----------------------------------------------------------------------------

_HDR = 4 + 100 + 8 + 4 + 100 + 100 + 8


def _text(v, n):
    v = b"" if v is None else (v if isinstance(v, bytes) else str(v).encode())
    return np.array([v], dtype=f"S{n}")


def sl_coefficients_to_file(f, obj):
    """One dump: time f4, filename S100, comp S8, nbodies i4, sph_file S100,
    model_file S100, lmax nmax i4, expcoef f8 ((lmax+1)^2 x nmax)."""
    np.array([obj.time if obj.time is not None else 0.0], dtype="<f4").tofile(f)
    _text(obj.filename, 100).tofile(f)
    _text(obj.comp, 8).tofile(f)
    np.array([obj.nbodies or 0], dtype="<i4").tofile(f)
    _text(obj.sph_file, 100).tofile(f)
    _text(obj.model_file, 100).tofile(f)
    np.array([obj.lmax, obj.nmax], dtype="<i4").tofile(f)
    np.asarray(obj.expcoef, dtype="<f8").reshape(-1).tofile(f)


def save_sl_coefficients(outfile, obj, verbose=0):
    """Append one dump to outfile (leading i4 dump count, then fixed-size dumps)."""
    if not os.path.exists(outfile):
        with open(outfile, "wb") as f:
            np.array([0], dtype="<i4").tofile(f)
    with open(outfile, "rb+") as f:
        nd = int(np.fromfile(f, dtype="<i4", count=1)[0]) + 1
        f.seek(0)
        np.array([nd], dtype="<i4").tofile(f)
        size = 8 * (obj.lmax * (obj.lmax + 2) + 1) * obj.nmax + _HDR
        f.seek(4 + (nd - 1) * size)
        sl_coefficients_to_file(f, obj)


def extract_sl_coefficients(f):
    o = SL_Object()
    o.time = float(np.fromfile(f, dtype="<f4", count=1)[0])
    o.filename = np.fromfile(f, dtype="S100", count=1)[0]
    o.comp = np.fromfile(f, dtype="S8", count=1)[0]
    o.nbodies = int(np.fromfile(f, dtype="<i4", count=1)[0])
    o.sph_file = np.fromfile(f, dtype="S100", count=1)[0]
    o.model_file = np.fromfile(f, dtype="S100", count=1)[0]
    o.lmax, o.nmax = (int(v) for v in np.fromfile(f, dtype="<i4", count=2))
    nrow = o.lmax * (o.lmax + 2) + 1
    o.expcoef = np.fromfile(f, dtype="<f8", count=nrow * o.nmax).reshape(nrow, o.nmax)
    return o


def restore_sl_coefficients(infile):
    """(last SL_Object, {round(time, 3): SL_Object})."""
    with open(infile, "rb") as f:
        nd = int(np.fromfile(f, dtype="<i4", count=1)[0])
        out = {}
        o = None
        for _ in range(nd):
            o = extract_sl_coefficients(f)
            out[np.round(o.time, 3)] = o
    return o, out
