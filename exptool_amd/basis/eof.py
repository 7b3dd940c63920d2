"""Cylindrical EOF basis: drop-in for exptool.basis.eof on the GPU.

Table I/O (host):
  eof_params        eof.py:98-200   (new YAML header 0x0C0A57A1 or old 76-byte header)
  parse_eof         eof.py:224-313
  set_table_params  eof.py:316-347
  write_eof_cache   (writer of the same layout; the reference has none)
Device work (libexptool_bfe.so through exptool_amd.bfe.BFESystem):
  accumulate               eof.py:492-640 (VAR subsample sums: one batched launch)
  compute_coefficients     eof.py:1158-1260
  compute_variance / compute_sn   eof.py:2135-2226 (host, coefficient-sized)
  accumulated_eval_particles / compute_forces   eof.py:989-1144, :1263-1317
  force_eval               eof.py:756-872 (one point or many per call)
  accumulated_eval         eof.py:874-929 (one point or many per call)
Coefficient files (host): save_eof_coefficients / restore_eof_coefficients
  eof.py:1646-1763
Return types follow the reference: float64 NumPy arrays and EOF_Object.
"""
from __future__ import annotations

import os

import numpy as np
import yaml

from .compatibility import r_to_xi, z_to_y

HMAGIC = 0x0C0A57A1


# ----------------------------------------------------------------------------
# cache file I/O
# ----------------------------------------------------------------------------

def _read_header(f):
    tmagic = np.fromfile(f, dtype="<i4", count=1)
    if tmagic.size and tmagic[0] == HMAGIC:
        ssize = int(np.fromfile(f, dtype="<i4", count=1)[0])
        data = yaml.safe_load(f.read(ssize).decode())
        cmap = data.get("cmap", data.get("cmapr"))
        hdr = dict(mmax=int(data["mmax"]), numx=int(data["numx"]), numy=int(data["numy"]),
                   nmax=int(data["nmax"]), norder=int(data["norder"]), dens=int(data["dens"]),
                   cmap=int(cmap), rmin=float(data["rmin"]), rmax=float(data["rmax"]),
                   ascale=float(data["ascl"]), hscale=float(data["hscl"]),
                   cylmass=float(data["cmass"]), time=float(data["time"]),
                   cmapz=int(data.get("cmapz", 0)), offset=8 + ssize, newstyle=True)
    else:
        f.seek(0)
        a = np.fromfile(f, dtype="<u4", count=7)
        b = np.fromfile(f, dtype="<f8", count=6)
        hdr = dict(mmax=int(a[0]), numx=int(a[1]), numy=int(a[2]), nmax=int(a[3]),
                   norder=int(a[4]), dens=int(a[5]), cmap=int(a[6]), rmin=float(b[0]),
                   rmax=float(b[1]), ascale=float(b[2]), hscale=float(b[3]),
                   cylmass=float(b[4]), time=float(b[5]), cmapz=0, offset=76, newstyle=False)
    return hdr


def eof_header(file):
    with open(file, "rb") as f:
        return _read_header(f)


def eof_params(file, verbose=0):
    """eof.py:98-200: returns rmin, rmax, numx, numy, mmax, norder, ascale,
    hscale, cmap, dens."""
    h = eof_header(file)
    if verbose:
        print("eof.eof_params:", {k: v for k, v in h.items() if k != "offset"})
    return (h["rmin"], h["rmax"], h["numx"], h["numy"], h["mmax"], h["norder"], h["ascale"],
            h["hscale"], h["cmap"], h["dens"])


def parse_eof(file, verbose=0):
    """eof.py:224-313: potC, rforcec, zforcec, densc, potS, rforces, zforces,
    denss, each (mmax+1, norder, numx+1, numy+1)."""
    h = eof_header(file)
    mmax, norder, numx, numy, dens = h["mmax"], h["norder"], h["numx"], h["numy"], h["dens"]
    shape = (mmax + 1, norder, numx + 1, numy + 1)
    out = [np.zeros(shape) for _ in range(8)]
    plane = (numx + 1) * (numy + 1)
    nq = 4 if dens == 1 else 3
    raw = np.fromfile(file, dtype="<f8", offset=h["offset"])
    pos = 0
    for kind, mlist in ((0, range(0, mmax + 1)), (4, range(1, mmax + 1))):
        for i in mlist:
            for j in range(norder):
                for q in range(nq):
                    out[kind + q][i, j] = raw[pos:pos + plane].reshape(numx + 1, numy + 1)
                    pos += plane
    return tuple(out)


def write_eof_cache(file, t, newstyle=True):
    """Write tables (dict from tablegen.make_eof_tables) in the EmpCylSL cache
    layout parse_eof reads: header, then for m=0..mmax, n<norder: potC,
    rforceC, zforceC[, densC]; then m=1..mmax: the S tables."""
    dens = int(t.get("dens", 1))
    with open(file, "wb") as f:
        if newstyle:
            hdr = dict(mmax=int(t["mmax"]), numx=int(t["numx"]), numy=int(t["numy"]),
                       nmax=int(t.get("nmax", t["norder"])), norder=int(t["norder"]), dens=dens,
                       cmap=int(t["cmap"]), rmin=float(t["rmin"]), rmax=float(t["rmax"]),
                       ascl=float(t["ascale"]), hscl=float(t["hscale"]),
                       cmass=float(t.get("cylmass", 1.0)), time=float(t.get("time", 0.0)))
            b = yaml.safe_dump(hdr).encode()
            np.array([HMAGIC, len(b)], dtype="<i4").tofile(f)
            f.write(b)
        else:
            np.array([t["mmax"], t["numx"], t["numy"], t.get("nmax", t["norder"]), t["norder"],
                      dens, t["cmap"]], dtype="<u4").tofile(f)
            np.array([t["rmin"], t["rmax"], t["ascale"], t["hscale"], t.get("cylmass", 1.0),
                      t.get("time", 0.0)], dtype="<f8").tofile(f)
        groups = ("pot", "rforce", "zforce", "dens")[:4 if dens == 1 else 3]
        for sfx, mlist in (("C", range(0, t["mmax"] + 1)), ("S", range(1, t["mmax"] + 1))):
            for i in mlist:
                for j in range(t["norder"]):
                    for g in groups:
                        np.asarray(t[g + sfx][i, j], dtype="<f8").tofile(f)


def load_eof(eof_file):
    """Read an EOF cache into the table dict used by BFESystem."""
    h = eof_header(eof_file)
    potC, rforceC, zforceC, densC, potS, rforceS, zforceS, densS = parse_eof(eof_file)
    return dict(mmax=h["mmax"], norder=h["norder"], numx=h["numx"], numy=h["numy"],
                cmap=h["cmap"], rmin=h["rmin"], rmax=h["rmax"], ascale=h["ascale"],
                hscale=h["hscale"], dens=h["dens"], potC=potC, rforceC=rforceC,
                zforceC=zforceC, densC=densC if h["dens"] == 1 else None, potS=potS,
                rforceS=rforceS, zforceS=zforceS, densS=densS if h["dens"] == 1 else None)


def set_table_params(RMAX=20.0, RMIN=0.001, ASCALE=0.01, HSCALE=0.001, NUMX=128, NUMY=64, CMAP=0):
    """eof.py:316-347."""
    Rtable = np.sqrt(0.5) * RMAX
    XMIN = r_to_xi(RMIN * ASCALE, CMAP, ASCALE)
    XMAX = r_to_xi(Rtable * ASCALE, CMAP, ASCALE)
    dX = (XMAX - XMIN) / NUMX
    YMIN = z_to_y(-Rtable * ASCALE, HSCALE)
    YMAX = z_to_y(Rtable * ASCALE, HSCALE)
    dY = (YMAX - YMIN) / NUMY
    return XMIN, XMAX, dX, YMIN, YMAX, dY


# ----------------------------------------------------------------------------
# objects
# ----------------------------------------------------------------------------

class EOF_Object(object):
    """eof.py:1627-1636."""
    time = None
    dump = None
    comp = None
    nbodies = None
    eof_file = None
    mmax = None
    nmax = None
    cos = None
    sin = None


_SYSTEMS = {}


def system_for(eof_file):
    """BFESystem holding the tables of eof_file (cached per file and mtime)."""
    from ..bfe import BFESystem
    key = (os.path.abspath(eof_file), os.path.getmtime(eof_file))
    s = _SYSTEMS.get(key)
    if s is None:
        s = BFESystem(eof=load_eof(eof_file))
        _SYSTEMS.clear()
        _SYSTEMS[key] = s
    return s


def _particle_arrays(P):
    """x, y, z, m from a particle.holder (xpos..mass) or an object with .data."""
    if hasattr(P, "data") and isinstance(getattr(P, "data"), dict):
        d = P.data
        return d["x"], d["y"], d["z"], d.get("m")
    return P.xpos, P.ypos, P.zpos, getattr(P, "mass", None)


def _nan_to_origin(x, y, z):
    """eof.py:1190-1204: particles with NaN coordinates are moved to the origin."""
    x, y, z = (np.array(a, dtype=np.float64, copy=True) for a in (x, y, z))
    bad = np.isnan(x) | np.isnan(y) | np.isnan(z)
    if bad.any():
        x[bad] = 0.0
        y[bad] = 0.0
        z[bad] = 0.0
    return x, y, z


# ----------------------------------------------------------------------------
# device-backed API
# ----------------------------------------------------------------------------

def accumulate_tables(system, x, y, z, m, VAR=0):
    """Coefficients (cos, sin) as float64 numpy arrays from a BFESystem; with
    VAR > 0 also the subsample sums (cos2, sin2), each (VAR, mmax+1, norder)."""
    from .._lib import f32
    x, y, z, m = (f32(np.asarray(a, dtype=np.float64), system.device) for a in (x, y, z, m))
    c, s = system.eof_accumulate(x, y, z, m)
    if not VAR:
        return c.cpu().numpy(), s.cpu().numpy()
    c2, s2 = variance_samples(system, x, y, z, m, VAR)
    return c.cpu().numpy(), s.cpu().numpy(), c2, s2


def variance_samples(system, x, y, z, m, VAR):
    """The VAR branch of eof.accumulate (eof.py:554-574): VAR subsamples of
    floor(sqrt(n)) particles drawn with replacement by np.random.randint, in
    the reference's call order on NumPy's global generator (so np.random.seed
    gives the reference's draws), each summed like the full accumulate
    (upscale 1).  The subsamples are gathered on the device and accumulated
    in one batched launch.  x, y, z, m are float32 device tensors.  Returns
    (cos2, sin2) float64 (VAR, mmax+1, norder) numpy arrays."""
    import torch
    VAR = int(VAR)
    n = x.numel()
    k = int(np.floor(np.sqrt(n)))
    use = np.concatenate([np.random.randint(n, size=k) for _ in range(VAR)]).astype(np.int64)
    idx = torch.from_numpy(use).to(system.device)
    g = [t.index_select(0, idx) for t in (x, y, z, m)]
    c2, s2 = system.eof_accumulate_batch(*g, [k] * VAR)
    return c2.cpu().numpy(), s2.cpu().numpy()


def compute_coefficients(PSPInput, eof_file, verbose=1, no_odd=False, nprocs_max=-1, VAR=False,
                         nanblock=False):
    """eof.py:1158-1260 on the GPU.  NaN particles are moved to the origin
    (the reference intends this at :1190-1204 but raises under NumPy >= 2);
    nanblock=True raises instead.  no_odd has no effect, as in the reference
    accumulate (eof.py:545-551).  VAR > 0 (True counts as 1) adds the
    subsample sums .cos2 / .sin2 of accumulate's VAR branch (eof.py:1256-1258,
    the single-process form :1245-1246)."""
    x, y, z, m = _particle_arrays(PSPInput)
    if nanblock and (np.isnan(x) | np.isnan(y) | np.isnan(z)).any():
        raise ValueError(f"eof.compute_coefficients: NaN values in {getattr(PSPInput, 'filename', '')}")
    x, y, z = _nan_to_origin(x, y, z)
    sysm = system_for(eof_file)
    out = EOF_Object()
    out.time = getattr(PSPInput, "time", None)
    out.filename = getattr(PSPInput, "filename", None)
    out.comp = getattr(PSPInput, "comp", None)
    out.nbodies = np.asarray(m).size
    out.eof_file = eof_file
    out.mmax = sysm.eof["mmax"]
    out.nmax = sysm.eof["norder"]
    if VAR:
        out.cos, out.sin, out.cos2, out.sin2 = accumulate_tables(sysm, x, y, z, m, VAR=int(VAR))
    else:
        out.cos, out.sin = accumulate_tables(sysm, x, y, z, m)
    return out


def accumulate(ParticleInstance, potC, potS, MMAX, NMAX, XMIN, dX, YMIN, dY, NUMX, NUMY, ASCALE,
               HSCALE, CMAP, verbose=0, no_odd=False, VAR=0):
    """eof.py:492-640 with tables passed as arrays (kept on the device
    between calls with the same arrays and grid).  VAR > 0 returns (cos,
    sin, cos2, sin2) as the reference (eof.py:554-574).  no_odd has no
    effect, as in the reference (:545-551)."""
    sysm = _point_system(dict(potC=potC, potS=potS), XMIN, dX, YMIN, dY, NUMX, NUMY, ASCALE,
                         HSCALE, CMAP)
    x, y, z, m = _particle_arrays(ParticleInstance)
    if VAR:
        c, sn, c2, s2 = accumulate_tables(sysm, x, y, z, m, VAR=int(VAR))
        return (c[:MMAX + 1, :NMAX], sn[:MMAX + 1, :NMAX], c2[:, :MMAX + 1, :NMAX],
                s2[:, :MMAX + 1, :NMAX])
    c, sn = accumulate_tables(sysm, x, y, z, m)
    return c[:MMAX + 1, :NMAX], sn[:MMAX + 1, :NMAX]


def compute_variance(ParticleInstance, accum_cos, accum_sin, accum_cos2, accum_sin2):
    """eof.py:2135-2193: with w = 1/sum(m), varC = cos2 w^2 - (w cos)^2 / n
    and b_Hall facC = (w cos)^2 / (varC/(n+1) + (w cos)^2 + 1e-10), the same
    for sin.  Returns (varC, varS, facC, facS)."""
    m = np.asarray(_particle_arrays(ParticleInstance)[3], dtype=np.float64)
    wgt = 1. / np.sum(m)
    nrm = wgt * wgt
    srm = 1. / float(m.size)
    sqrC = (np.asarray(accum_cos) * wgt) ** 2
    sqrS = (np.asarray(accum_sin) * wgt) ** 2
    varC = np.asarray(accum_cos2) * nrm - srm * sqrC
    varS = np.asarray(accum_sin2) * nrm - srm * sqrS
    facC = sqrC / (varC / (float(m.size) + 1.) + sqrC + 1.0e-10)
    facS = sqrS / (varS / (float(m.size) + 1.) + sqrS + 1.0e-10)
    return varC, varS, facC, facS


def compute_sn(ParticleInstance, accum_cos, accum_sin, accum_cos2, accum_sin2):
    """eof.py:2198-2226: signal to noise (coef^2 / var)^(1/2) with the
    variance of compute_variance.  Returns (snC, snS)."""
    varC, varS, _, _ = compute_variance(ParticleInstance, accum_cos, accum_sin, accum_cos2,
                                        accum_sin2)
    c, s = np.asarray(accum_cos), np.asarray(accum_sin)
    return (c * c / varC) ** 0.5, (s * s / varS) ** 0.5


# ----------------------------------------------------------------------------
# coefficient files (eof.py:1646-1741)
# ----------------------------------------------------------------------------

_EOF_HDR = 4 + 100 + 8 + 4 + 100 + 8      # time, filename, comp, nbodies, eof_file, mmax/nmax


def eof_coefficients_to_file(f, EOF_Object):
    """eof.py:1646-1668: f4 time, S100 filename, S8 comp, i4 nbodies,
    S100 eof_file, i4 mmax, nmax, then f8 cos and sin ((mmax+1) x nmax)."""
    np.array([EOF_Object.time], dtype='<f4').tofile(f)
    np.array([EOF_Object.filename], dtype='S100').tofile(f)
    np.array([EOF_Object.comp], dtype='S8').tofile(f)
    np.array([EOF_Object.nbodies], dtype='<i4').tofile(f)
    np.array([EOF_Object.eof_file], dtype='S100').tofile(f)
    np.array([EOF_Object.mmax, EOF_Object.nmax], dtype='<i4').tofile(f)
    np.asarray(EOF_Object.cos, dtype='<f8').reshape(-1).tofile(f)
    np.asarray(EOF_Object.sin, dtype='<f8').reshape(-1).tofile(f)


def save_eof_coefficients(outfile, EOF_Object, verbose=0):
    """eof.py:1672-1705: append one dump to outfile, whose leading i4 counts
    the dumps (each 16 (mmax+1) nmax + 224 bytes)."""
    if not os.path.exists(outfile):
        with open(outfile, 'wb') as f:
            np.array([0], dtype='<i4').tofile(f)
    with open(outfile, 'rb+') as f:
        nd = int(np.fromfile(f, dtype='<i4', count=1)[0]) + 1
        f.seek(0)
        np.array([nd], dtype='<i4').tofile(f)
        if verbose:
            print('eof.save_eof_coefficients: coefficient file currently has {0:d} dumps.'.format(nd))
        f.seek(4 + (nd - 1) * (16 * (EOF_Object.mmax + 1) * EOF_Object.nmax + _EOF_HDR))
        eof_coefficients_to_file(f, EOF_Object)


def extract_eof_coefficients(f):
    """eof.py:1743-1763: one dump from an open coefficient file."""
    o = EOF_Object()
    o.time = np.fromfile(f, dtype='<f4', count=1)[0]
    o.filename = np.fromfile(f, dtype='S100', count=1)[0]
    o.comp = np.fromfile(f, dtype='S8', count=1)[0]
    o.nbodies = int(np.fromfile(f, dtype='<i4', count=1)[0])
    o.eof_file = np.fromfile(f, dtype='S100', count=1)[0]
    o.mmax, o.nmax = (int(v) for v in np.fromfile(f, dtype='<i4', count=2))
    k = (o.mmax + 1) * o.nmax
    o.cos = np.fromfile(f, dtype='<f8', count=k).reshape(o.mmax + 1, o.nmax)
    o.sin = np.fromfile(f, dtype='<f8', count=k).reshape(o.mmax + 1, o.nmax)
    return o


def restore_eof_coefficients(infile):
    """eof.py:1708-1740: (last EOF_Object, {time: EOF_Object}) in file
    order."""
    from collections import OrderedDict
    with open(infile, 'rb') as f:
        nd = int(np.fromfile(f, dtype='<i4', count=1)[0])
        out = OrderedDict()
        o = None
        for _ in range(nd):
            o = extract_eof_coefficients(f)
            out[o.time] = o
    return o, out


def compute_forces(PSPInput, EOF_Object, verbose=1, nprocs=-1, m1=0, m2=1000, density=False):
    """eof.py:1263-1317 on the GPU: returns p0, p, fr, fp, fz, r (float64)
    or, with density, p0, p, d0, d, fr, fp, fz, r."""
    x, y, z, _ = _particle_arrays(PSPInput)
    s = system_for(EOF_Object.eof_file)
    o = s.eof_eval(x, y, z, EOF_Object.cos, EOF_Object.sin, m1=m1, m2=m2, density=density)
    g = {k: v.double().cpu().numpy() for k, v in o.items()}
    x, y = np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)
    R = (x * x + y * y + 1.e-10) ** 0.5
    if density:
        return g["p0"], g["p"], g["d0"], g["d"], g["fr"], g["fp"], g["fz"], R
    return g["p0"], g["p"], g["fr"], g["fp"], g["fz"], R


def accumulated_eval_particles(Particles, accum_cos, accum_sin, eof_file="", m1=0, m2=1000,
                               verbose=1, density=False, **kw):
    """eof.py:989-1144 on the GPU (tables given by eof_file)."""
    if not eof_file:
        raise ValueError("exptool_amd accumulated_eval_particles needs eof_file")
    obj = EOF_Object()
    obj.eof_file, obj.cos, obj.sin = eof_file, accum_cos, accum_sin
    return compute_forces(Particles, obj, m1=m1, m2=m2, density=density)


# ----------------------------------------------------------------------------
# point evaluators (eof.py:756-929)
# ----------------------------------------------------------------------------

def _point_system(tabs, rmin, dR, zmin, dZ, numx, numy, ASCALE, HSCALE, CMAP):
    """BFESystem holding the tables passed to a point evaluator (kept between
    calls with the same arrays and grid)."""
    from ..bfe import system_for_tables
    potC = np.asarray(tabs["potC"], dtype=np.float64)
    if potC.shape[2:] != (int(numx) + 1, int(numy) + 1):
        raise ValueError(f"eof tables are {potC.shape[2:]} nodes, not numx+1, numy+1 = "
                         f"{(int(numx) + 1, int(numy) + 1)}")
    tables = {k: (None if v is None else np.asarray(v, dtype=np.float64)) for k, v in tabs.items()}
    dens = int(tables.get("densC") is not None and tables.get("densS") is not None)
    scal = dict(mmax=potC.shape[0] - 1, norder=potC.shape[1], numx=int(numx), numy=int(numy),
                cmap=int(CMAP), ascale=float(ASCALE), hscale=float(HSCALE), XMIN=float(rmin),
                dX=float(dR), YMIN=float(zmin), dY=float(dZ), dens=dens)
    return system_for_tables("eof", tables, scal)


def _coef_rows(a, mmax, norder):
    c = np.asarray(a, dtype=np.float64)
    c = c.reshape(-1, c.shape[-1]) if c.ndim > 1 else c.reshape(1, -1)
    out = np.zeros((mmax + 1, norder))
    nr, nc = min(c.shape[0], mmax + 1), min(c.shape[1], norder)
    out[:nr, :nc] = c[:nr, :nc]
    return out


def _point_eval(sysm, r, z, phi, accum_cos, accum_sin, MMAX, NMAX, no_odd, compat_mask,
                density):
    mmax, norder = sysm.eof["mmax"], sysm.eof["norder"]
    scalar = all(np.ndim(v) == 0 for v in (r, z, phi))
    r, z, ph = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in (r, z, phi)))
    out = sysm.eof_eval_cyl(r.reshape(-1), z.reshape(-1), ph.reshape(-1),
                            _coef_rows(accum_cos, mmax, norder),
                            _coef_rows(accum_sin, mmax, norder), mmax_use=int(MMAX),
                            nmax_use=int(NMAX), no_odd=no_odd, compat_mask=compat_mask,
                            density=density).cpu().numpy()
    if scalar:
        return [float(v[0]) for v in out]
    return [v.reshape(r.shape) for v in out]


def force_eval(r, z, phi, accum_cos, accum_sin, potC, rforceC, zforceC, potS, rforceS, zforceS,
               rmin=0, dR=0, zmin=0, dZ=0, numx=0, numy=0, fac=1.0, MMAX=6, NMAX=18, ASCALE=0.0,
               HSCALE=0.0, CMAP=0, no_odd=False, perturb=False, compat=False):
    """eof.py:756-872 at one point or arrays of points (r the planar radius as
    given): (fr+fr0, fp, fz+fz0, p+p0, p0), fp = -dPhi/dphi, or with perturb
    (fr, fp, fz, p, p0, fr0, fz0) with the m >= 1 parts and the m = 0 ones
    apart.  Orders m > MMAX and n >= NMAX are dropped.  no_odd drops odd m;
    compat=True weights them by the reference's |cos(pi m/2)| (~6e-17,
    :833) instead.  fac is unused, as in the reference.  The tables are kept
    on the device between calls with the same arrays and grid."""
    sysm = _point_system(dict(potC=potC, rforceC=rforceC, zforceC=zforceC, potS=potS,
                              rforceS=rforceS, zforceS=zforceS),
                         rmin, dR, zmin, dZ, numx, numy, ASCALE, HSCALE, CMAP)
    p0, p, fr, fp, fz, _, _, fr0, fz0 = _point_eval(sysm, r, z, phi, accum_cos, accum_sin,
                                                    MMAX, NMAX, no_odd, compat, False)
    if perturb:
        return fr - fr0, fp, fz - fz0, p, p0, fr0, fz0
    return fr, fp, fz, p + p0, p0


def accumulated_eval(r, z, phi, accum_cos, accum_sin, potC, rforceC, zforceC, densC, potS,
                     rforceS, zforceS, densS, rmin=0, dR=0, zmin=0, dZ=0, numx=0, numy=0,
                     fac=1.0, MMAX=6, NMAX=18, ASCALE=0.0, HSCALE=0.0, CMAP=0, no_odd=False):
    """eof.py:874-929 at one point or arrays of points: (p0, p, fr, fp, fz,
    d0, d) with p and d the totals over m = 0..MMAX and p0, d0 the m = 0
    terms.  Every radial order is used (NMAX has no effect, as in the
    reference); no_odd skips odd m."""
    sysm = _point_system(dict(potC=potC, rforceC=rforceC, zforceC=zforceC, densC=densC,
                              potS=potS, rforceS=rforceS, zforceS=zforceS, densS=densS),
                         rmin, dR, zmin, dZ, numx, numy, ASCALE, HSCALE, CMAP)
    p0, p, fr, fp, fz, d0, d, _, _ = _point_eval(sysm, r, z, phi, accum_cos, accum_sin, MMAX,
                                                 -1, no_odd, False, True)
    return p0, p0 + p, fr, fp, fz, d0, d0 + d
