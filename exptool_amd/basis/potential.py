"""Frozen disc (EOF) + halo (SL) field: drop-in for exptool.basis.potential.Fields.

  Fields.__init__                 potential.py:56-96
  Fields.total_coefficients       potential.py:98-252 (PSP file or particle objects)
  Fields.prep_tables              potential.py:257-294
  Fields.set_field_parameters     potential.py:365-378
  Fields.reset_field_parameters   potential.py:380-387
  Fields.return_forces_cart       potential.py:445-497 (libexptool_bfe field kernel)
  Fields.return_forces_cyl        potential.py:389-440 (batch evaluators + host assembly)
  Fields.save_field, restore_field   potential.py:738-1041 (binary field file)
  get_fields                      potential.py:1243-1289

The tables and the folded field live in one BFESystem on the device; the field
is re-frozen (bfe_field_set) whenever coefficients or field parameters change.
"""
from __future__ import annotations

import numpy as np

from ..io import particle
from ..utils import halo_methods
from . import eof as eof_mod
from . import spheresl


class Fields:
    """Disc + halo field of one snapshot (potential.py:49-735, the hot-path part)."""

    def __init__(self, infile, eof_file, sph_file, model_file, nhalo=1000000, transform=False,
                 no_odd=False, centering=False, mutual_center=False, verbose=1, compat=False,
                 device=0):
        self.filename = infile
        self.eof_file = eof_file
        self.sph_file = sph_file
        self.model_file = model_file
        self.nhalo = nhalo
        self.transform = transform
        self.no_odd = no_odd
        self.centering = centering
        self.mutual_center = mutual_center
        self.verbose = verbose
        self.compat = compat
        self.device = device
        self.system = None
        self._frozen = None

    # ------------------------------------------------------------ construction
    @classmethod
    def from_tables(cls, eof_tables, sl_tables, eof_cos, eof_sin, sl_coef, halofac=1.0,
                    no_odd=False, compat=False, device=0):
        """Fields from in-memory tables (tablegen / load_eof / load_sl dicts)
        and coefficients, without files."""
        F = cls(None, None, None, None, no_odd=no_odd, compat=compat, device=device)
        F._set_tables(eof_tables, sl_tables)
        F.EOF = eof_mod.EOF_Object()
        F.EOF.cos, F.EOF.sin = np.asarray(eof_cos, np.float64), np.asarray(eof_sin, np.float64)
        F.EOF.mmax, F.EOF.nmax = F.mmax, F.norder
        F.SL = spheresl.SL_Object()
        F.SL.expcoef = np.asarray(sl_coef, np.float64)
        F.SL.lmax, F.SL.nmax = F.lmaxhalo, F.nmaxhalo
        F.halofac = float(halofac)
        F.reset_field_parameters()
        F.no_odd = no_odd
        return F

    def _set_tables(self, et, st):
        from ..bfe import BFESystem
        self.system = BFESystem(sl=st, eof=et, device=self.device)
        self.potC, self.rforceC, self.zforceC = et["potC"], et["rforceC"], et["zforceC"]
        self.potS, self.rforceS, self.zforceS = et["potS"], et["rforceS"], et["zforceS"]
        self.densC, self.densS = et.get("densC"), et.get("densS")
        self.rmindisk, self.rmaxdisk = et["rmin"], et["rmax"]
        self.numx, self.numy, self.mmax, self.norder = et["numx"], et["numy"], et["mmax"], et["norder"]
        self.ascale, self.hscale, self.cmapdisk = et["ascale"], et["hscale"], et["cmap"]
        self.densdisk = et.get("dens", 0)
        e = self.system.eof
        self.XMIN, self.dX, self.YMIN, self.dY = e["XMIN"], e["dX"], e["YMIN"], e["dY"]
        self.XMAX, self.YMAX = self.XMIN + self.numx * self.dX, self.YMIN + self.numy * self.dY
        self.disk_use_m, self.disk_use_n = self.mmax, self.norder
        self.lmaxhalo, self.nmaxhalo, self.numrhalo = st["lmax"], st["nmax"], st["numr"]
        self.cmaphalo, self.scalehalo = st["cmap"], st["scale"]
        self.rminhalo, self.rmaxhalo = st.get("rmin"), st.get("rmax")
        self.evtablehalo, self.eftablehalo = st["ev"], st["ef"]
        self.xihalo, self.p0halo, self.d0halo = st["xi"], st["p0"], st["d0"]
        self.rarrhalo = st.get("r")
        self.ltablehalo = np.asarray(st.get("ltable", np.arange(self.lmaxhalo + 1)), np.float64)
        self.halo_use_l, self.halo_use_n = self.lmaxhalo, self.nmaxhalo
        self._frozen = None

    def prep_tables(self):
        """potential.py:257-294: read the EOF cache and the SL cache + model
        into the device context."""
        self._set_tables(eof_mod.load_eof(self.eof_file),
                         halo_methods.load_sl(self.sph_file, self.model_file))

    def total_coefficients(self, disk=None, halo=None, return_forces=False):
        """potential.py:98-252.  Particles come from the PSP file `infile`
        (components 'star' and 'dark') or from the `disk` / `halo` objects
        (particle.holder, Input with .data).  transform rotates both into the
        bar frame of the disc (pattern.BarTransform, :121-148).  centering
        subtracts the mass centre of the 10 000 innermost particles of each
        component (the disc centre for both with mutual_center; with compat
        the halo centre uses the indices of the innermost disc particles, as
        the reference does at :199-209).  halofac is 1 (the reference's ratio
        of two identical reads, :138-143).  return_forces=True also returns
        return_forces_cart (fx_disk, fx_halo, ..., pot_halo; potential.py:
        445-497) at every particle used, disc particles then halo particles
        (in the bar frame and centred coordinates the coefficients used), with
        the current field parameters, from one bfe_forces_disc_halo call."""
        if disk is None:
            disk = particle.Input(self.filename, comp="star", verbose=self.verbose)
        if halo is None:
            halo = particle.Input(self.filename, comp="dark", verbose=self.verbose)
        self.time = getattr(disk, "time", None)
        if self.transform:
            from ..analysis import pattern
            disk = pattern.BarTransform(disk)
            halo = pattern.BarTransform(halo, bar_angle=disk.bar_angle)
            self.bar_angle = disk.bar_angle
        if self.system is None:
            self.prep_tables()
        dx, dy, dz, dm = (np.asarray(a, np.float64) for a in particle.particle_arrays(disk))
        hx, hy, hz, hm = (np.asarray(a, np.float64) for a in particle.particle_arrays(halo))
        self.halofac = 1.0
        if self.centering:
            def inner(x, y, z, ncenter=10000):
                return np.argsort(x * x + y * y + z * z)[:ncenter]

            def centre(x, y, z, m, idx):
                w = m[idx].sum()
                return tuple(float((v[idx] * m[idx]).sum() / w) for v in (x, y, z))
            didx = inner(dx, dy, dz)
            self.xcen_disk, self.ycen_disk, self.zcen_disk = centre(dx, dy, dz, dm, didx)
            if self.mutual_center:
                self.xcen_halo, self.ycen_halo, self.zcen_halo = \
                    self.xcen_disk, self.ycen_disk, self.zcen_disk
            else:
                hidx = didx if self.compat else inner(hx, hy, hz)
                self.xcen_halo, self.ycen_halo, self.zcen_halo = centre(hx, hy, hz, hm, hidx)
            dx, dy, dz = dx - self.xcen_disk, dy - self.ycen_disk, dz - self.zcen_disk
            hx, hy, hz = hx - self.xcen_halo, hy - self.ycen_halo, hz - self.zcen_halo
        else:
            self.xcen_disk = self.ycen_disk = self.zcen_disk = 0.0
            self.xcen_halo = self.ycen_halo = self.zcen_halo = 0.0
        dx, dy, dz = eof_mod._nan_to_origin(dx, dy, dz)
        forces = None
        if return_forces:
            ll, hn = self._halo_truncation()
            got = self.system.forces_disc_halo(
                (dx, dy, dz, dm), (hx, hy, hz, hm), halofac=self.halofac, no_odd=self.no_odd,
                disk_m=min(self.disk_use_m, self.mmax), disk_n=min(self.disk_use_n, self.norder),
                halo_l=ll, halo_n=hn, compat=self.compat, components=True)
            c, s, hcoef = got["cos"], got["sin"], got["coef"]
            forces = tuple(v for v in got["comp"].double().cpu().numpy())
        else:
            # disc and halo accumulate run concurrently on the device
            c, s, hcoef = self.system.accumulate_disc_halo((dx, dy, dz, dm), (hx, hy, hz, hm),
                                                           no_odd=self.no_odd)
        self.EOF = eof_mod.EOF_Object()
        self.EOF.cos, self.EOF.sin = c.cpu().numpy(), s.cpu().numpy()
        self.EOF.mmax, self.EOF.nmax, self.EOF.eof_file = self.mmax, self.norder, self.eof_file
        self.EOF.time, self.EOF.nbodies = self.time, dm.size
        self.SL = spheresl.SL_Object()
        self.SL.expcoef = hcoef.cpu().numpy()
        self.SL.lmax, self.SL.nmax = self.lmaxhalo, self.nmaxhalo
        self.SL.sph_file, self.SL.model_file = self.sph_file, self.model_file
        self.SL.time, self.SL.nbodies = self.time, hm.size
        self._frozen = None
        if return_forces:
            return forces

    # ------------------------------------------------------------ parameters
    def set_field_parameters(self, no_odd=False, halo_l=-1, halo_n=-1, disk_m=-1, disk_n=-1):
        """potential.py:365-378."""
        self.no_odd = no_odd
        if halo_l > -1:
            self.halo_use_l = halo_l
        if halo_n > -1:
            self.halo_use_n = halo_n
        if disk_m > -1:
            self.disk_use_m = disk_m
        if disk_n > -1:
            self.disk_use_n = disk_n
        self._frozen = None

    def reset_field_parameters(self):
        """potential.py:380-387."""
        self.no_odd = False
        self.halo_use_l, self.halo_use_n = self.lmaxhalo, self.nmaxhalo
        self.disk_use_m, self.disk_use_n = self.mmax, self.norder
        self._frozen = None

    def _halo_truncation(self):
        """spheresl.force_eval cuts l when halo_l < lmax and n when
        halo_n < nmax - 1 (spheresl.py:1134-1147)."""
        ll = self.halo_use_l if self.halo_use_l < self.lmaxhalo else self.lmaxhalo
        hn = self.halo_use_n if self.halo_use_n < self.nmaxhalo - 1 else self.nmaxhalo
        return ll, hn

    def freeze(self):
        """Fold the current coefficients and parameters into the device field."""
        key = (self.no_odd, self.halo_use_l, self.halo_use_n, self.disk_use_m, self.disk_use_n,
               self.halofac, id(self.EOF.cos), id(self.EOF.sin), id(self.SL.expcoef), self.compat)
        if self._frozen == key:
            return
        ll, hn = self._halo_truncation()
        self.system.set_field(self.EOF.cos, self.EOF.sin, self.SL.expcoef, halofac=self.halofac,
                              no_odd=self.no_odd, disk_m=min(self.disk_use_m, self.mmax),
                              disk_n=min(self.disk_use_n, self.norder), halo_l=ll, halo_n=hn,
                              compat=self.compat)
        self._frozen = key

    # ------------------------------------------------------------ evaluation
    def return_forces_cart(self, xval, yval, zval, rotpos=0.0):
        """potential.py:445-497: (fx_disk, fx_halo, fy_disk, fy_halo, fz_disk,
        fz_halo, pot_disk, pot_halo); floats for scalar input, float64 arrays
        otherwise."""
        scalar = np.ndim(xval) == 0 and np.ndim(yval) == 0 and np.ndim(zval) == 0
        x, y, z = (np.atleast_1d(np.asarray(v, np.float64)) for v in np.broadcast_arrays(xval, yval, zval))
        self.freeze()
        _, _, _, _, comp = self.system.field_forces(x.ravel(), y.ravel(), z.ravel(),
                                                    rotpos=float(rotpos), components=True)
        c = comp.double().cpu().numpy()
        if scalar:
            return tuple(float(v[0]) for v in c)
        return tuple(v.reshape(x.shape) for v in c)

    def get_cartesian_forces_array(self, arr):
        """Total force (fx, fy, fz) at arr = (x, y, z), the interface
        utils/rk4.py expects of a potential (rk4.py:50); arrays of points
        give (3, N)."""
        fxd, fxh, fyd, fyh, fzd, fzh, _, _ = self.return_forces_cart(arr[0], arr[1], arr[2])
        return np.array([fxd + fxh, fyd + fyh, fzd + fzh])

    # ------------------------------------------------------------ field files
    def save_field(self, filename=''):
        """potential.py:738-965: the field in the reference's binary layout:
        4 x S100 file names; i4 nhalo, transform, no_odd, centering,
        mutual_center, verbose; f4 time; i4 numx, numy, mmax, norder, cmap,
        dens; f4 rmin, rmax, ascale, hscale, XMIN, dX, YMIN, dY and the disc
        centre; f8 cos, sin, then potC, rforceC, zforceC, densC, potS,
        rforceS, zforceS, densS; f4 halofac, rmin, rmax, scale and the halo
        centre; i4 numr, cmap, lmax, nmax; f8 xi, p0, d0, ltable, ev, ef,
        expcoef."""
        if filename == '':
            print('potential.Fields.save_field: No filename specified.')
        shape = (self.mmax + 1, self.norder, self.numx + 1, self.numy + 1)

        def dens(a):
            return np.zeros(shape) if a is None else np.asarray(a, np.float64)
        with open(filename, 'wb') as f:
            np.array([self.filename, self.eof_file, self.sph_file, self.model_file],
                     dtype='S100').tofile(f)
            np.array([self.nhalo, self.transform, self.no_odd, self.centering, self.mutual_center,
                      self.verbose], dtype='i4').tofile(f)
            np.array([0.0 if getattr(self, 'time', None) is None else self.time],
                     dtype='f4').tofile(f)
            np.array([self.numx, self.numy, self.mmax, self.norder, self.cmapdisk, self.densdisk],
                     dtype='i4').tofile(f)
            cen = [getattr(self, k, 0.0) for k in ("xcen_disk", "ycen_disk", "zcen_disk",
                                                    "xcen_halo", "ycen_halo", "zcen_halo")]
            np.array([self.rmindisk, self.rmaxdisk, self.ascale, self.hscale, self.XMIN, self.dX,
                      self.YMIN, self.dY] + cen[:3], dtype='f4').tofile(f)
            for a in (self.EOF.cos, self.EOF.sin, self.potC, self.rforceC, self.zforceC,
                      dens(self.densC), self.potS, self.rforceS, self.zforceS, dens(self.densS)):
                np.asarray(a, dtype='f8').reshape(-1).tofile(f)
            np.array([self.halofac, self.rminhalo, self.rmaxhalo, self.scalehalo] + cen[3:],
                     dtype='f4').tofile(f)
            np.array([self.numrhalo, self.cmaphalo, self.lmaxhalo, self.nmaxhalo],
                     dtype='i4').tofile(f)
            for a in (self.xihalo, self.p0halo, self.d0halo, self.ltablehalo, self.evtablehalo,
                      self.eftablehalo, self.SL.expcoef):
                np.asarray(a, dtype='f8').reshape(-1).tofile(f)

    def return_forces_cyl(self, xval, yval, zval, rotpos=0.0):
        """potential.py:389-440: (fr_disk, fr_halo, fp_disk, -fp_halo, fz_disk,
        fz_halo, -pot_disk, pot_halo) with R = sqrt(x^2+y^2)+1e-10 and
        r = sqrt(R^2+z^2)+1e-10.  The azimuthal factor is applied by
        evaluating at the point rotated by rotpos."""
        scalar = np.ndim(xval) == 0 and np.ndim(yval) == 0 and np.ndim(zval) == 0
        x, y, z = (np.atleast_1d(np.asarray(v, np.float64)).ravel() for v in
                   np.broadcast_arrays(xval, yval, zval))
        cr, sr = np.cos(rotpos), np.sin(rotpos)
        xr, yr = x * cr - y * sr, x * sr + y * cr
        r2 = np.sqrt(x * x + y * y) + 1.e-10
        r3 = np.sqrt(r2 * r2 + z * z) + 1.e-10
        ec = np.array(self.EOF.cos, np.float64, copy=True)
        es = np.array(self.EOF.sin, np.float64, copy=True)
        ec[min(self.disk_use_m, self.mmax) + 1:] = 0.0
        es[min(self.disk_use_m, self.mmax) + 1:] = 0.0
        ec[:, self.disk_use_n:] = 0.0
        es[:, self.disk_use_n:] = 0.0
        if self.no_odd:
            ec[1::2] = 0.0
            es[1::2] = 0.0
        d = self.system.eof_eval(xr, yr, z, ec, es)
        ll, hn = self._halo_truncation()
        hc = spheresl._truncated(self.halofac * np.asarray(self.SL.expcoef, np.float64),
                                 self.halo_use_l, self.halo_use_n, self.lmaxhalo, self.nmaxhalo)
        h = self.system.sl_eval(xr, yr, z, hc, no_odd=self.no_odd, density=False,
                                cos_l=self.compat, compat=self.compat)
        g = {k: v.double().cpu().numpy() for k, v in d.items()}
        q = {k: v.double().cpu().numpy() for k, v in h.items()}
        dfp = g["fp"]
        hfp = q["potp"]
        guard = r3 < np.min(self.xihalo)
        dfp = np.where(guard, 0.0, dfp)
        hfp = np.where(guard, 0.0, hfp)
        hfr, hft = q["potr"], q["pott"]
        frhalo = -1.0 * (r2 * hfr + z * hft) / r3
        fzhalo = -1.0 * (z * hfr - r2 * hft) / r3
        out = (g["fr"], frhalo, dfp, -1.0 * hfp, g["fz"], fzhalo, -1.0 * (g["p"] + g["p0"]),
               q["pot1"] + q["pot0"])
        if scalar:
            return tuple(float(v[0]) for v in out)
        return out


def restore_field(filename='', compat=False, device=0):
    """potential.py:968-1041: a Fields from a save_field file, ready to
    evaluate (the state after total_coefficients + prep_tables).  As in the
    reference the file names come back as bytes, the table parameters as
    their float32 values (the disc grid uses the stored XMIN, dX, YMIN, dY)
    and the field parameters are reset to all orders."""
    with open(filename, 'rb') as f:
        infile, eof_file, sph_file, model_file = np.fromfile(f, dtype='S100', count=4)
        nhalo, transform, no_odd, centering, mutual_center, verbose = np.fromfile(f, dtype='i4',
                                                                                  count=6)
        [time] = np.fromfile(f, dtype='f4', count=1)
        F = Fields(infile, eof_file, sph_file, model_file, nhalo=nhalo, transform=transform,
                   no_odd=no_odd, centering=centering, mutual_center=mutual_center,
                   verbose=verbose, compat=compat, device=device)
        F.time = time
        numx, numy, mmax, norder, cmapdisk, densdisk = (int(v) for v in
                                                        np.fromfile(f, dtype='i4', count=6))
        (rmindisk, rmaxdisk, ascale, hscale, XMIN, dX, YMIN, dY,
         F.xcen_disk, F.ycen_disk, F.zcen_disk) = np.fromfile(f, dtype='f4', count=11)
        k = (mmax + 1) * norder
        cos = np.fromfile(f, dtype='f8', count=k).reshape(mmax + 1, norder)
        sin = np.fromfile(f, dtype='f8', count=k).reshape(mmax + 1, norder)
        shape = (mmax + 1, norder, numx + 1, numy + 1)
        nt = int(np.prod(shape))
        names = ("potC", "rforceC", "zforceC", "densC", "potS", "rforceS", "zforceS", "densS")
        et = {nm: np.fromfile(f, dtype='f8', count=nt).reshape(shape) for nm in names}
        et.update(mmax=mmax, norder=norder, numx=numx, numy=numy, cmap=cmapdisk, dens=densdisk,
                  rmin=float(rmindisk), rmax=float(rmaxdisk), ascale=float(ascale),
                  hscale=float(hscale), XMIN=float(XMIN), dX=float(dX), YMIN=float(YMIN),
                  dY=float(dY))
        (halofac, rminhalo, rmaxhalo, scalehalo,
         F.xcen_halo, F.ycen_halo, F.zcen_halo) = np.fromfile(f, dtype='f4', count=7)
        numr, cmaphalo, lmax, nmax = (int(v) for v in np.fromfile(f, dtype='i4', count=4))
        xi = np.fromfile(f, dtype='f8', count=numr)
        p0 = np.fromfile(f, dtype='f8', count=numr)
        d0 = np.fromfile(f, dtype='f8', count=numr)
        ltable = np.fromfile(f, dtype='f8', count=lmax + 1)
        ev = np.fromfile(f, dtype='f8', count=(lmax + 1) * nmax).reshape(lmax + 1, nmax)
        ef = np.fromfile(f, dtype='f8', count=(lmax + 1) * nmax * numr).reshape(lmax + 1, nmax, numr)
        expcoef = np.fromfile(f, dtype='f8', count=(lmax + 1) ** 2 * nmax).reshape((lmax + 1) ** 2,
                                                                                   nmax)
    st = dict(lmax=lmax, nmax=nmax, numr=numr, cmap=cmaphalo, rmin=float(rminhalo),
              rmax=float(rmaxhalo), scale=float(scalehalo), ltable=ltable, ev=ev, ef=ef, xi=xi,
              p0=p0, d0=d0)
    F._set_tables(et, st)
    F.XMIN, F.dX, F.YMIN, F.dY = (float(v) for v in (XMIN, dX, YMIN, dY))
    F.EOF = eof_mod.EOF_Object()
    F.EOF.cos, F.EOF.sin = cos, sin
    F.EOF.mmax, F.EOF.nmax, F.EOF.eof_file = mmax, norder, eof_file
    F.SL = spheresl.SL_Object()
    F.SL.expcoef = expcoef
    F.SL.lmax, F.SL.nmax = lmax, nmax
    F.SL.model_file, F.SL.sph_file = model_file, sph_file
    F.halofac = float(halofac)
    return F


def get_fields(simulation_directory, simulation_name, intime, eof_file, sph_file, model_file,
               bar_file='', nhalo=1000000, transform=True, fileprefix='OUT', compat=False,
               device=0):
    """potential.py:1243-1289: the field of dump <dir><prefix>.<name>.<intime
    %05i> (centred, mutual centre) and its pattern speed from bar_file.  As
    in the reference the pattern speed is the backward difference of the bar
    position at the tabulated time nearest the dump (find_barpattern with
    smth_order=None replaces the spline derivative computed just before).
    Returns (F, patt, rotfreq = patt / 2 pi)."""
    from ..analysis import pattern
    infile = simulation_directory + fileprefix + '.' + simulation_name + '.%05i' % intime
    if transform:
        B = pattern.BarDetermine()
        B.read_bar(bar_file)
        B.frequency_and_derivative(spline_derivative=2)
        P = particle.Input(infile, comp='star')
        patt = pattern.find_barpattern(P.time, B, smth_order=None)
        rotfreq = patt / (2. * np.pi)
    else:
        patt = 0.
        rotfreq = 0.
    F = Fields(infile, eof_file, sph_file, model_file, nhalo=nhalo, transform=transform,
               no_odd=False, centering=True, mutual_center=True, compat=compat, device=device)
    F.total_coefficients()      # reads the tables too (prep_tables)
    return F, patt, rotfreq
