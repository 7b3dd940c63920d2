"""Frozen disc (EOF) + halo (SL) field: drop-in for exptool.basis.potential.Fields.

  Fields.__init__                 potential.py:56-96
  Fields.total_coefficients       potential.py:98-252 (PSP file or particle objects)
  Fields.prep_tables              potential.py:257-294
  Fields.set_field_parameters     potential.py:365-378
  Fields.reset_field_parameters   potential.py:380-387
  Fields.return_forces_cart       potential.py:445-497 (libexptool_bfe field kernel)
  Fields.return_forces_cyl        potential.py:389-440 (batch evaluators + host assembly)

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
        self.halo_use_l, self.halo_use_n = self.lmaxhalo, self.nmaxhalo
        self._frozen = None

    def prep_tables(self):
        """potential.py:257-294: read the EOF cache and the SL cache + model
        into the device context."""
        self._set_tables(eof_mod.load_eof(self.eof_file),
                         halo_methods.load_sl(self.sph_file, self.model_file))

    def total_coefficients(self, disk=None, halo=None):
        """potential.py:98-252.  Particles come from the PSP file `infile`
        (components 'star' and 'dark') or from the `disk` / `halo` objects
        (particle.holder, Input with .data).  centering subtracts the mass
        centre of the 10 000 innermost particles of each component (the disc
        centre for both with mutual_center).  halofac is 1 (the reference's
        ratio of two identical reads, :138-143)."""
        if self.transform:
            raise NotImplementedError("transform=True (BarTransform) is outside this build")
        if disk is None:
            disk = particle.Input(self.filename, comp="star", verbose=self.verbose)
        if halo is None:
            halo = particle.Input(self.filename, comp="dark", verbose=self.verbose)
        self.time = getattr(disk, "time", None)
        if self.system is None:
            self.prep_tables()
        dx, dy, dz, dm = (np.asarray(a, np.float64) for a in particle.particle_arrays(disk))
        hx, hy, hz, hm = (np.asarray(a, np.float64) for a in particle.particle_arrays(halo))
        self.halofac = 1.0
        if self.centering:
            def centre(x, y, z, m, ncenter=10000):
                idx = np.argsort(x * x + y * y + z * z)[:ncenter]
                w = m[idx].sum()
                return tuple(float((v[idx] * m[idx]).sum() / w) for v in (x, y, z))
            self.xcen_disk, self.ycen_disk, self.zcen_disk = centre(dx, dy, dz, dm)
            if self.mutual_center:
                self.xcen_halo, self.ycen_halo, self.zcen_halo = \
                    self.xcen_disk, self.ycen_disk, self.zcen_disk
            else:
                self.xcen_halo, self.ycen_halo, self.zcen_halo = centre(hx, hy, hz, hm)
            dx, dy, dz = dx - self.xcen_disk, dy - self.ycen_disk, dz - self.zcen_disk
            hx, hy, hz = hx - self.xcen_halo, hy - self.ycen_halo, hz - self.zcen_halo
        else:
            self.xcen_disk = self.ycen_disk = self.zcen_disk = 0.0
            self.xcen_halo = self.ycen_halo = self.zcen_halo = 0.0
        dx, dy, dz = eof_mod._nan_to_origin(dx, dy, dz)
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
