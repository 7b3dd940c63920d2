"""Device-side BFE system: tables resident on one GPU, torch tensors in and out.

``BFESystem`` owns one library context with (optionally) one spherical SL
table set and one cylindrical EOF table set, the pairing the reference's
``Fields`` uses (exptool/basis/potential.py:49-294).  Particle arguments are
float32 device tensors (numpy arrays are copied to the device); coefficients
are float64 device tensors.  Every method enqueues on the current torch
stream and returns device tensors without synchronising.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from ._lib import BFEError, c64, f32, f64, host_ptr, ptr, stream_ptr


def eof_table_params(t):
    """set_table_params (exptool/basis/eof.py:316-347) for a table dict."""
    from .basis.compatibility import r_to_xi, z_to_y
    Rtable = math.sqrt(0.5) * t["rmax"]
    a, h = t["ascale"], t["hscale"]
    XMIN = float(r_to_xi(t["rmin"] * a, t["cmap"], a))
    XMAX = float(r_to_xi(Rtable * a, t["cmap"], a))
    dX = (XMAX - XMIN) / t["numx"]
    YMIN = float(z_to_y(-Rtable * a, h))
    YMAX = float(z_to_y(Rtable * a, h))
    dY = (YMAX - YMIN) / t["numy"]
    return XMIN, XMAX, dX, YMIN, YMAX, dY


def _array_key(a):
    """Identity of a table array for the system cache: its buffer, shape and
    dtype and a digest of its values (all of them up to 16 MiB, else a
    65 536-value stride sample plus the first and last 64 KiB)."""
    import hashlib
    a = np.asarray(a)
    b = np.ascontiguousarray(a)
    raw = b.reshape(-1).view(np.uint8)
    if raw.size <= (16 << 20):
        data = raw.tobytes()
    else:
        flat = b.reshape(-1)
        data = (np.ascontiguousarray(flat[::max(1, flat.size // 65536)]).tobytes() +
                raw[:65536].tobytes() + raw[-65536:].tobytes())
    return (a.__array_interface__["data"][0], a.shape, a.dtype.str,
            hashlib.blake2b(data, digest_size=16).hexdigest())


_TABLE_SYSTEMS = {}


def system_for_tables(kind, tables, scalars):
    """BFESystem holding the SL (kind "sl") or EOF (kind "eof") tables given
    as arrays, kept for later calls with the same arrays and parameters (the
    reference's point evaluators take the tables as arguments on every call;
    uploading them once keeps a per-point loop from re-uploading).  At most
    four table sets are kept."""
    key = (kind, tuple(sorted((k, _array_key(v)) for k, v in tables.items() if v is not None)),
           tuple(sorted(scalars.items())))
    s = _TABLE_SYSTEMS.get(key)
    if s is None:
        t = dict(scalars)
        t.update(tables)
        s = BFESystem(**{kind: t})
        if len(_TABLE_SYSTEMS) >= 4:
            _TABLE_SYSTEMS.pop(next(iter(_TABLE_SYSTEMS)))
        _TABLE_SYSTEMS[key] = s
    return s


class BFESystem:
    """SL + EOF tables on one device."""

    def __init__(self, sl=None, eof=None, device=0):
        self.ctx = _lib.Context(device)
        self.device = torch.device("cuda", self.ctx.device)
        self.sl = None
        self.eof = None
        if sl is not None:
            self.set_sl(sl)
        if eof is not None:
            self.set_eof(eof)

    # ------------------------------------------------------------------ tables
    def set_sl(self, t):
        """t: dict with lmax, nmax, numr, cmap, scale, xi, ev, ef, p0, d0."""
        lmax, nmax, numr = int(t["lmax"]), int(t["nmax"]), int(t["numr"])
        ev = c64(t["ev"]).reshape(lmax + 1, nmax)
        ef = c64(t["ef"]).reshape(lmax + 1, nmax, numr)
        xi, p0, d0 = c64(t["xi"]), c64(t["p0"]), c64(t["d0"])
        rc = self.ctx.lib.bfe_sl_tables_set(self.ctx.h, lmax, nmax, numr, int(t["cmap"]),
                                            float(t["scale"]), host_ptr(xi), host_ptr(ev),
                                            host_ptr(ef), host_ptr(p0), host_ptr(d0))
        self.ctx.check(rc, "bfe_sl_tables_set")
        self.sl = dict(t, lmax=lmax, nmax=nmax, numr=numr)

    def set_eof(self, t):
        """t: dict with mmax, norder, numx, numy, cmap, rmin, rmax, ascale,
        hscale, potC, rforceC, zforceC, potS, rforceS, zforceS[, densC, densS]
        (force tables may be None: zero).
        Grid parameters XMIN, dX, YMIN, dY, when present, are used as given
        instead of being derived from rmin/rmax (set_table_params)."""
        mmax, norder, numx, numy = (int(t[k]) for k in ("mmax", "norder", "numx", "numy"))
        if all(k in t for k in ("XMIN", "dX", "YMIN", "dY")):
            XMIN, dX, YMIN, dY = (float(t[k]) for k in ("XMIN", "dX", "YMIN", "dY"))
        else:
            XMIN, _, dX, YMIN, _, dY = eof_table_params(t)
        shape = (mmax + 1, norder, numx + 1, numy + 1)
        # a table given as None (accumulate needs only potC/potS) is zero
        arrs = {k: (c64(t[k]).reshape(shape) if t.get(k) is not None else np.zeros(shape))
                for k in ("potC", "rforceC", "zforceC", "potS", "rforceS", "zforceS")}
        dens = t.get("densC") is not None and t.get("densS") is not None and int(t.get("dens", 1))
        dC = c64(t["densC"]).reshape(shape) if dens else None
        dS = c64(t["densS"]).reshape(shape) if dens else None
        rc = self.ctx.lib.bfe_eof_tables_set(
            self.ctx.h, mmax, norder, numx, numy, int(t["cmap"]), float(t["ascale"]),
            float(t["hscale"]), XMIN, dX, YMIN, dY, host_ptr(arrs["potC"]),
            host_ptr(arrs["rforceC"]), host_ptr(arrs["zforceC"]), host_ptr(dC),
            host_ptr(arrs["potS"]), host_ptr(arrs["rforceS"]), host_ptr(arrs["zforceS"]),
            host_ptr(dS))
        self.ctx.check(rc, "bfe_eof_tables_set")
        self.eof = dict(t, mmax=mmax, norder=norder, numx=numx, numy=numy, XMIN=XMIN, dX=dX,
                        YMIN=YMIN, dY=dY, has_dens=bool(dens))

    # --------------------------------------------------------------- helpers
    def _xyz(self, x, y, z):
        x, y, z = (f32(a, self.device) for a in (x, y, z))
        n = x.numel()
        if y.numel() != n or z.numel() != n:
            raise ValueError("x, y, z must have the same length")
        return x, y, z, n

    def _need(self, what):
        if getattr(self, what) is None:
            raise BFEError(f"{what} tables are not set on this BFESystem")

    # ------------------------------------------------------------- accumulate
    def sl_accumulate(self, x, y, z, m, no_odd=False, bf16=False, stream=None):
        """Spherical coefficients ((lmax+1)^2, nmax) float64 (device).  bf16:
        the two-piece bf16 MFMA mode (about 1e-5 relative)."""
        self._need("sl")
        x, y, z, n = self._xyz(x, y, z)
        m = f32(m, self.device)
        if m.numel() != n:
            raise ValueError("mass must have the same length as x")
        out = torch.empty(((self.sl["lmax"] + 1) ** 2, self.sl["nmax"]), dtype=torch.float64,
                          device=self.device)
        rc = self.ctx.lib.bfe_sl_accumulate(self.ctx.h, ptr(x), ptr(y), ptr(z), ptr(m), n,
                                            (_lib.NO_ODD if no_odd else 0) |
                                            (_lib.ACC_BF16 if bf16 else 0), ptr(out),
                                            stream_ptr(stream))
        self.ctx.check(rc, "bfe_sl_accumulate")
        return out

    def eof_accumulate(self, x, y, z, m, bf16=False, stream=None):
        """Cylindrical coefficients (cos, sin), each (mmax+1, norder) float64.
        bf16: the bf16 MFMA mode (cell corner sums times the table as bf16
        hi + lo pieces; about 1e-5 relative)."""
        self._need("eof")
        x, y, z, n = self._xyz(x, y, z)
        m = f32(m, self.device)
        if m.numel() != n:
            raise ValueError("mass must have the same length as x")
        shape = (self.eof["mmax"] + 1, self.eof["norder"])
        cos = torch.empty(shape, dtype=torch.float64, device=self.device)
        sin = torch.empty(shape, dtype=torch.float64, device=self.device)
        rc = self.ctx.lib.bfe_eof_accumulate(self.ctx.h, ptr(x), ptr(y), ptr(z), ptr(m), n,
                                             _lib.ACC_BF16 if bf16 else 0, ptr(cos), ptr(sin),
                                             stream_ptr(stream))
        self.ctx.check(rc, "bfe_eof_accumulate")
        return cos, sin

    def accumulate_disc_halo(self, disc, halo, no_odd=False, bf16=False, stream=None):
        """Fields.total_coefficients for one snapshot: disc (x, y, z, m) -> EOF
        (cos, sin) and halo (x, y, z, m) -> SL coefficients in one call; the
        two accumulates run concurrently.  Returns (cos, sin, coef)."""
        self._need("eof")
        self._need("sl")
        dx, dy, dz, nd = self._xyz(*disc[:3])
        dm = f32(disc[3], self.device)
        hx, hy, hz, nh = self._xyz(*halo[:3])
        hm = f32(halo[3], self.device)
        if dm.numel() != nd or hm.numel() != nh:
            raise ValueError("mass must have the same length as x")
        shape = (self.eof["mmax"] + 1, self.eof["norder"])
        cos = torch.empty(shape, dtype=torch.float64, device=self.device)
        sin = torch.empty(shape, dtype=torch.float64, device=self.device)
        coef = torch.empty(((self.sl["lmax"] + 1) ** 2, self.sl["nmax"]), dtype=torch.float64,
                           device=self.device)
        rc = self.ctx.lib.bfe_accumulate_disc_halo(
            self.ctx.h, ptr(dx), ptr(dy), ptr(dz), ptr(dm), nd, ptr(hx), ptr(hy), ptr(hz), ptr(hm),
            nh, (_lib.NO_ODD if no_odd else 0) | (_lib.ACC_BF16 if bf16 else 0), ptr(cos), ptr(sin),
            ptr(coef), stream_ptr(stream))
        self.ctx.check(rc, "bfe_accumulate_disc_halo")
        return cos, sin, coef

    def forces_disc_halo(self, disc, halo, halofac=1.0, no_odd=False, disk_m=-1, disk_n=-1,
                         halo_l=-1, halo_n=-1, compat=False, rotpos=0.0, bf16=False,
                         components=False, stream=None):
        """Fields.total_coefficients then return_forces_cart at the same
        particles, in one call (bfe_forces_disc_halo): disc (x, y, z, m) and
        halo (x, y, z, m) give the coefficients, the frozen field and the
        forces on every particle, disc first.  Returns dict(cos, sin, coef,
        fx, fy, fz, pot[, comp]); the field stays set."""
        self._need("eof")
        self._need("sl")
        dx, dy, dz, nd = self._xyz(*disc[:3])
        dm = f32(disc[3], self.device)
        hx, hy, hz, nh = self._xyz(*halo[:3])
        hm = f32(halo[3], self.device)
        if dm.numel() != nd or hm.numel() != nh:
            raise ValueError("mass must have the same length as x")
        shape = (self.eof["mmax"] + 1, self.eof["norder"])
        out = dict(cos=torch.empty(shape, dtype=torch.float64, device=self.device),
                   sin=torch.empty(shape, dtype=torch.float64, device=self.device),
                   coef=torch.empty(((self.sl["lmax"] + 1) ** 2, self.sl["nmax"]),
                                    dtype=torch.float64, device=self.device))
        n = nd + nh
        for k in ("fx", "fy", "fz", "pot"):
            out[k] = torch.empty(n, dtype=torch.float32, device=self.device)
        comp = torch.empty((8, n), dtype=torch.float32, device=self.device) if components else None
        acc = (_lib.NO_ODD if no_odd else 0) | (_lib.ACC_BF16 if bf16 else 0)
        fld = (_lib.NO_ODD if no_odd else 0) | (_lib.COMPAT_ALL if compat else 0)
        rc = self.ctx.lib.bfe_forces_disc_halo(
            self.ctx.h, ptr(dx), ptr(dy), ptr(dz), ptr(dm), nd, ptr(hx), ptr(hy), ptr(hz), ptr(hm),
            nh, acc, ptr(out["cos"]), ptr(out["sin"]), ptr(out["coef"]), float(halofac),
            int(disk_m), int(disk_n), int(halo_l), int(halo_n), fld, float(rotpos),
            ptr(out["fx"]), ptr(out["fy"]), ptr(out["fz"]), ptr(out["pot"]), ptr(comp),
            stream_ptr(stream))
        self.ctx.check(rc, "bfe_forces_disc_halo")
        self._field_keep = (out["cos"], out["sin"], out["coef"])
        if components:
            out["comp"] = comp
        return out

    # ------------------------------------------------------------- batches
    @staticmethod
    def _offsets(counts_or_offsets, n=None, offsets=False):
        """int64 host offsets [nsnap + 1] from per-snapshot counts (or given offsets)."""
        a = np.asarray(counts_or_offsets, dtype=np.int64)
        off = a.copy() if offsets else np.concatenate([[0], np.cumsum(a)])
        if off[0] != 0 or np.any(np.diff(off) < 0):
            raise ValueError("snapshot offsets must start at 0 and not decrease")
        if n is not None and off[-1] != n:
            raise ValueError(f"snapshot sizes sum to {off[-1]}, not {n} particles")
        return np.ascontiguousarray(off, dtype=np.int64)

    def eof_accumulate_batch(self, x, y, z, m, counts, bf16=False, stream=None):
        """Coefficients of many snapshots in one launch: particles concatenated,
        counts[s] of snapshot s.  Returns (cos, sin), each (nsnap, mmax+1, norder).
        bf16: the bf16 MFMA mode, one (snapshots x nodes) x (nodes x norder)
        product per row over all snapshots."""
        self._need("eof")
        x, y, z, n = self._xyz(x, y, z)
        m = f32(m, self.device)
        off = self._offsets(counts, n)
        ns = off.size - 1
        shape = (ns, self.eof["mmax"] + 1, self.eof["norder"])
        cos = torch.empty(shape, dtype=torch.float64, device=self.device)
        sin = torch.empty(shape, dtype=torch.float64, device=self.device)
        rc = self.ctx.lib.bfe_eof_accumulate_batch(self.ctx.h, ptr(x), ptr(y), ptr(z), ptr(m),
                                                   host_ptr(off), ns, _lib.ACC_BF16 if bf16 else 0,
                                                   ptr(cos), ptr(sin), stream_ptr(stream))
        self.ctx.check(rc, "bfe_eof_accumulate_batch")
        return cos, sin

    def sl_accumulate_batch(self, x, y, z, m, counts, no_odd=False, stream=None):
        """SL coefficients of many snapshots: (nsnap, (lmax+1)^2, nmax)."""
        self._need("sl")
        x, y, z, n = self._xyz(x, y, z)
        m = f32(m, self.device)
        off = self._offsets(counts, n)
        ns = off.size - 1
        out = torch.empty((ns, (self.sl["lmax"] + 1) ** 2, self.sl["nmax"]), dtype=torch.float64,
                          device=self.device)
        rc = self.ctx.lib.bfe_sl_accumulate_batch(self.ctx.h, ptr(x), ptr(y), ptr(z), ptr(m),
                                                  host_ptr(off), ns, _lib.NO_ODD if no_odd else 0,
                                                  ptr(out), stream_ptr(stream))
        self.ctx.check(rc, "bfe_sl_accumulate_batch")
        return out

    def set_field_batch(self, eof_cos, eof_sin, sl_coef, halofac=1.0, no_odd=False, disk_m=-1,
                        disk_n=-1, halo_l=-1, halo_n=-1, compat=False, stream=None):
        """Freeze nsnap fields at once (eof_cos/eof_sin (nsnap, mmax+1, norder),
        sl_coef (nsnap, (lmax+1)^2, nmax))."""
        self._need("sl")
        self._need("eof")
        ec, es, sc = (f64(v, self.device) for v in (eof_cos, eof_sin, sl_coef))
        per_e = (self.eof["mmax"] + 1) * self.eof["norder"]
        per_s = (self.sl["lmax"] + 1) ** 2 * self.sl["nmax"]
        ns = ec.numel() // per_e
        if ec.numel() != ns * per_e or es.numel() != ec.numel() or sc.numel() != ns * per_s:
            raise ValueError("coefficients must be (nsnap, mmax+1, norder) x2 and (nsnap, (lmax+1)^2, nmax)")
        flags = (_lib.NO_ODD if no_odd else 0) | (_lib.COMPAT_ALL if compat else 0)
        rc = self.ctx.lib.bfe_field_set_batch(self.ctx.h, ptr(ec), ptr(es), ptr(sc), ns,
                                              float(halofac), int(disk_m), int(disk_n), int(halo_l),
                                              int(halo_n), flags, stream_ptr(stream))
        self.ctx.check(rc, "bfe_field_set_batch")
        self._batch_keep = (ec, es, sc)
        self._batch_nsnap = ns

    def field_forces_batch(self, x, y, z, counts, rotpos=0.0, stream=None):
        """Forces of the batch fields: the counts[s] points of snapshot s (points
        concatenated) in field s.  Returns (fx, fy, fz, pot) float32 device tensors."""
        x, y, z, n = self._xyz(x, y, z)
        off = self._offsets(counts, n)
        outs = [torch.empty(n, dtype=torch.float32, device=self.device) for _ in range(4)]
        rc = self.ctx.lib.bfe_field_eval_batch(self.ctx.h, ptr(x), ptr(y), ptr(z), host_ptr(off),
                                               off.size - 1, float(rotpos),
                                               *[ptr(o) for o in outs], stream_ptr(stream))
        self.ctx.check(rc, "bfe_field_eval_batch")
        return tuple(outs)

    # ---------------------------------------------------------------- tuning
    def set_tuning(self, name, value):
        self.ctx.check(self.ctx.lib.bfe_set_tuning(self.ctx.h, name.encode(), int(value)),
                       "bfe_set_tuning")

    def get_tuning(self, name):
        import ctypes as C
        v = C.c_int64()
        self.ctx.check(self.ctx.lib.bfe_get_tuning(self.ctx.h, name.encode(), C.byref(v)),
                       "bfe_get_tuning")
        return int(v.value)

    def kernel_times(self, name):
        """(total seconds, launches) of the timed compute kernel `name` since the
        last call (needs set_tuning("kernel_timing", 1))."""
        import ctypes as C
        ms, cnt = C.c_double(), C.c_int64()
        self.ctx.check(self.ctx.lib.bfe_kernel_times(self.ctx.h, name.encode(), C.byref(ms),
                                                     C.byref(cnt)), "bfe_kernel_times")
        return ms.value * 1e-3, int(cnt.value)

    # --------------------------------------------------------------- evaluate
    def sl_eval(self, x, y, z, coef, L1=-1000, L2=1000, no_odd=False, compat=False,
                density=True, cos_l=False, stream=None):
        """all_eval_particles outputs as float32 device tensors:
        dict(den0, den1, pot0, pot1, potr, pott, potp).  compat reproduces the
        density and first-cell quirks; cos_l the cos/sin(l phi) factor of
        spheresl.force_eval."""
        self._need("sl")
        x, y, z, n = self._xyz(x, y, z)
        coef = f64(coef, self.device)
        names = ("den0", "den1", "pot0", "pot1", "potr", "pott", "potp")
        out = {k: torch.empty(n, dtype=torch.float32, device=self.device) for k in names
               if density or not k.startswith("den")}
        flags = (_lib.NO_ODD if no_odd else 0)
        if compat:
            flags |= _lib.COMPAT_SL_DENS | _lib.COMPAT_SL_EDGE
        if cos_l:
            flags |= _lib.COMPAT_SL_COS_L
        rc = self.ctx.lib.bfe_sl_eval(self.ctx.h, ptr(x), ptr(y), ptr(z), n, ptr(coef),
                                      int(max(min(L1, 2**30), -2**30)),
                                      int(max(min(L2, 2**30), -2**30)), flags,
                                      *[ptr(out.get(k)) for k in names], stream_ptr(stream))
        self.ctx.check(rc, "bfe_sl_eval")
        return out

    def eof_eval(self, x, y, z, cos, sin, m1=0, m2=1000, density=False, stream=None):
        """accumulated_eval_particles outputs: dict(p0, p, fr, fp, fz[, d0, d])."""
        self._need("eof")
        x, y, z, n = self._xyz(x, y, z)
        cos, sin = f64(cos, self.device), f64(sin, self.device)
        names = ("p0", "p", "fr", "fp", "fz", "d0", "d")
        out = {k: torch.empty(n, dtype=torch.float32, device=self.device) for k in names
               if density or k not in ("d0", "d")}
        rc = self.ctx.lib.bfe_eof_eval(self.ctx.h, ptr(x), ptr(y), ptr(z), n, ptr(cos), ptr(sin),
                                       int(m1), int(min(m2, 2**30)), 0,
                                       *[ptr(out.get(k)) for k in names], stream_ptr(stream))
        self.ctx.check(rc, "bfe_eof_eval")
        return out

    # --------------------------------------------------------- point forms
    def sl_eval_sph(self, r, costh, phi, coef, lmax_use=-1, nmax_use=-1, no_odd=False,
                    density=False, edge=False, cos_l=False, stream=None):
        """The SL expansion at points given as (r, cos theta, phi) in float64
        (bfe_sl_eval_sph): (7, n) float64 device tensor of den0, den1 (l >= 1),
        pot0, pot1, potr, pott, potp.  edge: get_halo_dens_pot_force's
        first-cell stencil; cos_l: force_eval's cos/sin(l phi)."""
        self._need("sl")
        r, ct, ph = (f64(a, self.device).reshape(-1) for a in (r, costh, phi))
        n = r.numel()
        if ct.numel() != n or ph.numel() != n:
            raise ValueError("r, costh, phi must have the same length")
        coef = f64(coef, self.device)
        out = torch.empty((7, n), dtype=torch.float64, device=self.device)
        flags = ((_lib.NO_ODD if no_odd else 0) | (_lib.EVAL_DENS if density else 0) |
                 (_lib.COMPAT_SL_EDGE if edge else 0) | (_lib.COMPAT_SL_COS_L if cos_l else 0))
        rc = self.ctx.lib.bfe_sl_eval_sph(self.ctx.h, ptr(r), ptr(ct), ptr(ph), n, ptr(coef),
                                          int(lmax_use), int(nmax_use), flags, ptr(out),
                                          stream_ptr(stream))
        self.ctx.check(rc, "bfe_sl_eval_sph")
        return out

    def eof_eval_cyl(self, R, z, phi, cos, sin, mmax_use=-1, nmax_use=-1, no_odd=False,
                     compat_mask=False, density=False, stream=None):
        """The EOF expansion at points given as (R, z, phi) in float64
        (bfe_eof_eval_cyl): (9, n) float64 device tensor of p0, p (m >= 1),
        fr, fp, fz, d0, d (m >= 1), fr0, fz0.  compat_mask: with no_odd,
        force_eval's |cos(pi m/2)| weights for odd m."""
        self._need("eof")
        R, z, ph = (f64(a, self.device).reshape(-1) for a in (R, z, phi))
        n = R.numel()
        if z.numel() != n or ph.numel() != n:
            raise ValueError("R, z, phi must have the same length")
        cos, sin = f64(cos, self.device), f64(sin, self.device)
        out = torch.empty((9, n), dtype=torch.float64, device=self.device)
        flags = ((_lib.NO_ODD if no_odd else 0) | (_lib.EVAL_DENS if density else 0) |
                 (_lib.COMPAT_EOF_MASK if compat_mask else 0))
        rc = self.ctx.lib.bfe_eof_eval_cyl(self.ctx.h, ptr(R), ptr(z), ptr(ph), n, ptr(cos),
                                           ptr(sin), int(mmax_use), int(nmax_use), flags,
                                           ptr(out), stream_ptr(stream))
        self.ctx.check(rc, "bfe_eof_eval_cyl")
        return out

    # ------------------------------------------------------------------ field
    def set_field(self, eof_cos, eof_sin, sl_coef, halofac=1.0, no_odd=False, disk_m=-1,
                  disk_n=-1, halo_l=-1, halo_n=-1, compat=False, stream=None):
        """Freeze a disc+halo field (Fields.prep_tables + set_field_parameters)."""
        self._need("sl")
        self._need("eof")
        ec, es, sc = f64(eof_cos, self.device), f64(eof_sin, self.device), f64(sl_coef, self.device)
        flags = (_lib.NO_ODD if no_odd else 0) | (_lib.COMPAT_ALL if compat else 0)
        rc = self.ctx.lib.bfe_field_set(self.ctx.h, ptr(ec), ptr(es), ptr(sc), float(halofac),
                                        int(disk_m), int(disk_n), int(halo_l), int(halo_n),
                                        flags, stream_ptr(stream))
        self.ctx.check(rc, "bfe_field_set")
        self._field_keep = (ec, es, sc)

    def field_forces(self, x, y, z, rotpos=0.0, components=False, stream=None):
        """Return (fx, fy, fz, pot) float32 device tensors, and with
        components=True also the [8, n] return_forces_cart outputs."""
        x, y, z, n = self._xyz(x, y, z)
        fx, fy, fz, pot = (torch.empty(n, dtype=torch.float32, device=self.device) for _ in range(4))
        comp = torch.empty((8, n), dtype=torch.float32, device=self.device) if components else None
        rc = self.ctx.lib.bfe_field_eval_cart(self.ctx.h, ptr(x), ptr(y), ptr(z), n,
                                              float(rotpos), ptr(fx), ptr(fy), ptr(fz), ptr(pot),
                                              ptr(comp), stream_ptr(stream))
        self.ctx.check(rc, "bfe_field_eval_cart")
        if components:
            return fx, fy, fz, pot, comp
        return fx, fy, fz, pot

    def rk4(self, state, dt, nsteps, omegab=0.0, compat_frame=False, record_every=0, stream=None):
        """Advance state = (x, y, z, vx, vy, vz) float32 device tensors in place
        by nsteps RK4 steps in the frozen field (bfe_rk4); omegab != 0 adds the
        pseudo-forces of a frame rotating about z (the reference's signs with
        compat_frame).  Returns traj (nrec, 6, n) float32 or None."""
        x = state[0]
        n = x.numel()
        for t in state:
            if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n or t.device != self.device:
                raise ValueError("state tensors must be contiguous float32 on the context device")
        traj = None
        if record_every > 0:
            traj = torch.empty((nsteps // record_every + 1, 6, n), dtype=torch.float32,
                               device=self.device)
        flags = _lib.COMPAT_RK4_FRAME if compat_frame else 0
        rc = self.ctx.lib.bfe_rk4(self.ctx.h, *[ptr(t) for t in state], n, float(dt), int(nsteps),
                                  float(omegab), flags, int(record_every), ptr(traj),
                                  stream_ptr(stream))
        self.ctx.check(rc, "bfe_rk4")
        return traj

    def leapfrog(self, state, dt, nsteps, t0=0.0, rotfreq=0.0, energy_every=0,
                 record_every=0, use_graph=False, stream=None):
        """Advance state = (x, y, z, vx, vy, vz) float32 device tensors in place.
        dt is one step size for every orbit, or a float64 tensor (n,) of
        per-orbit step sizes (bfe_leapfrog_dt).
        Returns dict(drift=(n,) float64 max relative energy drift per orbit,
        energy=(checks,) float64 total energy log or None,
        traj=(nrec, 7, n) float32 or None)."""
        x, y, z, vx, vy, vz = state
        n = x.numel()
        for t in state:
            if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n or t.device != self.device:
                raise ValueError("state tensors must be contiguous float32 on the context device")
        drift = torch.empty(n, dtype=torch.float64, device=self.device)
        elog = None
        if energy_every > 0:
            elog = torch.empty(nsteps // energy_every + 1, dtype=torch.float64, device=self.device)
        traj = None
        if record_every > 0:
            traj = torch.empty((nsteps // record_every + 1, 7, n), dtype=torch.float32,
                               device=self.device)
        s = stream if stream is not None else torch.cuda.current_stream()
        if use_graph and s.cuda_stream == 0:
            raise ValueError("use_graph needs a non-default stream (torch.cuda.Stream())")
        if torch.is_tensor(dt):
            if dt.dtype != torch.float64 or dt.numel() != n or dt.device != self.device:
                raise ValueError("per-orbit dt must be a float64 tensor of n values on the context device")
            dt = dt.contiguous()
            rc = self.ctx.lib.bfe_leapfrog_dt(self.ctx.h, *[ptr(t) for t in state], n, ptr(dt),
                                              int(nsteps), float(t0), float(rotfreq), int(energy_every),
                                              ptr(drift), ptr(elog), int(record_every), ptr(traj),
                                              1 if use_graph else 0, stream_ptr(s))
            self.ctx.check(rc, "bfe_leapfrog_dt")
        else:
            rc = self.ctx.lib.bfe_leapfrog(self.ctx.h, *[ptr(t) for t in state], n, float(dt),
                                           int(nsteps), float(t0), float(rotfreq), int(energy_every),
                                           ptr(drift), ptr(elog), int(record_every), ptr(traj),
                                           1 if use_graph else 0, stream_ptr(s))
            self.ctx.check(rc, "bfe_leapfrog")
        return dict(drift=drift, energy=elog, traj=traj)
