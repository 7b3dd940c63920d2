"""Reader of EXP "outcoef" coefficient files: drop-in for exptool.io.outcoef.

  OutCoef                       io/outcoef.py:40-533
  CylCoefs                      io/outcoef.py:537-589

Formats (little-endian):
  new Cylinder   per record: u32 magic 202004387, u32 n, n bytes of YAML
                 (time, mmax, nmax), then for m = 0..mmax nmax f8 cos terms
                 and, for m > 0, nmax f8 sin terms
  new SphereSL   per record: u32 magic 202004386, u32 n, YAML (time, lmax,
                 nmax), then for n = 0..nmax-1 the (lmax+1)^2 f8 terms
  old Cylinder   per record: f8 time, u32 mmax, nmax, then as new Cylinder
  old SphereSL   per record: 64-byte title containing "Sphere SL", f8 time,
                 f8 scale, u32 nmax, lmax, then as new SphereSL
coefs is (T, 2, mmax+1, nmax) for cylinders (index 1 = sin, zero at m = 0)
and (T, (lmax+1)^2, nmax) for spheres.  Headers are read with
yaml.safe_load.
"""
from __future__ import annotations

import os

import numpy as np
import yaml

CYL_MAGIC = 202004387
SPH_MAGIC = 202004386


class EOF_Object(object):
    """Coefficient set of one time (io/outcoef.py:26-37)."""
    time = None
    dump = None
    comp = None
    nbodies = None
    mmax = None
    nmax = None
    eof_file = None
    cos = None
    sin = None


def _cyl_block(buf, off, mmax, nmax):
    """cos (mmax+1, nmax) and sin of one record starting at byte off."""
    cos = np.zeros((mmax + 1, nmax))
    sin = np.zeros((mmax + 1, nmax))
    rows = np.frombuffer(buf, dtype='<f8', count=(2 * mmax + 1) * nmax, offset=off)
    rows = rows.reshape(2 * mmax + 1, nmax)
    cos[0] = rows[0]
    cos[1:] = rows[1::2]
    sin[1:] = rows[2::2]
    return cos, sin, off + 8 * (2 * mmax + 1) * nmax


class OutCoef(object):
    """EXP outcoef file (io/outcoef.py:40-533): .basis ('Cylinder' or
    'SphereSL'), .T (times) and .coefs."""

    def __init__(self, filename):
        self.coeffile = filename
        with open(filename, 'rb') as f:
            buf = f.read()
        self._buf = buf
        cmagic = int(np.frombuffer(buf, dtype='<u4', count=1)[0]) if len(buf) >= 4 else 0
        if cmagic == SPH_MAGIC:
            self.basis = 'SphereSL'
            print('OutCoef: reading SphereSL coefficients . . .')
            self.read_binary_sl_coefficients()
        elif cmagic == CYL_MAGIC:
            self.basis = 'Cylinder'
            print('OutCoef: reading Cylinder coefficients . . .')
            self.read_binary_eof_coefficients()
        elif b'Sphere SL' in buf[:64]:
            self.basis = 'SphereSL'
            print('OutCoef: reading OLD SphereSL coefficients . . .')
            print('CAUION. These coefficients have a different normalisation scheme.')
            self.read_binary_sl_coefficients_old()
        else:
            self.basis = 'Cylinder'
            print('OutCoef: reading OLD Cylinder coefficients . . .')
            self.read_binary_eof_coefficients_old()
        del self._buf

    def read_binary_eof_coefficients(self):
        """io/outcoef.py:366-418 (new YAML records)."""
        buf, off = self._buf, 0
        T, C = [], []
        while off < len(buf):
            n = int(np.frombuffer(buf, dtype='<u4', count=1, offset=off + 4)[0])
            D = yaml.safe_load(buf[off + 8:off + 8 + n].split(b'\x00', 1)[0].decode())
            cos, sin, off = _cyl_block(buf, off + 8 + n, int(D['mmax']), int(D['nmax']))
            T.append(float(D['time']))
            C.append(np.stack([cos, sin]))
        self.T = np.array(T)
        self.coefs = np.array(C)

    def read_binary_sl_coefficients(self):
        """io/outcoef.py:292-364 (new YAML records)."""
        buf, off = self._buf, 0
        T, C = [], []
        while off < len(buf):
            n = int(np.frombuffer(buf, dtype='<u4', count=1, offset=off + 4)[0])
            D = yaml.safe_load(buf[off + 8:off + 8 + n].split(b'\x00', 1)[0].decode())
            L, N = int(D['lmax']), int(D['nmax'])
            k = (L + 1) ** 2
            a = np.frombuffer(buf, dtype='<f8', count=k * N, offset=off + 8 + n).reshape(N, k)
            T.append(float(D['time']))
            C.append(a.T.copy())
            off += 8 + n + 8 * k * N
        self.T = np.array(T)
        self.coefs = np.array(C)

    def read_binary_eof_coefficients_old(self):
        """io/outcoef.py:132-189: records of f8 time, u32 mmax, nmax and the
        coefficient rows; the record count is the file size over the record
        size of the first record."""
        buf = self._buf
        mmax, nmax = (int(v) for v in np.frombuffer(buf, dtype='<u4', count=2, offset=8))
        rec = 8 * (2 * mmax + 1) * nmax + 4 * 2 + 8
        T, C = [], []
        for t in range(len(buf) // rec):
            off = t * rec
            T.append(float(np.frombuffer(buf, dtype='<f8', count=1, offset=off)[0]))
            cos, sin, _ = _cyl_block(buf, off + 16, mmax, nmax)
            C.append(np.stack([cos, sin]))
        self.T = np.array(T)
        self.coefs = np.array(C)

    def read_binary_sl_coefficients_old(self):
        """io/outcoef.py:192-245: 64-byte title, f8 time, scale, u32 nmax,
        lmax, then nmax rows of (lmax+1)^2 terms per record."""
        buf = self._buf
        nmax, lmax = (int(v) for v in np.frombuffer(buf, dtype='<u4', count=2, offset=80))
        k = lmax * (lmax + 2) + 1
        rec = 8 * k * nmax + 4 * 2 + 8 * 2 + 64
        T, C = [], []
        for t in range(len(buf) // rec):
            off = t * rec
            T.append(float(np.frombuffer(buf, dtype='<f8', count=1, offset=off + 64)[0]))
            a = np.frombuffer(buf, dtype='<f8', count=k * nmax, offset=off + 88).reshape(nmax, k)
            C.append(a.T.copy())
        self.T = np.array(T)
        self.coefs = np.array(C)

    def _repackage_cylindrical_coefficients(self):
        """io/outcoef.py:421-441: self.C[m]['cos' | 'sin'][n] = series in time."""
        numt, _, numm, numn = self.coefs.shape
        self.C = {m: {c: {n: self.coefs[:, ic, m, n] for n in range(numn)}
                      for ic, c in enumerate(('cos', 'sin'))} for m in range(numm)}

    def _repackage_cylindrical_coefficients_compatibility(self):
        """io/outcoef.py:443-483: {time: EOF_Object} with placeholder names."""
        numt, _, numm, numn = self.coefs.shape
        out = dict()
        for tt in range(numt):
            o = EOF_Object()
            o.time = self.T[tt]
            o.mmax, o.nmax = numm - 1, numn
            o.filename, o.comp, o.nbodies, o.eof_file = '[redacted]', 'star', 0., '[redacted]'
            o.cos = self.coefs[tt, 0].copy()
            o.sin = self.coefs[tt, 1].copy()
            out[o.time] = o
        return out

    def _repackage_spherical_coefficients(self):
        """io/outcoef.py:485-533: self.C[l][m]['cos' | 'sin'][n], rows
        l^2 (m = 0), l^2 + 2m - 1 (cos) and l^2 + 2m (sin)."""
        numt, numl2, numn = self.coefs.shape
        numl = int(np.sqrt(numl2))
        self.C = dict()
        for l in range(numl):
            self.C[l] = dict()
            for m in range(l + 1):
                if m > 0:
                    self.C[l][m] = {
                        'cos': {n: self.coefs[:, l * l + 2 * m - 1, n] for n in range(numn)},
                        'sin': {n: self.coefs[:, l * l + 2 * m, n] for n in range(numn)}}
                else:
                    self.C[l][m] = {'cos': {n: self.coefs[:, l * l, n] for n in range(numn)}}


class CylCoefs(object):
    """io/outcoef.py:537-589 as intended: the reference sizes its arrays
    with the time vector and sums an undefined name, so normalise and the
    order sums fail there."""

    def __init__(self, filename):
        self.CylCoefs = OutCoef(filename)
        self.numt = self.CylCoefs.T.size
        self.mmax = self.CylCoefs.coefs.shape[2] - 1
        self.norder = self.CylCoefs.coefs.shape[3]

    def normalise(self):
        """Every coefficient over the (m, n) = (0, 0) cosine term of its time."""
        c = self.CylCoefs.coefs
        self.CylNormCoefs = c / c[:, 0:1, 0:1, 0:1]

    def total_order_power(self):
        c = self.CylCoefs.coefs
        return np.sum(c[:, 0] ** 2 + c[:, 1] ** 2, axis=2)

    def total_order_amplitude(self):
        return np.sqrt(self.total_order_power())
