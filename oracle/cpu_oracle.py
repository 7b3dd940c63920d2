"""ctypes binding of oracle/libbfe_cpu.so, the FP64 C++/OpenMP restatement
of the hot path (oracle/bfe_cpu.cpp).

TEST INFRASTRUCTURE ONLY: the multi-core CPU baseline of bench.py and a
second restatement that tests/ check against the NumPy oracle.  Functions
take the NumPy oracle's table objects (bfe_oracle.SLTable / EOFTable /
FieldSpec) and follow its semantics without compat flags.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbfe_cpu.so")
_lib = None

_D = C.POINTER(C.c_double)
_I, _I64, _F = C.c_int, C.c_int64, C.c_double


def load():
    """The library (built by `make -C oracle cpu`, which build() runs)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle cpu)")
        lib = C.CDLL(LIB_PATH)
        lib.cpu_threads.restype = _I
        lib.cpu_sl_accumulate.argtypes = [_I, _I, _I, _I, _F] + [_D] * 4 + [_D] * 4 + [_I64, _I, _D]
        lib.cpu_eof_accumulate.argtypes = [_I] * 5 + [_F] * 6 + [_D] * 2 + [_D] * 4 + [_I64, _D, _D]
        lib.cpu_field_eval.argtypes = ([_I, _I, _I, _I, _F] + [_D] * 4 + [_I] * 5 + [_F] * 6 +
                                       [_D] * 6 + [_D] * 3 + [_D] * 3 + [_I64] + [_D] * 4)
        _lib = lib
    return _lib


def threads():
    return int(load().cpu_threads())


def _p(a):
    return a.ctypes.data_as(_D)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _sl_args(t):
    keep = [_c(t.xi), _c(t.evtable), _c(t.eftable), _c(t.p0)]
    return keep, [t.lmax, t.nmax, t.numr, t.cmap, t.scale] + [_p(a) for a in keep]


def _eof_grid(t):
    return [t.mmax, t.norder, t.numx, t.numy, t.cmap, t.ascale, t.hscale, t.XMIN, t.dX, t.YMIN,
            t.dY]


def sl_accumulate(t, x, y, z, m, no_odd=False):
    """compute_coefficients_solitary (spheresl.py:567-656): ((lmax+1)^2, nmax)."""
    keep, args = _sl_args(t)
    pts = [_c(v) for v in (x, y, z, m)]
    out = np.zeros(((t.lmax + 1) ** 2, t.nmax))
    load().cpu_sl_accumulate(*args, *[_p(a) for a in pts], pts[0].size, int(bool(no_odd)), _p(out))
    return out


def eof_accumulate(t, x, y, z, m):
    """accumulate (eof.py:492-640): (accum_cos, accum_sin), each (mmax+1, norder)."""
    tabs = [_c(t.potC), _c(t.potS)]
    pts = [_c(v) for v in (x, y, z, m)]
    c = np.zeros((t.mmax + 1, t.norder))
    s = np.zeros((t.mmax + 1, t.norder))
    load().cpu_eof_accumulate(*_eof_grid(t), *[_p(a) for a in tabs], *[_p(a) for a in pts],
                              pts[0].size, _p(c), _p(s))
    return c, s


def total_force(F, x, y, z):
    """Fields.return_forces_cart summed over disc and halo (potential.py:445-497)
    for a FieldSpec with full truncations, halofac 1, no no_odd, rotpos 0:
    (fx, fy, fz, pot)."""
    if (F.halofac != 1.0 or F.no_odd or F.disk_m != F.eof.mmax or F.disk_n != F.eof.norder
            or F.halo_l != F.sl.lmax or F.halo_n != F.sl.nmax):
        raise ValueError("cpu_oracle.total_force: full truncations, halofac 1, no no_odd only")
    keep, sargs = _sl_args(F.sl)
    e = F.eof
    tabs = [_c(a) for a in (e.potC, e.rforceC, e.zforceC, e.potS, e.rforceS, e.zforceS)]
    coefs = [_c(F.eof_cos), _c(F.eof_sin), _c(F.sl_coef)]
    pts = [_c(v) for v in (x, y, z)]
    n = pts[0].size
    out = [np.zeros(n) for _ in range(4)]
    load().cpu_field_eval(*sargs, *_eof_grid(e), *[_p(a) for a in tabs], *[_p(a) for a in coefs],
                          *[_p(a) for a in pts], n, *[_p(a) for a in out])
    return tuple(out)
