"""Coordinate maps of the BFE tables (host side, float64 NumPy).

Same semantics as exptool/basis/compatibility.py:16-99; the device versions
live in exptool_amd/csrc/bfe_math.hpp and are compiled into the HIP library.
"""
from __future__ import annotations

import numpy as np


def r_to_xi(r, cmap, scale):
    """compatibility.py:16-47: cmap 1 -> (r/s-1)/(r/s+1); cmap 2 -> ln r;
    otherwise r.  Negative r maps to 0.  Scalars stay scalars."""
    a = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        if cmap == 1:
            out = (a / scale - 1.0) / (a / scale + 1.0)
        elif cmap == 2:
            out = np.log(a)
        else:
            out = np.array(a, copy=True)
    out = np.where(a < 0.0, 0.0, out)
    return out[()] if out.ndim == 0 else out


def xi_to_r(xi, cmap, scale):
    """compatibility.py:54-62."""
    if cmap == 1:
        return (1.0 + xi) / (1.0 - xi) * scale
    if cmap == 2:
        return np.exp(xi)
    return xi


def d_xi_to_r(xi, cmap, scale):
    """compatibility.py:65-80 (returns d xi / d r)."""
    if cmap == 1:
        return 0.5 * (1.0 - xi) * (1.0 - xi) / scale
    if cmap == 2:
        return np.exp(-xi)
    return 1.0


def z_to_y(z, hscale):
    """compatibility.py:83-91."""
    return (z / (np.abs(z) + 1.e-8)) * np.arcsinh(np.abs(z / hscale))


def y_to_z(y, hscale):
    """compatibility.py:94-99."""
    return hscale * np.sinh(y)
