"""Particle containers crossing the drop-in boundary, io/particle.py.

  holder  io/particle.py:142-157   attribute bag (xpos .. mass, pote, id)
  Input   io/particle.py:17-110    PSP reader front end; legacy=True exposes
                                   xpos/ypos/... attributes, else .data
"""
from __future__ import annotations

import numpy as np

from . import psp_io


class holder(object):
    """Attribute bag of particle quantities (io/particle.py:142-157)."""
    filename = None
    comp = None
    nbodies = None
    time = None
    xpos = None
    ypos = None
    zpos = None
    xvel = None
    yvel = None
    zvel = None
    mass = None
    pote = None
    id = None


class Input:
    """io/particle.py:17-110 for PSP ("OUT.") files."""

    def __init__(self, filename, comp=None, legacy=False, verbose=0):
        if "SPL." in str(filename):
            raise ValueError(f'SPL files are not supported: "{filename}"')
        I = psp_io.Input(filename, comp=comp, verbose=verbose)
        self.style = "OUT"
        self.header = I.header
        self.filename = I.filename
        self.time = I.time
        self.comp = I.comp
        if comp is None:
            return
        if legacy:
            d = I.data
            self.mass, self.xpos, self.ypos, self.zpos = d["m"], d["x"], d["y"], d["z"]
            self.xvel, self.yvel, self.zvel, self.pote = d["vx"], d["vy"], d["vz"], d["potE"]
            if "id" in d:
                self.id = d["id"]
            self.nbodies = d["x"].size
        else:
            self.data = I.data


def particle_arrays(P):
    """(x, y, z, m) of a holder / legacy Input (xpos .. mass) or of an object
    with a .data dict (x, y, z, m)."""
    d = getattr(P, "data", None)
    if isinstance(d, dict):
        return d["x"], d["y"], d["z"], d.get("m")
    return P.xpos, P.ypos, P.zpos, getattr(P, "mass", None)


def make_holder(x, y, z, m, **kw):
    """holder filled from arrays (velocities and the rest through kw)."""
    h = holder()
    h.xpos, h.ypos, h.zpos, h.mass = (np.asarray(a) for a in (x, y, z, m))
    h.nbodies = h.xpos.size
    for k, v in kw.items():
        setattr(h, k, v)
    return h
