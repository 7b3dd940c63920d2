"""Bar position, pattern speed and the bar frame: drop-in for the parts of
exptool.analysis.pattern the frozen-field path uses.

  BarTransform                 analysis/pattern.py:64-177
  BarDetermine                 analysis/pattern.py:180-471 (read / print /
                               track bar files, unwrap, pattern speed)
  find_barangle                analysis/pattern.py:507-546
  find_barpattern              analysis/pattern.py:549-581

The m = 2 phase of a particle set is one pass over the particles on the
device (float64 sums of cos 2phi and sin 2phi); the bar-file series are a few
thousand values and stay on the host.
"""
from __future__ import annotations

import numpy as np


def m2_phase(x, y, minr=0.0, maxr=1.0, device=None):
    """Half the phase of sum exp(2 i phi) over the particles with
    minr < R < maxr (pattern.py:159-177): the position angle of the m = 2
    distortion.  Computed on the device in float64."""
    import torch
    dev = device if device is not None else torch.device("cuda", 0)
    X = torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)
    Y = torch.as_tensor(np.asarray(y, dtype=np.float64), device=dev)
    R = torch.sqrt(X * X + Y * Y)
    w = (R > minr) & (R < maxr)
    ph = 2.0 * torch.atan2(Y[w], X[w])
    a = torch.cos(ph).sum().item()
    b = torch.sin(ph).sum().item()
    return float(np.arctan2(b, a) / 2.0)


def _data_of(P):
    """The .data dict of a particle object, or one built from a holder's
    xpos .. mass (absent velocities as zeros)."""
    d = getattr(P, 'data', None)
    if isinstance(d, dict):
        return d
    x = np.asarray(P.xpos)

    def get(name):
        v = getattr(P, name, None)
        return np.zeros_like(x) if v is None else np.asarray(v)
    return {'x': x, 'y': np.asarray(P.ypos), 'z': np.asarray(P.zpos), 'vx': get('xvel'),
            'vy': get('yvel'), 'vz': get('zvel'), 'm': np.asarray(P.mass),
            'potE': getattr(P, 'pote', None)}


class BarTransform:
    """pattern.py:64-157: the particles of ParticleInstanceIn (an object with
    .data x, y, z, vx, vy, vz, m, potE) rotated into the bar frame:
    bar_angle = -(m = 2 phase within maxr) + rel_bar_angle unless given, then
    (x, y) and (vx, vy) are rotated counter-clockwise by bar_angle.  The
    input is left unchanged; the result is in .data with .time, .filename,
    .comp copied."""

    def __init__(self, ParticleInstanceIn, bar_angle=None, rel_bar_angle=0., minr=0., maxr=1.):
        self.ParticleInstanceIn = ParticleInstanceIn
        self.bar_angle = bar_angle
        self.data = dict()
        if self.bar_angle is None:
            # the reference passes maxr only (minr stays 0, pattern.py:101)
            d = _data_of(ParticleInstanceIn)
            self.bar_angle = -1. * self.bar_fourier_compute(d['x'], d['y'], maxr=maxr)
        self.bar_angle += rel_bar_angle
        self.calculate_transform_and_return()

    def calculate_transform_and_return(self):
        """pattern.py:115-156."""
        d = _data_of(self.ParticleInstanceIn)
        c, s = np.cos(self.bar_angle), np.sin(self.bar_angle)
        self.data['x'] = d['x'] * c - d['y'] * s
        self.data['y'] = d['x'] * s + d['y'] * c
        self.data['z'] = np.copy(d['z'])
        self.data['vx'] = d['vx'] * c - d['vy'] * s
        self.data['vy'] = d['vx'] * s + d['vy'] * c
        self.data['vz'] = np.copy(d['vz'])
        self.data['m'] = d['m']
        self.data['potE'] = d.get('potE')
        self.time = getattr(self.ParticleInstanceIn, 'time', None)
        self.filename = getattr(self.ParticleInstanceIn, 'filename', None)
        self.comp = getattr(self.ParticleInstanceIn, 'comp', None)

    def bar_fourier_compute(self, posx, posy, minr=0., maxr=1.):
        """pattern.py:159-177."""
        return m2_phase(posx, posy, minr=minr, maxr=maxr)


class BarDetermine:
    """Bar position versus time (pattern.py:180-471): .time, .pos (phase of
    the m = 2 distortion) and .deriv (pattern speed)."""

    def __init__(self, **kwargs):
        if 'file' in kwargs:
            self.read_bar(kwargs['file'])

    def track_bar(self, filelist, verbose=0, maxr=1., apse=False, comp='star'):
        """pattern.py:202-248: the m = 2 phase within maxr of each PSP file
        listed (one file name a line) in filelist, then unwrapped and
        differenced."""
        if apse:
            raise NotImplementedError("apse bar tracking reads apsides files, which exptool_amd does not implement")
        from ..io import particle
        self.slist = filelist
        with open(filelist) as f:
            files = [ln.split()[0] for ln in f if ln.split()]      # parse_list, :218-228
        time, pos = [], []
        for fn in files:
            P = particle.Input(fn, comp=comp, verbose=verbose)
            time.append(P.time)
            pos.append(self.bar_fourier_compute(P.data['x'], P.data['y'], maxr=maxr))
        self.time = np.array(time, dtype=np.float64)
        self.pos = np.array(pos, dtype=np.float64)
        self.unwrap_bar_position()
        self.frequency_and_derivative()

    def unwrap_bar_position(self, jbuffer=-1., smooth=False, reverse=False, adjust=np.pi):
        """pattern.py:329-360: add `adjust` after every step that drops by
        more than |jbuffer| (rises, with reverse)."""
        if smooth:
            raise NotImplementedError("Savitzky-Golay smoothing of the bar position is not provided")
        step = np.diff(self.pos)
        jump = (step > -1. * jbuffer) if reverse else (step < jbuffer)
        jset = np.concatenate([[0], np.cumsum(np.where(jump, -1 if reverse else 1, 0))])
        self.pos = self.pos + jset * adjust

    def frequency_and_derivative(self, smth_order=None, fft_order=None, spline_derivative=None,
                                 verbose=0):
        """pattern.py:362-402: backward differences (deriv[0] = 0), then
        optionally a polynomial fit of order smth_order in time, or the
        derivative of a cubic smoothing spline with smoothing factor
        spline_derivative (also its second derivative in .dderiv)."""
        if smth_order and fft_order and verbose:
            print('Cannot assure proper functionality of both order smoothing and low pass filtering.')
        self.deriv = np.zeros_like(self.pos)
        self.deriv[1:] = np.diff(self.pos) / np.diff(self.time)
        if smth_order:
            self.deriv = np.poly1d(np.polyfit(self.time, self.deriv, smth_order))(self.time)
        if spline_derivative:
            from scipy.interpolate import UnivariateSpline
            spl = UnivariateSpline(self.time, self.pos, k=3, s=spline_derivative)
            self.deriv = spl.derivative()(self.time)
            self.dderiv = spl.derivative(2)(self.time)

    def bar_fourier_compute(self, posx, posy, maxr=1.):
        """pattern.py:404-415."""
        return m2_phase(posx, posy, maxr=maxr)

    def print_bar(self, outfile):
        """pattern.py:418-432: one 'time pos deriv' line per entry."""
        with open(outfile, 'w') as f:
            for i in range(len(self.time)):
                print(self.time[i], self.pos[i], self.deriv[i], file=f)

    def read_bar(self, infile):
        """pattern.py:439-471: columns time, pos[, deriv]; without a deriv
        column the backward differences are computed."""
        time, pos, deriv = [], [], []
        with open(infile) as f:
            for line in f:
                q = [float(d) for d in line.split()]
                time.append(q[0])
                pos.append(q[1])
                if len(q) > 2:
                    deriv.append(q[2])
        self.time = np.array(time)
        self.pos = np.array(pos)
        self.deriv = np.array(deriv)
        if len(self.deriv) < 1:
            self.frequency_and_derivative()


def find_barangle(time, BarInstance, interpolate=True):
    """pattern.py:507-546: minus the bar position at `time` (scalar or
    array), from the interpolating cubic spline through the non-NaN entries
    or, with interpolate=False, at the nearest tabulated time."""
    if interpolate:
        from scipy.interpolate import UnivariateSpline
        ok = ~np.isnan(BarInstance.pos)
        f = UnivariateSpline(BarInstance.time[ok], -BarInstance.pos[ok], s=0)
        if np.ndim(time) == 0:
            return f(time)
        return np.array([f(t) for t in time], dtype=np.float64)
    if np.ndim(time) == 0:
        return -BarInstance.pos[np.abs(time - BarInstance.time).argmin()]
    return np.array([-BarInstance.pos[np.abs(t - BarInstance.time).argmin()] for t in time])


def find_barpattern(intime, BarInstance, smth_order=2):
    """pattern.py:549-581: the pattern speed at the tabulated time nearest
    to intime (scalar or array), after frequency_and_derivative(smth_order)
    (which replaces any earlier derivative)."""
    BarInstance.frequency_and_derivative(smth_order=smth_order)
    if np.ndim(intime) == 0:
        return BarInstance.deriv[np.abs(intime - BarInstance.time).argmin()]
    return np.array([BarInstance.deriv[np.abs(t - BarInstance.time).argmin()] for t in intime])
