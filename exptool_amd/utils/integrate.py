"""Orbit integration in a frozen field: drop-in for exptool.utils.integrate.

  leapfrog_integrate   utils/integrate.py:53-190, one orbit or many per launch
  transform / clock_transform   utils/integrate.py:37-50
  Orbits               orbits/orbit.py:50-56 (dict of arrays)

The whole integration is one bfe_leapfrog call: the state stays in registers
on the device and the trajectory is recorded every `record_every` steps.
"""
from __future__ import annotations

import time

import numpy as np


class Orbits(dict):
    """Dictionary of orbit arrays (orbits/orbit.py:50-56)."""


def transform(xarray, yarray, thetas):
    """Counter-clockwise planar rotation (integrate.py:37-41)."""
    return (np.cos(thetas) * xarray - np.sin(thetas) * yarray,
            np.sin(thetas) * xarray + np.cos(thetas) * yarray)


def clock_transform(xarray, yarray, thetas):
    """Clockwise planar rotation (integrate.py:44-48)."""
    return (np.cos(thetas) * xarray + np.sin(thetas) * yarray,
            -1.0 * np.sin(thetas) * xarray + np.cos(thetas) * yarray)


def _apse_stop(X, Y, ap_max):
    """Number of recorded steps the reference keeps with apse=True
    (integrate.py:145-153): an apocentre is counted at step s when the planar
    radius at s-1 exceeds those at s-2 and s; the loop ends after the step
    that brings the count to ap_max."""
    r2 = X * X + Y * Y
    n = 0
    for s in range(2, r2.size):
        if r2[s - 1] > r2[s - 2] and r2[s - 1] > r2[s]:
            n += 1
            if n >= ap_max:
                return s + 1
    return r2.size


def leapfrog_integrate(FieldInstance, nint, dt, initpos, initvel, rotfreq=0., no_odd=False,
                       halo_l=-1, halo_n=-1, disk_m=-1, disk_n=-1, verbose=0, force=False,
                       ap_max=1000, apse=False, record_every=1, max_record_bytes=32 << 30):
    """integrate.py:53-190 on the GPU.

    FieldInstance is an exptool_amd.basis.potential.Fields.  initpos/initvel
    are (3,) for one orbit, as in the reference, or (3, N) for N orbits in
    one launch (then every array in the result is (N, nrec)).  The update is
    the reference's: x += v dt + a dt^2/2; a' = F(x', 2 pi rotfreq t);
    v += (a + a') dt/2, with t_k = k dt.  State is float32 on the device,
    energies and the field combination float64.  record_every > 1 keeps
    every record_every-th step (the reference keeps all)."""
    import torch
    FieldInstance.set_field_parameters(no_odd=no_odd, halo_l=halo_l, halo_n=halo_n, disk_m=disk_m,
                                       disk_n=disk_n)
    FieldInstance.freeze()
    sysm = FieldInstance.system
    pos = np.asarray(initpos, dtype=np.float64)
    vel = np.asarray(initvel, dtype=np.float64)
    single = pos.ndim == 1
    pos = pos.reshape(3, -1)
    vel = vel.reshape(3, -1)
    norb = pos.shape[1]
    nint = int(nint)
    if nint < 1:
        raise ValueError("nint must be >= 1")
    nrec = (nint - 1) // record_every + 1
    if nrec * 7 * norb * 4 > max_record_bytes:
        raise ValueError(f"trajectory of {nrec} records x {norb} orbits exceeds {max_record_bytes} "
                         "bytes; raise record_every or max_record_bytes")
    t0 = time.time()
    state = [torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device=sysm.device)
             for a in (pos[0], pos[1], pos[2], vel[0], vel[1], vel[2])]
    out = sysm.leapfrog(state, dt=float(dt), nsteps=nint - 1, t0=0.0, rotfreq=float(rotfreq),
                        energy_every=0, record_every=int(record_every))
    traj = out["traj"].double().cpu().numpy()           # (nrec, 7, norb)
    if verbose:
        print('{0:4.3f} seconds to integrate.'.format(time.time() - t0))
    steps = np.arange(nrec) * record_every
    times = steps * dt
    barpos = 2.0 * np.pi * rotfreq * times
    O = Orbits()
    keys = ("X", "Y", "Z", "VX", "VY", "VZ", "P")
    arr = {k: traj[:, j, :].T for j, k in enumerate(keys)}        # (norb, nrec)
    if force:
        # accelerations at the recorded positions: the field at angle
        # phi + barpos equals the field at the point rotated by barpos, with
        # the Cartesian force rotated back
        X, Y, Z = arr["X"], arr["Y"], arr["Z"]
        cb, sb = np.cos(barpos)[None, :], np.sin(barpos)[None, :]
        xr, yr = X * cb - Y * sb, X * sb + Y * cb
        fx, fy, fz, _ = sysm.field_forces(xr.ravel(), yr.ravel(), Z.ravel())
        fx = fx.double().cpu().numpy().reshape(X.shape)
        fy = fy.double().cpu().numpy().reshape(X.shape)
        arr["FX"] = fx * cb + fy * sb
        arr["FY"] = -fx * sb + fy * cb
        arr["FZ"] = fz.double().cpu().numpy().reshape(X.shape)
    nkeep = np.full(norb, nrec)
    if apse:
        if record_every != 1:
            raise ValueError("apse needs record_every=1")
        nkeep = np.array([_apse_stop(arr["X"][i], arr["Y"][i], ap_max) for i in range(norb)])
    T = np.broadcast_to(times, (norb, nrec)).copy()
    bp = np.broadcast_to(barpos, (norb, nrec))
    tf = transform if np.min(barpos) < 0.0 else clock_transform
    TX, TY = tf(arr["X"], arr["Y"], bp)
    VTX, VTY = tf(arr["VX"], arr["VY"], bp)
    full = dict(T=T, **arr, TX=TX, TY=TY, VTX=VTX, VTY=VTY)
    if apse:
        for k, v in full.items():
            for i in range(norb):
                v[i, nkeep[i]:] = np.nan
    if single:
        n = int(nkeep[0])
        for k, v in full.items():
            O[k] = v[0, :n]
    else:
        for k, v in full.items():
            O[k] = v
        O["NSTEPS"] = nkeep
    return O
