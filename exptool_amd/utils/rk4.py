"""Fourth-order Runge-Kutta orbits in a frozen field: drop-in for exptool.utils.rk4.

  rk4                        utils/rk4.py:26-63
  add_rotation_acceleration  utils/rk4.py:66-90
  rk4_rotation               utils/rk4.py:94-135

The reference integrates one orbit in any potential object that has
get_cartesian_forces_array.  Here the potential is the frozen disc + halo
field of an exptool_amd Fields (or a BFESystem with a field set), and the
integration is one bfe_rk4 call for one orbit or for a batch of orbits: the
state stays in registers on the device, four field evaluations per step.
"""
from __future__ import annotations

import numpy as np


def add_rotation_acceleration(xarr, vxarr, omegab=1):
    """rk4.py:66-90 as the reference computes it, in the plane:
    -omegab^2 (x, y, 0) + 2 omegab z^ x (vx, vy, 0).  Both terms have the
    opposite sign of the pseudo-forces of a rotating frame (Binney &
    Tremaine 3.117: +Omega^2 R - 2 Omega x v)."""
    e1 = np.array([0, 0, 1])
    x, y = xarr[0], xarr[1]
    vx, vy = vxarr[0], vxarr[1]
    coriolis = 2 * np.cross(omegab * e1, np.array([vx, vy, 0]))
    centrifugal = omegab * omegab * np.array([x, y, 0])
    return -centrifugal + coriolis


def _system(P):
    from ..bfe import BFESystem
    if isinstance(P, BFESystem):
        return P
    if hasattr(P, "freeze") and hasattr(P, "system"):
        P.freeze()
        return P.system
    raise TypeError("rk4 integrates in the frozen field of an exptool_amd Fields or BFESystem")


def _integrate(P, ri, vi, dt, n, omegab, compat_frame):
    import torch
    sysm = _system(P)
    pos = np.asarray(ri, dtype=np.float64)
    vel = np.asarray(vi, dtype=np.float64)
    single = pos.ndim == 1
    pos, vel = pos.reshape(3, -1), vel.reshape(3, -1)
    n = int(n)
    if n < 1:
        raise ValueError("n must be >= 1")
    state = [torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device=sysm.device)
             for a in (pos[0], pos[1], pos[2], vel[0], vel[1], vel[2])]
    traj = sysm.rk4(state, dt=float(dt), nsteps=n - 1, omegab=float(omegab),
                    compat_frame=compat_frame, record_every=1)
    t = traj.double().cpu().numpy()                  # (n, 6, norb)
    r = np.transpose(t[:, 0:3, :], (1, 2, 0))       # (3, norb, n)
    v = np.transpose(t[:, 3:6, :], (1, 2, 0))
    if single:
        return r[:, 0, :], v[:, 0, :]
    return r, v


def rk4(P, ri, vi, dt, n):
    """rk4.py:26-63: n states (step 0 included) of classical RK4 with step dt
    in the field of P.  ri, vi are (3,) for one orbit, as in the reference
    (returns r, v each (3, n)), or (3, N) for N orbits (each (3, N, n))."""
    return _integrate(P, ri, vi, dt, n, 0.0, False)


def rk4_rotation(P, ri, vi, dt, n, omegab=1, compat=None):
    """rk4.py:94-135: rk4 in a frame rotating at omegab about z, the field of
    P fixed in that frame.  compat (default: P.compat) uses the reference's
    add_rotation_acceleration signs; otherwise the physical pseudo-forces
    +omegab^2 R - 2 omegab z^ x v."""
    if compat is None:
        compat = bool(getattr(P, "compat", False))
    return _integrate(P, ri, vi, dt, n, omegab, bool(compat))
