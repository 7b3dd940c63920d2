"""Orbit integration in a frozen field: drop-in for exptool.utils.integrate.

  leapfrog_integrate   utils/integrate.py:53-190, one orbit or many per launch
  transform / clock_transform   utils/integrate.py:37-50
  gen_init_step        utils/integrate.py:196-214
  compute_timestep     utils/integrate.py:217-272, many orbits per call
  do_integrate_multi, multi_compute_integration(_3D), integrate_grid(_launchy,
  _3D, _3D_launchy), integrate_grid_star*   utils/integrate.py:290-393,760-967
  redistribute_arrays, re_form_orbit_arrays(_3D)   utils/integrate.py:440-510
  print_orbit_array(_3D)   utils/integrate.py:513-568 (text orbit maps)
  read_integrations(_3D)   utils/integrate.py:970-1039
  run_time, run_time_mod   utils/integrate.py:571-758 (grids from a dump)
  Orbits               orbits/orbit.py:50-56 (dict of arrays)

The whole integration is one bfe_leapfrog call: the state stays in registers
on the device and the trajectory is recorded every `record_every` steps.  An
orbit grid (every radius x velocity [x z x vz] start) is one bfe_leapfrog_dt
call in which each orbit carries its own step size, in place of the
reference's process pool looping over orbits one at a time.
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


# ---------------------------------------------------------------------------
# orbit grids (integrate.py:196-567, 760-1028)
# ---------------------------------------------------------------------------

def gen_init_step(xpos, vtan, z0=0.0, zvel0=0.):
    """Start at planar apocentre on the x axis (integrate.py:196-214)."""
    return [xpos, 0.0, z0], [0.0, vtan, zvel0]


_TS_EPS = 1.e-10
_TS_NAMES = ("dtd", "dtv", "dta", "dtA")


def _timestep_criteria(comp, pos, vel, compat):
    """The four EXP time-scale criteria of integrate.py:241-259 on the device
    in float64: comp is the [8, n] return_forces_cart output at pos (3, n).
    compat=True keeps the reference's vertical term dfz*hfz (:249, :251) in
    place of the total vertical force dfz+hfz.  Returns (4, n)."""
    import torch
    c = comp.double()
    eps = _TS_EPS
    vtot = torch.sqrt((vel * vel).sum(0)) + eps
    fx, fy = c[0] + c[1], c[2] + c[3]
    fz = c[4] * c[5] if compat else c[4] + c[5]
    dtr = pos[0] * fx + pos[1] * fy + pos[2] * fz
    atot = torch.sqrt(fx * fx + fy * fy + fz * fz) + eps
    ptot = torch.abs(c[6] + c[7])
    return torch.stack((1.0 / torch.sqrt(vtot + eps), torch.sqrt(vtot / (atot + eps)),
                        ptot / (torch.abs(dtr) + eps), torch.sqrt(ptot / (atot * atot + eps))))


def _timesteps(FieldInstance, pos, vel, dyn_res, compat):
    """Per-orbit compute_timestep on the device: pos, vel float64 (3, n)
    arrays.  Returns ((n,) float64 device tensor of dt, (4, n) criteria)."""
    import torch
    if compat is None:
        compat = bool(getattr(FieldInstance, "compat", False))
    FieldInstance.freeze()
    sysm = FieldInstance.system
    p = torch.tensor(np.ascontiguousarray(pos, dtype=np.float64), device=sysm.device)
    v = torch.tensor(np.ascontiguousarray(vel, dtype=np.float64), device=sysm.device)
    # return_forces_cart at the start point, field not rotated (integrate.py:247)
    _, _, _, _, comp = sysm.field_forces(p[0].float(), p[1].float(), p[2].float(), rotpos=0.0,
                                         components=True)
    crit = _timestep_criteria(comp, p, v, compat)
    return crit.min(0).values / float(dyn_res), crit


def compute_timestep(FieldInstance, start_pos, start_vel, dyn_res=100., verbose=False, compat=None):
    """integrate.py:217-272: the smallest of the EXP criteria
    dtd = 1/sqrt(|v|), dtv = sqrt(|v|/|a|), dta = |Phi|/|x.a|,
    dtA = sqrt(|Phi|/|a|^2) (each with the reference's 1e-10 guards), divided
    by dyn_res.  Forces and potential come from the frozen field at the start
    point (float32 position on the device).

    start_pos/start_vel are (3,) for one orbit (returns a float) or (3, N)
    for N orbits in one field evaluation (returns an (N,) float64 array).
    compat (default: FieldInstance.compat) selects the reference's vertical
    term dfz*hfz; compat=False uses the total vertical force dfz+hfz.
    verbose prints the chosen criterion (the reference's print raises
    TypeError on Python 3: it indexes dict keys)."""
    pos = np.asarray(start_pos, dtype=np.float64)
    vel = np.asarray(start_vel, dtype=np.float64)
    single = pos.ndim == 1
    dt, crit = _timesteps(FieldInstance, pos.reshape(3, -1), vel.reshape(3, -1), dyn_res, compat)
    dt = dt.cpu().numpy()
    if verbose:
        which = crit.argmin(0).cpu().numpy()
        for i in range(dt.size):
            print('The chosen time criteria is {0:s}, dt={1:6.5f}'.format(_TS_NAMES[which[i]], dt[i]))
    return float(dt[0]) if single else dt


ROWS_2D = ("X", "Y", "TX", "TY", "VX", "VY", "T")
ROWS_3D = ("X", "Y", "Z", "TX", "TY", "VX", "VY", "VZ", "T")


def _grid_starts(rads, vels, zs=None, vzs=None, launch="x"):
    """Start points of an orbit grid in the reference's loop order (radius
    outermost, then velocity, z, vz): x launch [rad, 0, z], [0, vel, vz]
    (integrate.py:868-869, :784-785); y launch [0, rad, z], [0, vel, vz]
    (:906-907, :828-829).  Returns (shape, pos (3, N), vel (3, N))."""
    rads = np.atleast_1d(np.asarray(rads, dtype=np.float64))
    vels = np.atleast_1d(np.asarray(vels, dtype=np.float64))
    if zs is None:
        zs_, vzs_ = np.zeros(1), np.zeros(1)
        shape = (rads.size, vels.size)
    else:
        zs_ = np.atleast_1d(np.asarray(zs, dtype=np.float64))
        vzs_ = np.atleast_1d(np.asarray(vzs, dtype=np.float64))
        shape = (rads.size, vels.size, zs_.size, vzs_.size)
    R, V, Z, VZ = (g.ravel() for g in np.meshgrid(rads, vels, zs_, vzs_, indexing="ij"))
    zero = np.zeros_like(R)
    if launch == "x":
        pos = np.stack([R, zero, Z])
    elif launch == "y":
        pos = np.stack([zero, R, Z])
    else:
        raise ValueError(f"launch must be 'x' or 'y', not {launch!r}")
    return shape, pos, np.stack([zero, V, VZ])


def integrate_orbit_grid(F, pos, vel, nint, dt, rotfreq=0., no_odd=False, halo_l=-1, disk_m=-1,
                         dyn_res=100., threedee=False, compat=None, max_device_bytes=16 << 30):
    """Integrate N orbits (pos, vel (3, N) start points), each with step
    dtime = max(compute_timestep, dt) (integrate.py:870-871), for nint
    recorded steps, all in bfe_leapfrog_dt launches.  Returns the orbit rows
    (N, 7 | 9, nint) float64 in the integrate_grid layout (ROWS_2D, or
    ROWS_3D with threedee) and the (N,) step sizes.

    As in the reference every orbit runs the full nint steps: integrate_grid
    passes ap_max to leapfrog_integrate without apse=True, so no orbit stops
    early and the zero padding of the grid arrays never applies.  The field
    parameters (no_odd, halo_l, disk_m) are set before the time steps are
    computed; the reference's first orbit of each worker computes its time
    step with the parameters the field had before the call."""
    import torch
    F.set_field_parameters(no_odd=no_odd, halo_l=halo_l, disk_m=disk_m)
    F.freeze()
    sysm = F.system
    nint = int(nint)
    if nint < 1:
        raise ValueError("nint must be >= 1")
    pos = np.asarray(pos, dtype=np.float64).reshape(3, -1)
    vel = np.asarray(vel, dtype=np.float64).reshape(3, -1)
    norb = pos.shape[1]
    rows = ROWS_3D if threedee else ROWS_2D
    out = np.zeros((norb, len(rows), nint))
    dts = np.zeros(norb)
    if norb == 0:
        return out, dts
    per_orbit = nint * (7 * 4 + len(rows) * 8)
    chunk = max(1, int(max_device_bytes // per_orbit))
    tf = transform if rotfreq < 0.0 else clock_transform    # sign of barpos (integrate.py:180)
    for a in range(0, norb, chunk):
        b = min(norb, a + chunk)
        dt_i, _ = _timesteps(F, pos[:, a:b], vel[:, a:b], dyn_res, compat)
        dt_i = torch.clamp(dt_i, min=float(dt))
        state = [torch.tensor(np.ascontiguousarray(v, dtype=np.float32), device=sysm.device)
                 for v in (pos[0, a:b], pos[1, a:b], pos[2, a:b], vel[0, a:b], vel[1, a:b], vel[2, a:b])]
        res = sysm.leapfrog(state, dt=dt_i, nsteps=nint - 1, rotfreq=float(rotfreq), record_every=1)
        traj = res["traj"].double().permute(2, 1, 0)               # (n, 7, nint)
        T = dt_i[:, None] * torch.arange(nint, dtype=torch.float64, device=sysm.device)[None, :]
        barpos = 2.0 * np.pi * float(rotfreq) * T
        cb, sb = torch.cos(barpos), torch.sin(barpos)
        X, Y, Z, VX, VY, VZ = (traj[:, j] for j in range(6))
        if tf is transform:
            TX, TY = cb * X - sb * Y, sb * X + cb * Y
        else:
            TX, TY = cb * X + sb * Y, -1.0 * sb * X + cb * Y
        d = dict(X=X, Y=Y, Z=Z, TX=TX, TY=TY, VX=VX, VY=VY, VZ=VZ, T=T)
        out[a:b] = torch.stack([d[k] for k in rows], 1).cpu().numpy()
        dts[a:b] = dt_i.cpu().numpy()
    return out, dts


def _grid(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, zs, vzs, launch,
          threedee):
    if threedee and (zs is None or vzs is None):
        raise ValueError('ERROR: 3D orbit specified, but no z or vz values passed!')
    shape, pos, vel = _grid_starts(rads, vels, zs if threedee else None, vzs if threedee else None,
                                   launch)
    out, _ = integrate_orbit_grid(F, pos, vel, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res,
                                  threedee=threedee)
    return out.reshape(shape + out.shape[1:])


def integrate_grid(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, ap_max,
                   verbose):
    """integrate.py:848-884: (nrad, nvel, 7, nint) rows X, Y, TX, TY, VX, VY,
    T of orbits started at [rad, 0, 0], [0, vel, 0]; dt is the smallest
    step an orbit takes.  ap_max has no effect (see integrate_orbit_grid)."""
    return _grid(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, None, None,
                 "x", False)


def integrate_grid_launchy(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res,
                           ap_max, verbose):
    """integrate.py:886-922: integrate_grid launched from [0, rad, 0]."""
    return _grid(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, None, None,
                 "y", False)


def integrate_grid_3D(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, ap_max,
                      zs, vzs, verbose):
    """integrate.py:760-802: (nrad, nvel, nz, nvz, 9, nint) rows X, Y, Z, TX,
    TY, VX, VY, VZ, T of orbits started at [rad, 0, z], [0, vel, vz].  Missing
    zs or vzs raises ValueError (the reference prints and then fails on an
    unbound name)."""
    return _grid(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, zs, vzs,
                 "x", True)


def integrate_grid_3D_launchy(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res,
                              ap_max, zs, vzs, verbose):
    """integrate.py:808-846: integrate_grid_3D launched from [0, rad, z]."""
    return _grid(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, zs, vzs,
                 "y", True)


def integrate_grid_star(a_b):
    """integrate.py:379-381."""
    return integrate_grid(*a_b)


def integrate_grid_star_launchy(a_b):
    """integrate.py:383-385."""
    return integrate_grid_launchy(*a_b)


def integrate_grid_star_3D(a_b):
    """integrate.py:387-389."""
    return integrate_grid_3D(*a_b)


def integrate_grid_star_3D_launchy(a_b):
    """integrate.py:391-393."""
    return integrate_grid_3D_launchy(*a_b)


def redistribute_arrays(rads, divisions):
    """integrate.py:440-466: split rads into `divisions` consecutive parts of
    floor(n/divisions) values, the first part taking the remainder."""
    rads = np.asarray(rads)
    subrads = [[] for _ in range(divisions)]
    per = int(np.floor(rads.size / divisions))
    first = int(rads.size - per * (divisions - 1))
    print('Each processor has {} radii.'.format(per))
    for i in range(divisions):
        if i > 0:
            subrads[i] = rads[(i - 1) * per + first: i * per + first]
        else:
            subrads[i] = rads[0:first]
    return subrads


def re_form_orbit_arrays(array):
    """integrate.py:469-488: stack the per-part grids along the radius axis
    as float32 (the reference also prints the whole input first)."""
    return np.concatenate([np.asarray(a) for a in array], axis=0).astype('f4')


def re_form_orbit_arrays_3D(array):
    """integrate.py:490-510 (same for the 3D grids)."""
    return re_form_orbit_arrays(array)


def multi_compute_integration(subrads, nprocs, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m,
                              dyn_res, ap_max, launch, verbose=5):
    """integrate.py:338-376: the grid of every part of subrads, returned as
    a list of per-part arrays.  All parts are integrated in one device
    launch; nprocs only names how many parts are used."""
    parts = [np.atleast_1d(np.asarray(subrads[i], dtype=np.float64)) for i in range(nprocs)]
    full = _grid(np.concatenate(parts), vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m,
                 dyn_res, None, None, launch, False)
    cut = np.cumsum([0] + [p.size for p in parts])
    return [full[cut[i]:cut[i + 1]] for i in range(nprocs)]


def multi_compute_integration_3D(subrads, nprocs, vels, F, nint, dt, rotfreq, no_odd, halo_l,
                                 disk_m, dyn_res, ap_max, zs, vzs, launch, verbose=5):
    """integrate.py:924-967 (3D grids), one device launch."""
    parts = [np.atleast_1d(np.asarray(subrads[i], dtype=np.float64)) for i in range(nprocs)]
    full = _grid(np.concatenate(parts), vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m,
                 dyn_res, zs, vzs, launch, True)
    cut = np.cumsum([0] + [p.size for p in parts])
    return [full[cut[i]:cut[i + 1]] for i in range(nprocs)]


def do_integrate_multi(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, ap_max,
                       verbose=0, nprocs=-1, threedee=False, zs=None, vzs=None, launch='x'):
    """integrate.py:290-335: the orbit grid over rads x vels (x zs x vzs with
    threedee) as a float32 array (nrad, nvel, [nz, nvz,] 7 | 9, nint).  The
    reference splits rads over a process pool; here every orbit of the grid
    is integrated in one bfe_leapfrog_dt launch and nprocs is ignored."""
    rads = np.atleast_1d(np.asarray(rads, dtype=np.float64))
    vels = np.atleast_1d(np.asarray(vels, dtype=np.float64))
    if verbose > 0:
        print('Beginning integration of {} orbits on the device.'.format(
            rads.size * vels.size * (np.size(zs) * np.size(vzs) if threedee else 1)))
    t1 = time.time()
    arr = _grid(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, disk_m, dyn_res, zs, vzs, launch,
                bool(threedee)).astype('f4')
    el = time.time() - t1
    norb = max(1, int(np.prod(arr.shape[:-2])))
    print('Total integration calculation took {0:3.2f} seconds, or {1:3.2f} seconds per orbit.'.format(
        el, el / float(norb)))
    return arr


def _line(head, series):
    """One orbit-map line as the reference's print calls write it
    (integrate.py:518-536): the header values and every series value, each
    followed by one space, then a newline."""
    vals = [str(v) for v in head]
    for s in series:
        vals.extend(str(v) for v in s)
    return ' '.join(vals) + ' \n'


def print_orbit_array(f, OrbitArray):
    """integrate.py:513-536: one line per (rad, vel) orbit of a 2D grid:
    nsteps, R_0 = X[0], V_0 = VY[0], dT = T[1], then the X, Y, TX, TY, VX, VY
    series."""
    nsteps = OrbitArray.shape[-1]
    for rad in range(OrbitArray.shape[0]):
        for vel in range(OrbitArray.shape[1]):
            o = OrbitArray[rad, vel]
            f.write(_line((nsteps, o[0, 0], o[5, 0], o[-1, 1]), (o[j] for j in range(6))))


def print_orbit_array_3D(f, OrbitArray):
    """integrate.py:539-568: nsteps, R_0 = X[0], V_0 = VY[0], Z_0 = Z[0],
    VZ_0 = VZ[0], dT = T[1], then X, Y, Z, TX, TY, VX, VY, VZ per orbit."""
    nsteps = OrbitArray.shape[-1]
    for rad in range(OrbitArray.shape[0]):
        for vel in range(OrbitArray.shape[1]):
            for z in range(OrbitArray.shape[2]):
                for vz in range(OrbitArray.shape[3]):
                    o = OrbitArray[rad, vel, z, vz]
                    f.write(_line((nsteps, o[0, 0], o[6, 0], o[2, 0], o[7, 0], o[-1, 1]),
                                  (o[j] for j in range(8))))


def _read(infile, head, series):
    D = {k: {} for k in head + series}
    nh = len(head)
    with open(infile, 'r') as f:
        for linenum, line in enumerate(f):
            d = [float(q) for q in line.split()]
            npoints = int(d[0])
            for j, k in enumerate(head):
                D[k][linenum] = d[1 + j]
            for j, k in enumerate(series):
                D[k][linenum] = d[j * npoints + nh + 1:(j + 1) * npoints + nh + 1]
    return D


def read_integrations(infile):
    """integrate.py:970-999: the print_orbit_array format back as dicts
    keyed by line number (R_0, V_0, dT scalars; X ... VY lists)."""
    return _read(infile, ["R_0", "V_0", "dT"], ["X", "Y", "TX", "TY", "VX", "VY"])


def read_integrations_3D(infile):
    """integrate.py:1001-1039 (the print_orbit_array_3D format)."""
    return _read(infile, ["R_0", "V_0", "Z_0", "VZ_0", "dT"],
                 ["X", "Y", "Z", "TX", "TY", "VX", "VY", "VZ"])


# ---------------------------------------------------------------------------
# drivers from a simulation dump (integrate.py:571-758)
# ---------------------------------------------------------------------------

def run_time(simulation_directory, simulation_name, eof_file, sph_file, model_file, intime, rads,
             vels, nint, dt, no_odd, halo_l, max_m, dyn_res, ap_max, verbose, nprocs=-1,
             omegap=-1., orbitfile='', transform=True, fileprefix='OUT', launch='x'):
    """integrate.py:571-635: the field of one dump (potential.get_fields),
    the pattern speed (omegap >= 0 overrides it; the frame turns at
    -|patt| / 2 pi), the orbit grid, written as an orbit map to orbitfile or
    <dir>omap_<intime>.txt.  As in the reference no bar file is passed to
    get_fields, so transform=True needs one and fails."""
    from ..basis import potential
    if verbose:
        print('exptool.integrate.run_time: in directory {}, run {} at output {}, with transform={}'
              .format(simulation_directory, simulation_name, intime, transform))
    F, patt, rotfreq = potential.get_fields(simulation_directory, simulation_name, intime, eof_file,
                                            sph_file, model_file, transform=transform,
                                            fileprefix=fileprefix)
    if omegap >= 0.:
        patt = omegap
    rotfreq = -1. * abs(patt / (2. * np.pi))
    OrbitArray = do_integrate_multi(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, max_m, dyn_res,
                                    ap_max, verbose=verbose, nprocs=nprocs, launch=launch)
    path = orbitfile if orbitfile != '' else simulation_directory + 'omap_' + str(intime) + '.txt'
    with open(path, 'w') as f:
        print_orbit_array(f, OrbitArray)
    return OrbitArray


def run_time_mod(simulation_directory, simulation_name, eof_file, sph_file, model_file, intime,
                 rads, vels, nint, dt, no_odd, halo_l, max_m, dyn_res, ap_max, verbose, nprocs=-1,
                 omegap=-1., orbitfile='', transform=False, save_field=True, field_file_name='',
                 field_file=None, bar_file='', fileprefix='OUT', threedee=False, zs=None, vzs=None,
                 launch='x'):
    """integrate.py:638-758: run_time with the field saved to (or, when
    field_file is given, restored from) field_file_name, the pattern speed
    from bar_file, and 2D or 3D grids (orbit map omap_<intime>.txt or
    omap_3D_<intime>.txt)."""
    from ..analysis import pattern
    from ..basis import potential
    from ..io import particle
    if verbose:
        print('exptool.integrate.run_time: in directory {}, run {} at output {}, with transform={}'
              .format(simulation_directory, simulation_name, intime, transform))
    if transform and bar_file == '':
        print('error - no bar file supplied for transform')
        return None
    if field_file is None:
        F, patt, rotfreq = potential.get_fields(simulation_directory, simulation_name, intime,
                                                eof_file, sph_file, model_file, transform=transform,
                                                bar_file=bar_file, fileprefix=fileprefix)
        if save_field:
            if str(field_file_name) != '':
                F.save_field(str(field_file_name))
            else:
                print('saving field file with default name (field_file) to local directory')
                F.save_field('field_file')
    else:
        print('field file supplied! File name:')
        print(field_file_name)
        F = potential.restore_field(str(field_file_name))
        if transform:
            B = pattern.BarDetermine()
            B.read_bar(bar_file)
            B.frequency_and_derivative(spline_derivative=2)
            infile = simulation_directory + fileprefix + '.' + simulation_name + '.{0:05d}'.format(intime)
            P = particle.Input(infile, comp='star')
            patt = pattern.find_barpattern(P.time, B, smth_order=None)
        else:
            patt = 0.
    if omegap >= 0.:
        patt = omegap
    rotfreq = -1. * abs(patt / (2. * np.pi))
    if not threedee:
        print('launching from ', launch, ' axis')
        OrbitArray = do_integrate_multi(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, max_m,
                                        dyn_res, ap_max, verbose=verbose, nprocs=nprocs, launch=launch)
        path = orbitfile if orbitfile != '' else simulation_directory + 'omap_' + str(intime) + '.txt'
        with open(path, 'w') as f:
            print_orbit_array(f, OrbitArray)
        return OrbitArray
    OrbitArray = do_integrate_multi(rads, vels, F, nint, dt, rotfreq, no_odd, halo_l, max_m, dyn_res,
                                    ap_max, verbose=verbose, nprocs=nprocs, threedee=threedee, zs=zs,
                                    vzs=vzs, launch=launch)
    path = orbitfile if orbitfile != '' else simulation_directory + 'omap_3D_' + str(intime) + '.txt'
    with open(path, 'w') as f:
        print_orbit_array_3D(f, OrbitArray)
    return OrbitArray
