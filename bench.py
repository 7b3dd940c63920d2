#!/usr/bin/env python
"""Benchmark of the BFE hot path on one MI355X.

Workload (one "step"): 1M exponential-disc particles and 1M Hernquist-halo
particles resident in HBM as float32 SoA.  A step accumulates the disc EOF
coefficients (mmax 6, norder 18, 256x128 tables) and the halo SL coefficients
(lmax 6, nmax 20, numr 2000), freezes the disc+halo field (folds the
coefficients into the tables) and evaluates Phi, F at all 2M particles
(Fields.return_forces_cart semantics).  value = particles / s through the
whole step (accumulate + force evaluation of every particle).

Also reported: per-kernel HIP-event timings, the roofline record of the
dominant kernel, orbit-steps/s of the plain-BFE leapfrog in the frozen field,
the CPU baselines on the host cores (the reference's decomposition restated
over a process pool, the C++ FP64 OpenMP restatement, the NumPy oracle on
one core) and the step's error against FP64 on a sample of its particles.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0          # dense bf16 MFMA
PMC_SUMMARY = os.path.join(REPO, "profiles", "pmc_summary.json")


# Reference-formulation work per unit (M = mmax, N = norder, L = lmax, n = nmax):
# FLOPs of the unfolded sums the reference evaluates, and the compulsory HBM
# bytes (float32 SoA in, float32 out).
def eof_acc_flops(M, N):
    return 10 * (2 * M + 1) * N


def sl_acc_flops(L, n):
    return 2 * (L + 1) ** 2 * n + 4 * (L + 1) * n


def field_ref_flops(M, N, L, n):
    return 32 * (2 * M + 1) * N + 8 * (L + 1) ** 2 * n + 12 * (L + 1) * n


def pmc_summary():
    """Executed-flop counts and HBM traffic per dispatch measured with
    rocprofv3 PMC counters for this workload (profiles/pmc_summary.json)."""
    try:
        with open(PMC_SUMMARY) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-disc", type=int, default=1_000_000)
    ap.add_argument("--n-halo", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pool-cores", type=int, default=min(16, os.cpu_count() or 1),
                    help="pool workers of the reference-path CPU baseline (the host's CPU share)")
    ap.add_argument("--orbit-steps", type=int, default=10_000)
    ap.add_argument("--orbits", type=int, default=1_000_000)
    ap.add_argument("--mlp", type=int, default=1, help="train and time the MLP surrogate (0/1)")
    ap.add_argument("--mlp-points", type=int, default=16_000_000)
    ap.add_argument("--mlp-disc-points", type=int, default=4_000_000,
                    help="training points drawn from the disc model, beside the uniform ones")
    ap.add_argument("--mlp-train-seconds", type=float, default=120.0,
                    help="guard on the training time (the step count ends it first)")
    ap.add_argument("--mlp-steps", type=int, default=60000)
    ap.add_argument("--batch", type=int, default=1, help="time the batched-snapshot workload (0/1)")
    ap.add_argument("--batch-nsnap", type=int, default=64)
    ap.add_argument("--tune", type=int, default=1, help="autotune accumulate launch shapes (0/1)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (reference for PMC byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "exptool_amd runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    import torch
    from exptool_amd.bfe import BFESystem
    from exptool_amd.utils import workloads as W
    from exptool_amd.utils import samplers as S

    if a.copy_calibration > 0:
        nel = a.copy_calibration * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    t_setup = time.time()
    sl_t, eof_t = W.disc_halo_tables(seed=1)
    (dxyzm, hxyzm) = W.disc_halo_particles(a.n_disc, a.n_halo, seed=0)
    sysm = BFESystem(sl=sl_t, eof=eof_t, device=0)
    dev = sysm.device
    T = lambda v: torch.as_tensor(np.asarray(v, np.float32), device=dev)
    dx, dy, dz, dm = (T(v) for v in dxyzm)
    hx, hy, hz, hm = (T(v) for v in hxyzm)
    ax, ay, az = (torch.cat([p, q]) for p, q in ((dx, hx), (dy, hy), (dz, hz)))
    ntot = ax.numel()
    setup_s = time.time() - t_setup

    tuned = None
    if a.tune:
        from exptool_amd.autotune import autotune
        tuned = autotune(sysm, disc=(dx, dy, dz, dm), halo=(hx, hy, hz, hm), cache=False)

    def step():
        # Fields.total_coefficients (disc and halo accumulate, run concurrently,
        # each folded into the frozen field as soon as it is reduced), then
        # return_forces_cart at every particle: one bfe_forces_disc_halo call
        return sysm.forces_disc_halo((dx, dy, dz, dm), (hx, hy, hz, hm))

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    value = ntot / ((t1 - t0) / a.steps)
    progress(f"step: {ms_per_step:.3f} ms")

    # ---- per-kernel timings (HIP events on the launch stream) ---------------
    def timed(fn, reps):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps * 1e-3

    c, s = sysm.eof_accumulate(dx, dy, dz, dm)
    h = sysm.sl_accumulate(hx, hy, hz, hm)
    reps = max(3, a.steps)
    kt = {
        "eof_accumulate": timed(lambda: sysm.eof_accumulate(dx, dy, dz, dm), reps),
        "sl_accumulate": timed(lambda: sysm.sl_accumulate(hx, hy, hz, hm), reps),
        "field_set": timed(lambda: sysm.set_field(c, s, h), reps),
        "field_eval_cart": timed(lambda: sysm.field_forces(ax, ay, az), reps),
    }
    M, N = eof_t["mmax"], eof_t["norder"]
    L, n = sl_t["lmax"], sl_t["nmax"]
    # units per launch, reference-formulation flops and compulsory HBM bytes
    units = {
        "eof_accumulate": (a.n_disc, eof_acc_flops(M, N), 16),
        "sl_accumulate": (a.n_halo, sl_acc_flops(L, n), 16),
        "field_eval_cart": (ntot, field_ref_flops(M, N, L, n), 28),
    }
    # the compute kernel alone (HIP events recorded by the library on the
    # launch stream around that kernel; cell sort and reductions excluded)
    lib_name = {"eof_accumulate": "eof_accumulate", "sl_accumulate": "sl_accumulate",
                "field_eval_cart": "field_eval"}
    sysm.set_tuning("kernel_timing", 1)
    for k in lib_name.values():
        sysm.kernel_times(k)
    for _ in range(reps):
        sysm.eof_accumulate(dx, dy, dz, dm)
        sysm.sl_accumulate(hx, hy, hz, hm)
        sysm.field_forces(ax, ay, az)
    ko = {}
    for k, nm in lib_name.items():
        tot, cnt = sysm.kernel_times(nm)
        ko[k] = tot / max(cnt, 1)
    sysm.set_tuning("kernel_timing", 0)
    # the two accumulate modes of each basis at the full 1M (call time; the
    # step runs the float32 ones): float32 sums, and bf16 MFMA products of
    # hi + lo operand pieces (errors against FP64 in accuracy_vs_fp64)
    modes = {
        "eof_f32": timed(lambda: sysm.eof_accumulate(dx, dy, dz, dm), reps),
        "eof_bf16_mfma": timed(lambda: sysm.eof_accumulate(dx, dy, dz, dm, bf16=True), reps),
        "sl_f32": timed(lambda: sysm.sl_accumulate(hx, hy, hz, hm), reps),
        "sl_bf16_mfma": timed(lambda: sysm.sl_accumulate(hx, hy, hz, hm, bf16=True), reps),
    }
    sysm.set_field(c, s, h)
    dom = max(units, key=lambda k: ko[k])
    roof = roofline(dom, units[dom], ko[dom], pmc_summary())

    progress(f"kernels timed, dominant {dom}")
    # ---- orbit integration in the frozen field (plain BFE and MLP surrogate) --
    orbit = None
    surrogate = None
    if a.orbit_steps > 0 and a.orbits > 0:
        sysm.set_field(c, s, h)
        no = a.orbits
        ox, oy, oz, _ = S.exponential_disc(no, seed=5, **W.DISC)
        Rr = np.sqrt(ox * ox + oy * oy)
        fx, fy, fz, _ = sysm.field_forces(Rr, np.zeros_like(Rr), np.zeros_like(Rr))
        vx, vy, vz = S.circular_orbits(ox, oy, oz, fx.cpu().numpy().astype(np.float64), 0.1, seed=6)
        st0 = [T(v) for v in (ox, oy, oz, vx, vy, vz)]
        sysm.leapfrog([t.clone() for t in st0], dt=1e-3, nsteps=10, energy_every=0)
        torch.cuda.synchronize()
        st = [t.clone() for t in st0]
        t0 = time.perf_counter()
        out = sysm.leapfrog(st, dt=1e-3, nsteps=a.orbit_steps, energy_every=100)
        torch.cuda.synchronize()
        dt_o = time.perf_counter() - t0
        dr = out["drift"].cpu().numpy()
        orbit = {"orbit_steps_per_s": no * a.orbit_steps / dt_o, "orbits": no,
                 "steps": a.orbit_steps, "dt": 1e-3, "seconds": dt_o,
                 "energy_drift_median": float(np.median(dr)),
                 "energy_drift_p99": float(np.quantile(dr, 0.99)),
                 "energy_drift_max": float(np.max(dr))}
        progress(f"orbits: {orbit['orbit_steps_per_s']:.3e} orbit-steps/s")
        # per-orbit step sizes (orbit grids) and RK4 (four evaluations a step)
        st = [t.clone() for t in st0]
        dtv = torch.full((no,), 1e-3, dtype=torch.float64, device=st[0].device)
        t0 = time.perf_counter()
        sysm.leapfrog(st, dt=dtv, nsteps=a.orbit_steps, energy_every=100)
        torch.cuda.synchronize()
        orbit["per_orbit_dt_steps_per_s"] = no * a.orbit_steps / (time.perf_counter() - t0)
        nrk = max(1, a.orbit_steps // 4)
        st = [t.clone() for t in st0]
        sysm.rk4([t.clone() for t in st0], dt=4e-3, nsteps=4)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sysm.rk4(st, dt=4e-3, nsteps=nrk)
        torch.cuda.synchronize()
        orbit["rk4_steps_per_s"] = no * nrk / (time.perf_counter() - t0)
        progress(f"per-orbit dt: {orbit['per_orbit_dt_steps_per_s']:.3e} orbit-steps/s, "
                 f"rk4: {orbit['rk4_steps_per_s']:.3e} steps/s")
        if a.mlp:
            surrogate = mlp_stage(a, sysm, st0, orbit)
            progress(f"surrogate: {surrogate['orbit_steps_per_s']:.3e} orbit-steps/s")

    # ---- batched snapshots ------------------------------------------------------
    batch = batch_stage(a, sysm) if a.batch else None
    if batch:
        progress(f"batch: {batch['particles_per_s']:.3e} particles/s")

    # ---- CPU baseline: the FP64 oracle on a bounded sample --------------------
    cpu = cpu_all = acc = None
    if not a.no_cpu:
        cpu, cpu_all, ref = cpu_baselines(sl_t, eof_t, dxyzm, hxyzm, a.cpu_seconds,
                                          a.pool_cores)
        acc = accuracy(sysm, dxyzm, hxyzm, ref)
        progress(f"accuracy: force max {acc['force_rel_max']:.2e}, coef "
                 f"{acc['coef_eof_block_rel_max']:.2e} / {acc['coef_sl_block_rel_max']:.2e}")

    line = {
        "metric": "particle force evals/s (accumulate + Phi,F eval, disc EOF + halo SL)",
        "value": value, "unit": "particles/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 geometry and cross-block sums)",
        "data": "synthetic (seeded exponential disc + Hernquist halo)",
        "config": {"workload": "disc EOF mmax6 norder18 256x128 + halo SL lmax6 nmax20 numr2000, "
                               "1M+1M particles, accumulate + return_forces_cart at 2M",
                   "n_disc": a.n_disc, "n_halo": a.n_halo, "parallelism": "single GPU"},
        "call_seconds": kt, "kernel_seconds": ko, "accumulate_mode_seconds": modes,
        "roofline": roof, "tuning": tuned, "orbits": orbit,
        "surrogate": surrogate, "batch": batch,
        "cpu_baseline": cpu, "cpu_baselines": cpu_all, "accuracy_vs_fp64": acc,
        "setup_seconds": setup_s,
    }
    print(json.dumps(line))


def roofline(kernel, unit, seconds, pmc):
    """Roofline record of the dominant kernel.  achieved = executed flops per
    launch (the kernels evaluate exact rewrites of the reference sums, so their
    own work, counted by PMC, is what the hardware did) / HIP-event time; the
    bound is the vector ALU of the precision that dominates the executed flops.
    traffic = HBM bytes per launch from the FETCH_SIZE/WRITE_SIZE passes."""
    nunits, ref_flops, hbm_bytes = unit
    rec = {"kernel": kernel, "units_per_launch": nunits,
           "reference_flops_per_unit": ref_flops, "algorithmic_hbm_bytes_per_unit": hbm_bytes,
           "algorithmic_hbm_gbs": nunits * hbm_bytes / seconds / 1e9}
    ex = (pmc or {}).get("executed_flops", {}).get(kernel)
    tr = (pmc or {}).get("traffic", {}).get(kernel)
    if ex:
        # both precisions share the vector ALU: the float32 flops count at the
        # ratio of the peaks, so frac is the share of the kernel's time the
        # executed work needs at the published f64 and f32 vector peaks
        f64, f32 = ex["f64_flops_per_unit"], ex["f32_flops_per_unit"]
        use64 = f64 >= f32
        if use64:
            per_unit, peak = f64 + f32 * FP64_PEAK_TFLOPS / FP32_PEAK_TFLOPS, FP64_PEAK_TFLOPS
        else:
            per_unit, peak = f32 + f64 * FP32_PEAK_TFLOPS / FP64_PEAK_TFLOPS, FP32_PEAK_TFLOPS
        achieved = nunits * per_unit / seconds / 1e12
        rec.update({"bound": "valu_f64" if use64 else "valu_f32", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak, "flops_per_unit": per_unit,
                    "flops_per_unit_f64": f64, "flops_per_unit_f32": f32,
                    "flops_precision": ("f64-equivalent (f64 + f32 x 78.6/157.3)" if use64
                                        else "f32-equivalent (f32 + f64 x 157.3/78.6)"),
                    "flops_source": "executed, rocprofv3 SQ_INSTS_VALU_* (profiles/bench_step_pmc_valu.md)"})
    else:
        achieved = rec["algorithmic_hbm_gbs"]
        rec.update({"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS})
    rec["traffic"] = tr["bytes"] if tr else None
    if tr:
        rec["traffic_source"] = "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated (profiles/bench_step_pmc_traffic.md)"
        rec["traffic_gbs"] = tr["bytes"] / seconds / 1e9
    return rec


def batch_stage(a, sysm):
    """Batched snapshots: nsnap disc + halo snapshots (disc counts log-uniform
    in [200k, 4M], barred discs), every snapshot's EOF and SL coefficients
    accumulated in one launch each, every snapshot's field frozen in one call
    and every particle evaluated in its own snapshot's field."""
    import torch
    from exptool_amd.autotune import tune_accumulate
    from exptool_amd.utils import workloads as W
    t0 = time.perf_counter()
    b = W.batch_snapshots(nsnap=a.batch_nsnap, seed=0, device=sysm.device)
    torch.cuda.synchronize()
    t_sample = time.perf_counter() - t0
    D, H = b["disc"], b["halo"]
    dc, hc = b["disc_counts"], b["halo_counts"]
    tuned = {}
    if a.tune:
        for kind, P, cnt in (("eof", D, dc), ("sl", H, hc)):
            r = tune_accumulate(sysm, *P, kind=kind, counts=cnt, reps=1, candidates=(2048, 4096))
            tuned.update({k: v for k, v in r.items() if k != "timings"})
    stages = {}

    def run(rec):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        ev[0].record()
        c, s = sysm.eof_accumulate_batch(*D, dc)
        ev[1].record()
        h = sysm.sl_accumulate_batch(*H, hc)
        ev[2].record()
        sysm.set_field_batch(c, s, h)
        ev[3].record()
        sysm.field_forces_batch(D[0], D[1], D[2], dc)
        ev[4].record()
        sysm.field_forces_batch(H[0], H[1], H[2], hc)
        ev[5].record()
        if rec:
            torch.cuda.synchronize()
            for k, (i, j) in {"eof_accumulate_batch": (0, 1), "sl_accumulate_batch": (1, 2),
                              "field_set_batch": (2, 3), "field_eval_batch_disc": (3, 4),
                              "field_eval_batch_halo": (4, 5)}.items():
                stages.setdefault(k, []).append(ev[i].elapsed_time(ev[j]) * 1e-3)

    run(False)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        run(True)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    ntot = int(dc.sum() + hc.sum())
    return {"workload": f"{a.batch_nsnap} disc+halo snapshots, disc counts log-uniform in [2e5, 4e6] "
                        "(m=2 bar, A in [0, 0.3]), halo counts equal; same bases as the step; every "
                        "particle accumulated once and evaluated in its own snapshot's field",
            "particles": ntot, "snapshots": a.batch_nsnap, "seconds_per_pass": sec,
            "particles_per_s": ntot / sec,
            "stage_seconds": {k: float(np.mean(v)) for k, v in stages.items()},
            "sample_seconds": t_sample, "tuning": tuned}


def mlp_stage(a, sysm, st0, orbit):
    """Train the MLP surrogate on BFE evaluations (time-bounded), measure its
    force error on held-out points and integrate the same orbits with it."""
    import torch
    from exptool_amd.surrogate.mlp import MLPField, sample_domain
    from exptool_amd.utils import samplers as S
    from exptool_amd.utils import workloads as W
    # conservative field (forces = -grad of the learned potential + base);
    # the 16M uniform points of the training domain plus points drawn from
    # the disc the orbits start in, all labelled by the BFE
    m = MLPField(sysm, seed=0, conservative=True)
    ex, ey, ez, _ = S.exponential_disc(a.mlp_disc_points, seed=123, **W.DISC)
    rep = m.fit(npoints=a.mlp_points, steps=a.mlp_steps, batch=65536, lr=2e-3,
                time_budget=a.mlp_train_seconds, seed=11, extra=(ex, ey, ez))
    hx, hy, hz = sample_domain(1_000_000, m.rmax, m.zmax, seed=99, device=sysm.device)
    ferr, perr = m.force_error(hx, hy, hz)
    no = st0[0].numel()
    m.leapfrog([t.clone() for t in st0], dt=1e-3, nsteps=10, energy_every=0)
    torch.cuda.synchronize()
    st = [t.clone() for t in st0]
    t0 = time.perf_counter()
    out = m.leapfrog(st, dt=1e-3, nsteps=a.orbit_steps, energy_every=100, check_period=100)
    torch.cuda.synchronize()
    dt_o = time.perf_counter() - t0
    dr = out["drift"].cpu().numpy()
    me = out["maxerr"].cpu().numpy()
    ck = out["checked"].cpu().numpy()
    fb = out["fallback"].cpu().numpy()
    rate = no * a.orbit_steps / dt_o
    # the analytic base alone through the same kernels (zero network)
    keep = (m.params.copy(), m.cfg.copy())
    m.params = np.zeros_like(m.params)
    m.set_scaling([1, 1, 1, 1], [0, 0, 0, 0])
    m.upload()
    berr, _ = m.force_error(hx, hy, hz)
    m.params, m.cfg = keep
    m.upload()
    # network cost per evaluation: the forward pass 2 (3*128 + 3*128^2 + 128*4)
    # = 100 096 flops; the conservative mode also carries the three tangent
    # directions through the hidden layers (3 x 3 x 2*128^2) and the
    # potential row of the output (3 x 2*128): 395 776 flops.  At the dense
    # bf16 MFMA peak that bounds the surrogate's step rate.
    flops_eval = 100096.0 + 3 * (3 * 2 * 128 * 128 + 2 * 128)
    ceiling = BF16_PEAK_TFLOPS * 1e12 / flops_eval
    return {"mode": "conservative (F = -grad of the learned potential + analytic base)",
            "orbit_steps_per_s": rate, "speedup_vs_bfe": rate / orbit["orbit_steps_per_s"],
            "network_flops_per_eval": flops_eval,
            "mfma_frac": rate * flops_eval / (BF16_PEAK_TFLOPS * 1e12),
            "speedup_ceiling_at_mfma_peak": ceiling / orbit["orbit_steps_per_s"],
            "speedup_ceiling_direct_mode": BF16_PEAK_TFLOPS * 1e12 / 100096.0 / orbit["orbit_steps_per_s"],
            "train_disc_points": a.mlp_disc_points,
            "base": rep.get("base"), "base_fit_seconds": rep.get("base_seconds"),
            "force_rel_err_base_alone_median": float(np.median(berr)),
            "seconds": dt_o, "train_seconds": rep["train_seconds"], "train_steps": rep["steps"],
            "train_samples": rep["samples"], "train_points": rep["npoints"],
            "force_rel_err_heldout": {"median": float(np.median(ferr)),
                                      "p99": float(np.quantile(ferr, 0.99)),
                                      "max": float(np.max(ferr))},
            "pot_rel_err_heldout_median": float(np.median(perr)),
            "checked_force_err": {"orbits_checked": int((ck > 0).sum()),
                                  "median_of_max": float(np.median(me[ck > 0])) if (ck > 0).any() else None,
                                  "max": float(np.max(me[ck > 0])) if (ck > 0).any() else None},
            "fallback_evals_fraction": float(fb.sum()) / (no * (a.orbit_steps + 1)),
            "energy_drift_median": float(np.median(dr)),
            "energy_drift_p99": float(np.quantile(dr, 0.99)),
            "energy_drift_max": float(np.max(dr))}


def host_info():
    """Host CPU identity for the baseline records."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = None
    return {"os_cpu_count": os.cpu_count(), "affinity_cpus": aff, "model": model}


def _oracle_tables(sl_t, eof_t):
    from oracle import bfe_oracle as O
    slo = O.SLTable(sl_t["lmax"], sl_t["nmax"], sl_t["numr"], sl_t["cmap"], sl_t["rmin"],
                    sl_t["rmax"], sl_t["scale"], sl_t["ev"], sl_t["ef"], sl_t["xi"], sl_t["p0"],
                    sl_t["d0"])
    eo = O.EOFTable(eof_t["mmax"], eof_t["norder"], eof_t["numx"], eof_t["numy"], eof_t["cmap"],
                    eof_t["rmin"], eof_t["rmax"], eof_t["ascale"], eof_t["hscale"], eof_t["potC"],
                    eof_t["rforceC"], eof_t["zforceC"], eof_t["potS"], eof_t["rforceS"],
                    eof_t["zforceS"])
    return slo, eo


def cpu_baselines(sl_t, eof_t, dxyzm, hxyzm, budget_s, pool_cores):
    """The CPU baselines, timed on this host:
      reference path  the reference's decomposition restated (oracle/ref_path.py):
                      a pool of pool_cores workers over contiguous blocks,
                      per-particle SL loops, vectorised EOF blocks, per-point
                      field evaluation; stage costs measured on samples and
                      extrapolated linearly to the full workload;
      openmp          the C++ FP64 OpenMP restatement (oracle/bfe_cpu.cpp) on
                      all threads, on a bounded sample;
      numpy_1core     the vectorised NumPy oracle on one thread.
    Returns (reference-path record, dict of all records, the OpenMP outputs on
    its sample for the accuracy check)."""
    from oracle import bfe_oracle as O
    from oracle import cpu_oracle as CO
    from oracle import ref_path as RP
    slo, eo = _oracle_tables(sl_t, eof_t)
    host = host_info()
    nd, nh = dxyzm[0].size, hxyzm[0].size
    # ---- OpenMP, bounded sample of the same particles
    k = min(200_000, nd, nh)
    d = [v[:k] for v in dxyzm]
    h = [v[:k] for v in hxyzm]
    t0 = time.perf_counter()
    c, s = CO.eof_accumulate(eo, *d)
    hc = CO.sl_accumulate(slo, *h)
    F = O.FieldSpec(eo, c, s, slo, hc)
    pts = [np.concatenate([a, b]) for a, b in zip(d[:3], h[:3])]
    forces = CO.total_force(F, *pts)
    t_omp = time.perf_counter() - t0
    omp = {"value": 2 * k / t_omp, "unit": "particles/s", "cores": CO.threads(), "kind": "port",
           "sample": f"{k} disc + {k} halo particles of the workload, C++ FP64 OpenMP "
                     f"(accumulate + return_forces_cart), {t_omp:.2f} s on {CO.threads()} threads",
           "host": host}
    progress(f"cpu openmp: {omp['value']:.3e} particles/s")
    # ---- reference decomposition over a process pool, extrapolated
    nc = pool_cores
    st = RP.time_reference_path(slo, eo, dxyzm, hxyzm, (c, s, hc), nc, n_eof=nc * 20000,
                                n_sl=nc * 600, n_field=nc * 200)
    full = (st["eof"] * nd / st["n_eof"] + st["sl"] * nh / st["n_sl"]
            + st["field"] * (nd + nh) / st["n_field"])
    ref = {"value": (nd + nh) / full, "unit": "particles/s", "cores": nc, "kind": "port",
           "extrapolated": True, "seconds_full_workload": full,
           "stage_seconds": {k2: st[k2] for k2 in ("eof", "sl", "field")},
           "sample": f"the reference's decomposition restated (oracle/ref_path.py): {nc} pool "
                     f"workers, contiguous blocks; EOF accumulate of {st['n_eof']} disc, SL "
                     f"accumulate of {st['n_sl']} halo particles (per-particle loop), field at "
                     f"{st['n_field']} points (per-point force_eval); each stage extrapolated "
                     f"linearly to {nd} + {nh} particles",
           "host": host}
    progress(f"cpu reference path: {ref['value']:.3e} particles/s (extrapolated, {nc} workers)")
    # ---- NumPy oracle, one thread
    num = cpu_baseline_numpy(slo, eo, dxyzm, hxyzm, budget_s)
    num["host"] = host
    progress(f"cpu numpy 1 core: {num['value']:.3e} particles/s")
    return ref, {"reference_path": ref, "openmp": omp, "numpy_1core": num}, (k, c, s, hc, forces)


def cpu_baseline_numpy(slo, eo, dxyzm, hxyzm, budget_s):
    """Time the NumPy oracle's accumulate + force evaluation on a sample sized
    to roughly budget_s seconds of one core; report particles/s."""
    from oracle import bfe_oracle as O

    def run(k):
        d = [v[:k] for v in dxyzm]
        h = [v[:k] for v in hxyzm]
        t0 = time.perf_counter()
        c, s = O.eof_accumulate(eo, *d)
        hc = O.sl_accumulate(slo, *h)
        F = O.FieldSpec(eo, c, s, slo, hc)
        for chunk in range(0, k, 4096):
            O.total_force(F, d[0][chunk:chunk + 4096], d[1][chunk:chunk + 4096], d[2][chunk:chunk + 4096])
            O.total_force(F, h[0][chunk:chunk + 4096], h[1][chunk:chunk + 4096], h[2][chunk:chunk + 4096])
        return time.perf_counter() - t0

    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        return _cpu_sized(run, budget_s)


def _cpu_sized(run, budget_s):
    k = 2000
    t = run(k)
    k2 = int(min(max(k * budget_s / max(t, 1e-3), k), 200000))
    t2 = run(k2)
    return {"value": 2 * k2 / t2, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"{k2} disc + {k2} halo particles of the same workload, FP64 NumPy oracle "
                      f"(accumulate + return_forces_cart), {t2:.1f} s, single thread"}


def accuracy(sysm, dxyzm, hxyzm, ref):
    """The bench step's call (bfe_forces_disc_halo, cell-ordered kernels) on
    the OpenMP baseline's sample, against those FP64 results: coefficient
    error per block (max |d a| / max |a| over each m or l block), force error
    per particle (|dF| / |F|, max and RMS) and potential error."""
    import torch
    k, c, s, hc, forces = ref
    T = lambda v: torch.as_tensor(np.asarray(v[:k], np.float32), device=sysm.device)
    omin = sysm.get_tuning("order_min")
    sysm.set_tuning("order_min", 1)
    try:
        out = sysm.forces_disc_halo([T(v) for v in dxyzm], [T(v) for v in hxyzm])
        torch.cuda.synchronize()
    finally:
        sysm.set_tuning("order_min", omin)
    g = {q: v.double().cpu().numpy() for q, v in out.items()}

    def blocks(got, ref_, rows):
        return max(float(np.max(np.abs(got[r] - ref_[r])) / np.max(np.abs(ref_[r])))
                   for r in rows if np.max(np.abs(ref_[r])) > 0)

    L = int(round(np.sqrt(hc.shape[0]))) - 1
    fr = np.stack(forces[:3])
    fg = np.stack([g["fx"], g["fy"], g["fz"]])
    e = np.linalg.norm(fg - fr, axis=0) / np.linalg.norm(fr, axis=0)
    eof_rows = ([slice(m, m + 1) for m in range(c.shape[0])], [slice(m, m + 1) for m in range(1, s.shape[0])])
    sl_rows = [slice(l * l, (l + 1) ** 2) for l in range(L + 1)]
    bc, bs = (v.double().cpu().numpy() for v in sysm.eof_accumulate(*[T(v) for v in dxyzm], bf16=True))
    bh = sysm.sl_accumulate(*[T(v) for v in hxyzm], bf16=True).double().cpu().numpy()
    return {"sample": f"{k} disc + {k} halo particles of the workload; FP64 reference: the "
                      "C++ OpenMP oracle (matches the NumPy oracle to 1e-12)",
            "coef_eof_block_rel_max": max(blocks(g["cos"], c, eof_rows[0]), blocks(g["sin"], s, eof_rows[1])),
            "coef_sl_block_rel_max": blocks(g["coef"], hc, sl_rows),
            "coef_eof_block_rel_max_bf16_mfma": max(blocks(bc, c, eof_rows[0]), blocks(bs, s, eof_rows[1])),
            "coef_sl_block_rel_max_bf16_mfma": blocks(bh, hc, sl_rows),
            "force_rel_max": float(np.max(e)), "force_rel_rms": float(np.sqrt(np.mean(e * e))),
            "pot_rel_max": float(np.max(np.abs(g["pot"] - forces[3]) / np.abs(forces[3])))}


if __name__ == "__main__":
    main()
