"""Launch-shape autotuner for the accumulate kernels (no counterpart in exptool).

The accumulate kernels are persistent: a launch runs `eof_acc_blocks` /
`sl_acc_blocks` blocks, each owning a run of consecutive particle tiles, and
a batch of snapshots is cut into segments of `eof_seg_tiles` / `sl_seg_tiles`
tiles.  The best values depend on the particle count, the table shapes and
the device.
The cell-ordered kernels also come in several lane shapes (`eof_dep_shape`,
`sl_dep_shape`: how many coefficient slots a lane owns and how many radial
orders each slot spans).
`tune_accumulate` times each candidate on the caller's own particles with
HIP events on the launch stream and keeps the fastest; `autotune` does this
for both kernels (single and batched form), applies the choices to the
BFESystem and caches them in a JSON file keyed by device, table shapes and
particle-count bucket.

Results depend on the launch shape only through the float32 partial sums
inside a segment (the float64 reductions are in a fixed order for a given
shape), so every choice meets the same accuracy bound.
"""
from __future__ import annotations

import json
import math
import os

import torch

BLOCK_CANDIDATES = (256, 512, 1024, 2048, 4096, 8192)
SEG_CANDIDATES = (0, 1, 2, 4, 8, 16)
# lane shapes (slots per lane, radial orders per slot) of the cell-ordered
# accumulate kernels, by tuning index (eof_dep_shape / sl_dep_shape)
EOF_SHAPES = ((1, 4), (1, 8), (1, 16), (2, 2), (2, 4), (2, 8), (2, 16), (4, 8))
SL_SHAPES = ((1, 8), (1, 16), (1, 24), (2, 8), (2, 16), (2, 24), (4, 16))


def fitting_shapes(system, kind):
    """Indices of the lane shapes that cover the coefficient rows of the
    tables: rows * ceil(norder / orders per slot) <= 64 * slots per lane."""
    if kind == "eof":
        rows, n, shapes = 2 * system.eof["mmax"] + 1, system.eof["norder"], EOF_SHAPES
    else:
        rows, n, shapes = (system.sl["lmax"] + 1) ** 2, system.sl["nmax"], SL_SHAPES
    return [k for k, (ppt, ch) in enumerate(shapes) if rows * -(-n // ch) <= 64 * ppt]


def _cache_path():
    return os.environ.get("EXPTOOL_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "exptool_amd", "tuning.json"))


def _time(fn, reps):
    stream = torch.cuda.current_stream()
    fn()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / reps * 1e-3


def _key(system, kind, n, nsnap):
    name = torch.cuda.get_device_name(system.device)
    if kind == "eof":
        shape = "eof:m%d:n%d:%dx%d" % (system.eof["mmax"], system.eof["norder"], system.eof["numx"],
                                       system.eof["numy"])
    else:
        shape = "sl:l%d:n%d:r%d" % (system.sl["lmax"], system.sl["nmax"], system.sl["numr"])
    bucket = int(round(math.log2(max(n, 1))))
    return f"{name}|{shape}|n2^{bucket}|snap{nsnap}"


def tune_accumulate(system, x, y, z, m, kind="eof", counts=None, reps=3, candidates=None,
                    shapes=True):
    """Time the accumulate of (x, y, z, m) for every candidate launch shape and
    apply the fastest: first the block count (and, with counts, the segment
    length) with the current lane shape, then, with shapes, every fitting lane
    shape of the cell-ordered kernels at that block count.  counts:
    per-snapshot counts for the batched form.  Returns dict(knob=value, ...,
    seconds=..., timings={...})."""
    knob = "eof_acc_blocks" if kind == "eof" else "sl_acc_blocks"
    seg = "eof_seg_tiles" if kind == "eof" else "sl_seg_tiles"
    shape_knob = "eof_dep_shape" if kind == "eof" else "sl_dep_shape"
    if counts is None:
        call = (lambda: system.eof_accumulate(x, y, z, m)) if kind == "eof" else \
            (lambda: system.sl_accumulate(x, y, z, m))
    else:
        call = (lambda: system.eof_accumulate_batch(x, y, z, m, counts)) if kind == "eof" else \
            (lambda: system.sl_accumulate_batch(x, y, z, m, counts))
    saved = {k: system.get_tuning(k) for k in (knob, seg)}
    timings = {}
    seg_list = SEG_CANDIDATES if counts is not None else (saved[seg],)
    for b in (candidates or BLOCK_CANDIDATES):
        for sgt in seg_list:
            system.set_tuning(knob, b)
            system.set_tuning(seg, sgt)
            timings[(b, sgt)] = _time(call, reps)
    (b, sgt), t = min(timings.items(), key=lambda kv: kv[1])
    system.set_tuning(knob, b)
    system.set_tuning(seg, sgt if counts is not None else saved[seg])
    out = {knob: b, "seconds": t, "timings": {f"{k[0]}/{k[1]}": v for k, v in timings.items()}}
    if counts is not None:
        out[seg] = sgt
    if shapes:
        auto = system.get_tuning(shape_knob)
        st = {auto: t}
        for k in fitting_shapes(system, kind):
            if k == auto:
                continue
            system.set_tuning(shape_knob, k)
            st[k] = _time(call, reps)
        best = min(st, key=st.get)
        system.set_tuning(shape_knob, best)
        out[shape_knob] = best
        out["seconds"] = st[best]
        out["timings"].update({f"shape{k}": v for k, v in st.items()})
    return out


def autotune(system, disc=None, halo=None, disc_counts=None, halo_counts=None, reps=3,
             cache=True):
    """Tune the EOF (disc) and SL (halo) accumulate launch shapes on the given
    particles ((x, y, z, m) device tensors; counts for batches) and apply them.
    Cached choices are applied without timing.  Returns the chosen knobs."""
    path = _cache_path()
    store = {}
    if cache and os.path.exists(path):
        try:
            with open(path) as f:
                store = json.load(f)
        except (OSError, ValueError):
            store = {}
    chosen = {}
    for kind, parts, counts in (("eof", disc, disc_counts), ("sl", halo, halo_counts)):
        if parts is None or (kind == "eof" and system.eof is None) or (kind == "sl" and system.sl is None):
            continue
        n = parts[0].numel()
        key = _key(system, kind, n, 1 if counts is None else len(counts))
        if key in store:
            rec = store[key]
            for k, v in rec.items():
                if k in ("eof_acc_blocks", "sl_acc_blocks", "eof_seg_tiles", "sl_seg_tiles",
                         "eof_dep_shape", "sl_dep_shape"):
                    system.set_tuning(k, v)
        else:
            rec = tune_accumulate(system, *parts, kind=kind, counts=counts, reps=reps)
            store[key] = {k: v for k, v in rec.items() if k != "timings"}
        chosen[key] = {k: v for k, v in rec.items() if k != "timings"}
    if cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(store, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return chosen
