"""Batched snapshots (bfe_*_batch): many snapshots' coefficients in one
accumulate launch and many frozen fields evaluated in one launch.

Each snapshot's coefficients are checked against the FP64 oracle (<= 1e-6
per block, as the single-snapshot path) and against the single-snapshot call;
the batched field evaluation must equal bfe_field_set + bfe_field_eval_cart
per snapshot bit for bit (same folded tables, same per-point code).
"""
import numpy as np
import pytest

from oracle import bfe_oracle as O
from tests.test_gpu_parity import block_rel_eof, block_rel_sl, disc_halo, f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sysm():
    from exptool_amd.basis import tablegen as tg
    from exptool_amd.bfe import BFESystem
    from exptool_amd.utils import samplers as S
    slt = tg.make_sl_tables(6, 20, 2000, 1e-3, 200.0, 1, 5.0, model_mass=10.0)
    d = S.exponential_disc(30000, seed=3)
    et = tg.make_eof_tables(d, mmax=6, norder=18, numx=256, numy=128, lmax_s=16, nmax_s=12)
    slo = O.SLTable(slt["lmax"], slt["nmax"], slt["numr"], slt["cmap"], slt["rmin"], slt["rmax"],
                    slt["scale"], slt["ev"], slt["ef"], slt["xi"], slt["p0"], slt["d0"])
    eo = O.EOFTable(et["mmax"], et["norder"], et["numx"], et["numy"], et["cmap"], et["rmin"],
                    et["rmax"], et["ascale"], et["hscale"], et["potC"], et["rforceC"],
                    et["zforceC"], et["potS"], et["rforceS"], et["zforceS"])
    return BFESystem(sl=slt, eof=et), slo, eo


COUNTS = [3000, 0, 1500, 40000, 257, 9000]
COUNTS_BIG = [50000, 30000, 0, 20000]        # >= 65536 in total: cell-ordered accumulate
# the segmented sort's tile edges (2048 items): whole, one over, one under,
# a single item and an empty snapshot, >= 65536 in total
COUNTS_EDGE = [2048, 1, 0, 2049, 4095, 60000]


def _snapshots(counts=COUNTS):
    parts_d, parts_h = [], []
    for k, n in enumerate(counts):
        d, h = disc_halo(max(n, 1), 100 + k)
        parts_d.append([v[:n] for v in d])
        parts_h.append([v[:n] for v in h])
    cat = lambda P: [np.concatenate([p[j] for p in P]) for j in range(4)]
    return parts_d, parts_h, f32(*cat(parts_d)), f32(*cat(parts_h))


@pytest.mark.parametrize("counts", [COUNTS, COUNTS_BIG, COUNTS_EDGE])
def test_batch_accumulate_matches_oracle_and_single(sysm, counts):
    s, slo, eo = sysm
    pd, ph, D, H = _snapshots(counts)
    ref = {}
    for k, n in enumerate(counts):
        if n:
            d, hh = f32(*pd[k]), f32(*ph[k])
            ref[k] = (d, O.eof_accumulate(eo, *d), O.sl_accumulate(slo, *hh))
    for tiles in (0, 1, 7):
        s.set_tuning("eof_seg_tiles", tiles)
        s.set_tuning("sl_seg_tiles", tiles)
        c, sn = (v.cpu().numpy() for v in s.eof_accumulate_batch(*D, counts))
        h = s.sl_accumulate_batch(*H, counts).cpu().numpy()
        assert c.shape == (len(counts), 7, 18) and h.shape == (len(counts), 49, 20)
        for k, n in enumerate(counts):
            if n == 0:
                assert np.all(c[k] == 0) and np.all(sn[k] == 0) and np.all(h[k] == 0)
                continue
            d, (rc, rs), rh = ref[k]
            assert block_rel_eof(c[k], rc) <= 1e-6, (tiles, k)
            assert block_rel_eof(sn[k][1:], rs[1:]) <= 1e-6, (tiles, k)
            assert block_rel_sl(h[k], rh, 6) <= 1e-6, (tiles, k)
            c1, s1 = (v.cpu().numpy() for v in s.eof_accumulate(*d))
            assert block_rel_eof(c[k], c1) <= 1e-6
    s.set_tuning("eof_seg_tiles", 0)
    s.set_tuning("sl_seg_tiles", 0)


@pytest.mark.parametrize("counts", [COUNTS, COUNTS_EDGE])
def test_batch_field_equals_per_snapshot_field(sysm, counts):
    s, _, _ = sysm
    pd, ph, D, H = _snapshots(counts)
    c, sn = s.eof_accumulate_batch(*D, counts)
    h = s.sl_accumulate_batch(*H, counts)
    # evaluate each snapshot's field at its disc particles
    x, y, z = D[0], D[1], D[2]
    for no_odd in (False, True):
        s.set_field_batch(c, sn, h, no_odd=no_odd)
        got = [v.cpu().numpy() for v in s.field_forces_batch(x, y, z, counts, rotpos=0.3)]
        off = np.concatenate([[0], np.cumsum(counts)])
        for k, n in enumerate(counts):
            if n == 0:
                continue
            s.set_field(c[k], sn[k], h[k], no_odd=no_odd)
            ref = [v.cpu().numpy() for v in s.field_forces(x[off[k]:off[k + 1]], y[off[k]:off[k + 1]],
                                                            z[off[k]:off[k + 1]], rotpos=0.3)]
            for j in range(4):
                assert np.array_equal(got[j][off[k]:off[k + 1]], ref[j]), (no_odd, k, j)


# more than 2 048 sort tiles (2 048 items each): the segmented tile prefix
# takes two rounds, the second starting inside the third snapshot
COUNTS_ROUNDS = [2_600_000, 0, 1_700_001, 7]


def test_batch_two_round_segmented_sort(sysm):
    """4.3M points in one batch: coefficients as the single-snapshot calls
    within the float32 mode's 1e-6 per block (same kernels; other segment
    plans regroup the float32 run sums, 2e-8 seen), and every point
    evaluated in its own snapshot's field bit for bit."""
    s, _, _ = sysm
    counts = COUNTS_ROUNDS
    pd, ph, D, H = _snapshots(counts)
    c, sn = (v.cpu().numpy() for v in s.eof_accumulate_batch(*D, counts))
    h = s.sl_accumulate_batch(*H, counts).cpu().numpy()
    for k, n in enumerate(counts):
        if n == 0:
            assert np.all(c[k] == 0) and np.all(sn[k] == 0) and np.all(h[k] == 0)
            continue
        c1, s1 = (v.cpu().numpy() for v in s.eof_accumulate(*f32(*pd[k])))
        h1 = s.sl_accumulate(*f32(*ph[k])).cpu().numpy()
        assert block_rel_eof(c[k], c1) <= 1e-6, k
        assert block_rel_eof(sn[k][1:], s1[1:]) <= 1e-6, k
        assert block_rel_sl(h[k], h1, 6) <= 1e-6, k
    s.set_field_batch(c, sn, h)
    x, y, z = D[0], D[1], D[2]
    got = [v.cpu().numpy() for v in s.field_forces_batch(x, y, z, counts)]
    off = np.concatenate([[0], np.cumsum(counts)])
    for k, n in enumerate(counts):
        if n == 0:
            continue
        s.set_field(c[k], sn[k], h[k])
        ref = [v.cpu().numpy() for v in s.field_forces(x[off[k]:off[k + 1]], y[off[k]:off[k + 1]],
                                                        z[off[k]:off[k + 1]])]
        for j in range(4):
            assert np.array_equal(got[j][off[k]:off[k + 1]], ref[j]), (k, j)


def test_batch_sampler_and_errors(sysm):
    s, _, _ = sysm
    from exptool_amd._lib import BFEError
    from exptool_amd.utils import workloads as W
    b = W.batch_snapshots(nsnap=4, lo=2000, hi=20000, seed=5)
    assert b["disc"][0].numel() == b["disc_counts"].sum()
    assert b["halo"][0].numel() == b["halo_counts"].sum()
    off = np.concatenate([[0], np.cumsum(b["disc_counts"])])
    m = b["disc"][3].double().cpu().numpy()
    for k in range(4):
        assert m[off[k]:off[k + 1]].sum() == pytest.approx(1.0, rel=1e-5)
    R = np.hypot(b["disc"][0].cpu().numpy(), b["disc"][1].cpu().numpy())
    assert R.max() <= 10.0 + 1e-4
    b2 = W.batch_snapshots(nsnap=4, lo=2000, hi=20000, seed=5)
    assert np.array_equal(b["disc"][0].cpu().numpy(), b2["disc"][0].cpu().numpy())
    D = [v[:100] for v in b["disc"]]
    with pytest.raises(ValueError):
        s.eof_accumulate_batch(*D, [50, 40])          # does not sum to n
    with pytest.raises(BFEError):
        s.set_tuning("eof_acc_blocks", 0)
    with pytest.raises(BFEError):
        s.set_tuning("no_such_knob", 1)


def test_autotune_applies_choice_within_tolerance(sysm, tmp_path, monkeypatch):
    s, slo, eo = sysm
    from exptool_amd import autotune as A
    monkeypatch.setenv("EXPTOOL_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    pd, ph, D, H = _snapshots()
    base = s.eof_accumulate(*D)[0].cpu().numpy()
    got = A.autotune(s, disc=[__import__("torch").tensor(v, dtype=__import__("torch").float32,
                                                        device="cuda") for v in D], reps=1)
    (key, rec), = got.items()
    assert rec["eof_acc_blocks"] in A.BLOCK_CANDIDATES and s.get_tuning("eof_acc_blocks") == rec["eof_acc_blocks"]
    assert (tmp_path / "tune.json").exists()
    after = s.eof_accumulate(*D)[0].cpu().numpy()
    assert block_rel_eof(after, base) <= 1e-6
    s.set_tuning("eof_acc_blocks", 2048)
    again = A.autotune(s, disc=[__import__("torch").tensor(v, dtype=__import__("torch").float32,
                                                          device="cuda") for v in D], reps=1)
    assert again[key]["eof_acc_blocks"] == rec["eof_acc_blocks"]      # from the cache
    assert s.get_tuning("eof_acc_blocks") == rec["eof_acc_blocks"]
    s.set_tuning("eof_acc_blocks", 2048)


def test_batch_field_invalidated_by_new_tables():
    """Table uploads drop the folded batch fields: evaluating them afterwards
    is an error (no launch), not a read of fields folded for other tables."""
    from exptool_amd._lib import BFEError
    from exptool_amd.basis import tablegen as tg
    from exptool_amd.bfe import BFESystem
    from exptool_amd.utils import samplers as S
    slt = tg.make_sl_tables(2, 4, 200, 1e-3, 50.0, 1, 1.0)
    d = S.exponential_disc(5000, seed=1)
    et = tg.make_eof_tables(d, mmax=2, norder=4, numx=16, numy=8, lmax_s=6, nmax_s=4)
    s = BFESystem(sl=slt, eof=et)
    counts = [10, 20]
    c = np.zeros((2, 3, 4)); h = np.zeros((2, 9, 4))
    s.set_field_batch(c, c, h)
    x = np.ones(30)
    s.field_forces_batch(x, x, x, counts)
    big = tg.make_sl_tables(3, 6, 400, 1e-3, 50.0, 1, 1.0)
    s.set_sl(big)
    with pytest.raises(BFEError, match="no batch field"):
        s.field_forces_batch(x, x, x, counts)
    s.set_sl(slt)
    s.set_field_batch(c, c, h)
    s.set_eof(et)
    with pytest.raises(BFEError, match="no batch field"):
        s.field_forces_batch(x, x, x, counts)


def test_batch_barred_snapshots_against_oracle():
    """The benchmark's batch workload at reduced snapshot count: barred discs
    (m = 2, A ~ U[0, 0.3]) of 200k-500k particles and their halos, sampled on
    the device, on the benchmark's bases.  Every snapshot's coefficients
    against the FP64 oracle (<= 1e-6 per block) and each snapshot's field at
    a subsample of its own particles against the oracle's Cartesian field."""
    import torch
    from exptool_amd.bfe import BFESystem
    from exptool_amd.utils import workloads as W
    st, et = W.disc_halo_tables(seed=1)
    s = BFESystem(sl=st, eof=et)
    slo = O.SLTable(st["lmax"], st["nmax"], st["numr"], st["cmap"], st["rmin"], st["rmax"],
                    st["scale"], st["ev"], st["ef"], st["xi"], st["p0"], st["d0"])
    eo = O.EOFTable(et["mmax"], et["norder"], et["numx"], et["numy"], et["cmap"], et["rmin"],
                    et["rmax"], et["ascale"], et["hscale"], et["potC"], et["rforceC"],
                    et["zforceC"], et["potS"], et["rforceS"], et["zforceS"])
    b = W.batch_snapshots(nsnap=3, lo=200_000, hi=500_000, seed=4)
    dc, hc = b["disc_counts"], b["halo_counts"]
    print(f"barred batch: disc counts {list(dc)}, bar amplitudes {np.round(b['amp'], 3)}", flush=True)
    assert dc.min() >= 200_000 and b["amp"].max() > 0.1
    c, sn = s.eof_accumulate_batch(*b["disc"], dc)
    h = s.sl_accumulate_batch(*b["halo"], hc)
    s.set_field_batch(c, sn, h)
    fd = [v.cpu().numpy().astype(np.float64) for v in s.field_forces_batch(*b["disc"][:3], dc)]
    torch.cuda.synchronize()
    c, sn, h = (v.cpu().numpy() for v in (c, sn, h))
    D = [v.cpu().numpy().astype(np.float64) for v in b["disc"]]
    H = [v.cpu().numpy().astype(np.float64) for v in b["halo"]]
    od, oh = np.concatenate([[0], np.cumsum(dc)]), np.concatenate([[0], np.cumsum(hc)])
    rng = np.random.default_rng(7)
    for k in range(3):
        dk = [v[od[k]:od[k + 1]] for v in D]
        hk = [v[oh[k]:oh[k + 1]] for v in H]
        rc, rs = O.eof_accumulate(eo, *dk)
        assert block_rel_eof(c[k], rc) <= 1e-6, k
        assert block_rel_eof(sn[k][1:], rs[1:]) <= 1e-6, k
        # the bar shows in the m = 2 block in proportion to its amplitude
        assert np.max(np.abs(rc[2])) > 0.1 * b["amp"][k] * np.max(np.abs(rc[0]))
        rh = O.sl_accumulate(slo, *hk)
        assert block_rel_sl(h[k], rh, slo.lmax) <= 1e-6, k
        sub = rng.choice(dc[k], 3000, replace=False)
        ref = O.total_force(O.FieldSpec(eo, rc, rs, slo, rh), *(v[sub] for v in dk[:3]))
        Fg = np.stack([fd[j][od[k]:od[k + 1]][sub] for j in range(3)])
        Fr = np.stack(ref[:3])
        e = np.linalg.norm(Fg - Fr, axis=0) / np.linalg.norm(Fr, axis=0)
        assert np.max(e) <= 1e-5, (k, np.max(e))
        assert np.max(np.abs(fd[3][od[k]:od[k + 1]][sub] - ref[3]) / np.abs(ref[3])) <= 1e-5


def test_eof_bf16_mfma_mode(sysm):
    """The EOF accumulate's bf16 MFMA mode (corner sums of each cell, node
    sums, then per row the snapshots x nodes x radial-orders product with the
    node sums and the table as bf16 hi + lo pieces): one snapshot (sorted and
    small) and batches of 6 and 20 snapshots (two 16-snapshot tiles, an
    empty snapshot) within 1e-4 per m block of the FP64 oracle."""
    s, slo, eo = sysm
    for n in (90000, 3000):
        (dx, dy, dz, dm), _ = disc_halo(n, 71)
        d = f32(dx, dy, dz, dm)
        c, sn = (v.cpu().numpy() for v in s.eof_accumulate(*d, bf16=True))
        rc, rs = O.eof_accumulate(eo, *d)
        e = max(block_rel_eof(c, rc), block_rel_eof(sn[1:], rs[1:]))
        print(f"bf16 MFMA mode, {n} particles: block error {e:.2e}", flush=True)
        assert e <= 1e-4 and np.all(sn[0] == 0)
        c2, _ = (v.cpu().numpy() for v in s.eof_accumulate(*d, bf16=True))
        assert np.max(np.abs(c2 - c)) <= 1e-12 * np.max(np.abs(c))
    for counts in (COUNTS, [700 + 37 * k for k in range(19)] + [0]):
        pd, ph, D, H = _snapshots(counts)
        c, sn = (v.cpu().numpy() for v in s.eof_accumulate_batch(*D, counts, bf16=True))
        worst = 0.0
        for k, n in enumerate(counts):
            if n == 0:
                assert np.all(c[k] == 0) and np.all(sn[k] == 0)
                continue
            rc, rs = O.eof_accumulate(eo, *f32(*pd[k]))
            worst = max(worst, block_rel_eof(c[k], rc), block_rel_eof(sn[k][1:], rs[1:]))
        print(f"bf16 MFMA mode, {len(counts)} snapshots: worst block error {worst:.2e}", flush=True)
        assert worst <= 1e-4
