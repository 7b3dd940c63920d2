"""Parity of the HIP path (libexptool_bfe.so via exptool_amd.bfe) with the FP64
oracle, on the golden tables and on tables of the workload's shape.

Tolerances (float32 tables and state, float64 geometry and cross-block sums):
  coefficients  max|d a| / max|a| per l (SL) or m (EOF) block   <= 1e-6
  potential, forces, per particle |dF| / |F| (Cartesian field)   <= 1e-5
  batch evaluator outputs, max|d| / max|ref| per output          <= 1e-5
"""
import numpy as np
import pytest

from oracle import bfe_oracle as O
from tests.test_golden import eof_table, sl_table

pytestmark = pytest.mark.gpu

COEF_TOL = 1e-6
FORCE_TOL = 1e-5


def block_rel_sl(got, ref, lmax):
    worst = 0.0
    for l in range(lmax + 1):
        rows = slice(l * l, (l + 1) * (l + 1))
        den = np.max(np.abs(ref[rows]))
        if den > 0:
            worst = max(worst, np.max(np.abs(got[rows] - ref[rows])) / den)
    return worst


def block_rel_eof(got, ref):
    worst = 0.0
    for m in range(ref.shape[0]):
        den = np.max(np.abs(ref[m]))
        if den > 0:
            worst = max(worst, np.max(np.abs(got[m] - ref[m])) / den)
    return worst


def relmax(a, b):
    b = np.asarray(b, float)
    return np.max(np.abs(np.asarray(a, float) - b)) / max(np.max(np.abs(b)), 1e-300)


def f32(*a):
    return [np.asarray(v, dtype=np.float32).astype(np.float64) for v in a]


def sl_dict(t):
    return dict(lmax=t.lmax, nmax=t.nmax, numr=t.numr, cmap=t.cmap, scale=t.scale, rmin=t.rmin,
                rmax=t.rmax, xi=t.xi, ev=t.evtable, ef=t.eftable, p0=t.p0, d0=t.d0)


def eof_dict(t):
    return dict(mmax=t.mmax, norder=t.norder, numx=t.numx, numy=t.numy, cmap=t.cmap, rmin=t.rmin,
                rmax=t.rmax, ascale=t.ascale, hscale=t.hscale, potC=t.potC, rforceC=t.rforceC,
                zforceC=t.zforceC, potS=t.potS, rforceS=t.rforceS, zforceS=t.zforceS,
                densC=t.densC, densS=t.densS, dens=1 if t.densC is not None else 0)


@pytest.fixture(scope="module")
def small(golden):
    from exptool_amd.bfe import BFESystem
    slt, eoft = sl_table(golden), eof_table(golden)
    return BFESystem(sl=sl_dict(slt), eof=eof_dict(eoft)), slt, eoft


@pytest.fixture(scope="module")
def big():
    """Tables of the workload's shape: halo SL lmax=6 nmax=20 numr=2000 (cmap 1,
    scale 5); disc EOF mmax=6 norder=18 on a 256x128 grid."""
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
                    et["zforceC"], et["potS"], et["rforceS"], et["zforceS"], et["densC"],
                    et["densS"])
    return BFESystem(sl=slt, eof=et), slo, eo


def disc_halo(n, seed):
    from exptool_amd.utils import samplers as S
    d = S.exponential_disc(n, seed=seed)
    h = S.hernquist_sphere(n, a=5.0, rtrunc=100.0, mass=10.0, seed=seed + 1)
    return d, h


# ----------------------------------------------------------------------------
def test_sl_accumulate_golden_particles(small, golden):
    s, slt, _ = small
    x, y, z, m = f32(*(golden[k] for k in ("px", "py", "pz", "pm")))
    for no_odd in (False, True):
        got = s.sl_accumulate(x, y, z, m, no_odd=no_odd).cpu().numpy()
        ref = O.sl_accumulate(slt, x, y, z, m, no_odd=no_odd)
        assert block_rel_sl(got, ref, slt.lmax) <= COEF_TOL


def test_eof_accumulate_golden_particles(small, golden):
    s, _, eoft = small
    x, y, z, m = f32(*(golden[k] for k in ("px", "py", "pz", "pm")))
    c, sn = (v.cpu().numpy() for v in s.eof_accumulate(x, y, z, m))
    rc, rs = O.eof_accumulate(eoft, x, y, z, m)
    assert block_rel_eof(c, rc) <= COEF_TOL
    assert block_rel_eof(sn, rs) <= COEF_TOL
    assert np.all(sn[0] == 0)


def test_accumulate_workload_tables(big):
    s, slo, eo = big
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(40000, 5)
    dx, dy, dz, dm = f32(dx, dy, dz, dm)
    hx, hy, hz, hm = f32(hx, hy, hz, hm)
    c, sn = (v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, dm))
    rc, rs = O.eof_accumulate(eo, dx, dy, dz, dm)
    assert block_rel_eof(c, rc) <= COEF_TOL
    assert block_rel_eof(sn[1:], rs[1:]) <= COEF_TOL
    got = s.sl_accumulate(hx, hy, hz, hm).cpu().numpy()
    ref = O.sl_accumulate(slo, hx, hy, hz, hm)
    assert block_rel_sl(got, ref, slo.lmax) <= COEF_TOL


@pytest.mark.parametrize("compat", [False, True])
def test_sl_eval_golden(small, golden, compat):
    s, slt, _ = small
    x, y, z = f32(*(golden[k] for k in ("px", "py", "pz")))
    coef = golden["sl_coef"]
    got = s.sl_eval(x, y, z, coef, compat=compat)
    with np.errstate(all="ignore"):
        ref = O.sl_all_eval_particles(slt, x, y, z, coef, compat=compat)
    for k, r in zip(("den0", "den1", "pot0", "pot1", "potr", "pott", "potp"), ref):
        assert relmax(got[k].cpu().numpy(), r) <= FORCE_TOL, k
    got = s.sl_eval(x, y, z, coef, L1=1, L2=2, no_odd=True, compat=compat)
    with np.errstate(all="ignore"):
        ref = O.sl_all_eval_particles(slt, x, y, z, coef, L1=1, L2=2, no_odd=True, compat=compat)
    for k, r in zip(("den0", "den1", "pot0", "pot1", "potr", "pott", "potp"), ref):
        assert relmax(got[k].cpu().numpy(), r) <= FORCE_TOL, k


def test_sl_eval_matches_reference_vectors(small, golden):
    """compat evaluation against the reference's own all_eval_particles output
    (positions rounded to float32 on the device side are re-used by the oracle
    in the other tests; here the float64 reference vectors themselves)."""
    s, _, _ = small
    x, y, z = (golden[k] for k in ("px", "py", "pz"))
    got = s.sl_eval(x, y, z, golden["sl_coef"], compat=True)
    ref = golden["sl_alleval"]
    for j, k in enumerate(("den0", "den1", "pot0", "pot1", "potr", "pott", "potp")):
        assert relmax(got[k].cpu().numpy(), ref[j]) <= 1e-4, k


def test_eof_eval_golden(small, golden):
    s, _, eoft = small
    x, y, z = f32(*(golden[k] for k in ("px", "py", "pz")))
    c, sn = golden["eof_cos"], golden["eof_sin"]
    for m1, m2 in ((0, 1000), (1, 1), (0, 0)):
        got = s.eof_eval(x, y, z, c, sn, m1=m1, m2=m2, density=True)
        ref = O.eof_eval_particles(eoft, x, y, z, c, sn, m1=m1, m2=m2, density=True)
        for k, r in zip(("p0", "p", "d0", "d", "fr", "fp", "fz"), ref[:7]):
            if np.max(np.abs(r)) == 0:
                assert np.max(np.abs(got[k].cpu().numpy())) == 0
            else:
                assert relmax(got[k].cpu().numpy(), r) <= FORCE_TOL, (k, m1, m2)


def _field_rel(got4, ref4):
    fx, fy, fz, _ = got4
    rfx, rfy, rfz, _ = ref4
    num = np.sqrt((fx - rfx) ** 2 + (fy - rfy) ** 2 + (fz - rfz) ** 2)
    den = np.sqrt(rfx ** 2 + rfy ** 2 + rfz ** 2)
    return num / den


@pytest.mark.parametrize("compat", [False, True])
@pytest.mark.parametrize("no_odd", [False, True])
def test_field_cart_golden(small, golden, compat, no_odd):
    s, slt, eoft = small
    x, y, z = f32(*(golden[k] for k in ("px", "py", "pz")))
    s.set_field(golden["eof_cos"], golden["eof_sin"], golden["sl_coef"], no_odd=no_odd,
                compat=compat)
    F = O.FieldSpec(eoft, golden["eof_cos"], golden["eof_sin"], slt, golden["sl_coef"],
                    no_odd=no_odd)
    for rot in (0.0, 0.7):
        fx, fy, fz, pot, comp = s.field_forces(x, y, z, rotpos=rot, components=True)
        got = [v.cpu().numpy().astype(np.float64) for v in (fx, fy, fz, pot)]
        ref = O.total_force(F, x, y, z, rot, compat=compat)
        e = _field_rel(got, ref)
        # the r = 0 particle has F = 0 analytically; compare it absolutely
        ok = np.isfinite(e) & (np.sqrt(ref[0] ** 2 + ref[1] ** 2 + ref[2] ** 2) > 1e-6)
        assert np.max(e[ok]) <= FORCE_TOL
        assert relmax(got[3], ref[3]) <= FORCE_TOL
        refc = O.return_forces_cart(F, x, y, z, rot, compat=compat)
        compn = comp.cpu().numpy()
        for j in range(8):
            assert relmax(compn[j], refc[j]) <= FORCE_TOL, j


def test_field_cart_matches_reference_vectors(small, golden):
    s, _, _ = small
    s.set_field(golden["eof_cos"], golden["eof_sin"], golden["sl_coef"], compat=True)
    x, y, z = (golden[k][:120] for k in ("px", "py", "pz"))
    rot = golden["fields_rotpos"]
    ref = golden["fields_cart"]
    for r0 in np.unique(rot):
        sel = rot == r0
        _, _, _, _, comp = s.field_forces(x[sel], y[sel], z[sel], rotpos=float(r0), components=True)
        c = comp.cpu().numpy()
        for j in range(8):
            assert relmax(c[j], ref[sel, j]) <= 1e-4, j


def test_field_workload(big):
    s, slo, eo = big
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(20000, 7)
    c, sn = s.eof_accumulate(*f32(dx, dy, dz, dm))
    hc = s.sl_accumulate(*f32(hx, hy, hz, hm))
    s.set_field(c, sn, hc)
    x = np.concatenate([dx[:3000], hx[:3000]])
    y = np.concatenate([dy[:3000], hy[:3000]])
    z = np.concatenate([dz[:3000], hz[:3000]])
    x, y, z = f32(x, y, z)
    got = [v.cpu().numpy().astype(np.float64) for v in s.field_forces(x, y, z)]
    F = O.FieldSpec(eo, c.cpu().numpy(), sn.cpu().numpy(), slo, hc.cpu().numpy())
    ref = O.total_force(F, x, y, z)
    e = _field_rel(got, ref)
    assert np.max(e) <= FORCE_TOL, np.max(e)
    assert relmax(got[3], ref[3]) <= FORCE_TOL


def test_leapfrog_matches_oracle(small, golden):
    s, slt, eoft = small
    import torch
    s.set_field(golden["eof_cos"], golden["eof_sin"], golden["sl_coef"])
    F = O.FieldSpec(eoft, golden["eof_cos"], golden["eof_sin"], slt, golden["sl_coef"])
    init = golden["leap_init"].astype(np.float32).astype(np.float64)
    st = [torch.tensor(init[:, k], dtype=torch.float32, device="cuda") for k in range(6)]
    out = s.leapfrog(st, dt=0.01, nsteps=199, rotfreq=0.05, energy_every=0, record_every=1)
    traj = out["traj"].cpu().numpy()        # (200, 7, 3)
    ref = O.leapfrog(F, 200, 0.01, init[:, :3], init[:, 3:], rotfreq=0.05)
    for j, k in enumerate(("X", "Y", "Z", "VX", "VY", "VZ", "P")):
        assert relmax(traj[:, j, :].T, ref[k]) <= 1e-4, k


def test_leapfrog_graph_equals_stream(small, golden):
    s, _, _ = small
    import torch
    s.set_field(golden["eof_cos"], golden["eof_sin"], golden["sl_coef"])
    rng = np.random.default_rng(0)
    pos = rng.normal(0, 2, (4096, 3)).astype(np.float32)
    vel = rng.normal(0, 0.3, (4096, 3)).astype(np.float32)
    res = []
    for graph in (False, True):
        st = [torch.tensor(np.ascontiguousarray(a[:, k]), device="cuda")
              for a in (pos, vel) for k in range(3)]
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            out = s.leapfrog(st, dt=0.005, nsteps=300, energy_every=100, use_graph=graph,
                             stream=stream)
        stream.synchronize()
        res.append((torch.stack(st).cpu().numpy(), out["drift"].cpu().numpy(),
                    out["energy"].cpu().numpy()))
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_determinism_and_empty(big):
    s, _, _ = big
    (dx, dy, dz, dm), _ = disc_halo(50000, 9)
    a = [v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, dm)]
    b = [v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, dm)]
    assert all(np.array_equal(u, v) for u, v in zip(a, b))
    e = np.zeros(0)
    c, sn = s.eof_accumulate(e, e, e, e)
    assert np.all(c.cpu().numpy() == 0) and np.all(sn.cpu().numpy() == 0)
    assert np.all(s.sl_accumulate(e, e, e, e).cpu().numpy() == 0)


def test_invariants_full_size(big):
    """Size-independent properties at 1M particles: z-rotation by alpha
    rotates each (cos, sin) pair by m alpha; mass linearity; split sums."""
    s, _, _ = big
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(1_000_000, 21)
    c, sn = (v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, dm))
    al = 0.37
    rx = np.cos(al) * dx - np.sin(al) * dy
    ry = np.sin(al) * dx + np.cos(al) * dy
    c2, s2 = (v.cpu().numpy() for v in s.eof_accumulate(rx, ry, dz, dm))
    m = np.arange(c.shape[0])[:, None]
    # rotating particles by al: a'_c = a_c cos(m al) - a_s sin(m al); a'_s = a_s cos + a_c sin
    pc = c * np.cos(m * al) - sn * np.sin(m * al)
    ps = sn * np.cos(m * al) + c * np.sin(m * al)
    scale = np.max(np.abs(c), axis=1, keepdims=True)
    assert np.max(np.abs(c2 - pc) / scale) < 1e-4
    assert np.max(np.abs(s2 - ps) / scale) < 1e-4
    c3, s3 = (v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, 2.0 * dm))
    assert block_rel_eof(c3, 2 * c) < 1e-6
    h = s.sl_accumulate(hx, hy, hz, hm).cpu().numpy()
    h1 = s.sl_accumulate(hx[:400000], hy[:400000], hz[:400000], hm[:400000]).cpu().numpy()
    h2 = s.sl_accumulate(hx[400000:], hy[400000:], hz[400000:], hm[400000:]).cpu().numpy()
    assert block_rel_sl(h1 + h2, h, 6) < 1e-6


def test_errors_raise(small):
    from exptool_amd._lib import BFEError
    from exptool_amd.bfe import BFESystem
    s = BFESystem()
    with pytest.raises(BFEError):
        s.sl_accumulate(np.zeros(3), np.zeros(3), np.zeros(3), np.ones(3))
    with pytest.raises(BFEError):
        s.ctx.check(s.ctx.lib.bfe_field_eval_cart(s.ctx.h, None, None, None, 1, 0.0, None, None,
                                                  None, None, None, None), "field")


def test_cell_ordered_evaluation_is_per_point(big):
    """Batches of >= 65536 points are evaluated in table-cell order; every
    point's result must equal the unordered evaluation bit for bit (chunks
    below the threshold run in input order)."""
    s, _, _ = big
    import torch
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(70000, 31)
    c, sn = s.eof_accumulate(dx, dy, dz, dm)
    s.set_field(c, sn, s.sl_accumulate(hx, hy, hz, hm))
    x, y, z = (np.concatenate(p) for p in ((dx, hx), (dy, hy), (dz, hz)))
    whole = s.field_forces(x, y, z, rotpos=0.2, components=True)
    parts = [s.field_forces(x[k:k + 50000], y[k:k + 50000], z[k:k + 50000], rotpos=0.2,
                            components=True) for k in range(0, x.size, 50000)]
    for j in range(4):
        a = whole[j].cpu().numpy()
        b = np.concatenate([p[j].cpu().numpy() for p in parts])
        assert np.array_equal(a, b), j
    comp = np.concatenate([p[4].cpu().numpy() for p in parts], axis=1)
    assert np.array_equal(whole[4].cpu().numpy(), comp)
    torch.cuda.synchronize()


def test_leapfrog_independent_of_orbit_order(big):
    """Orbits advance in cell order (rebuilt every chunk); the result of an
    orbit and the energy log must not depend on the input order."""
    s, _, _ = big
    import torch
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(100000, 41)
    c, sn = s.eof_accumulate(dx, dy, dz, dm)
    s.set_field(c, sn, s.sl_accumulate(hx, hy, hz, hm))
    rng = np.random.default_rng(3)
    n = 20000
    R = rng.uniform(0.5, 8.0, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    pos = np.stack([R * np.cos(ph), R * np.sin(ph), rng.normal(0, 0.1, n)])
    vel = np.stack([-np.sin(ph), np.cos(ph), np.zeros(n)]) * 1.2 + rng.normal(0, 0.1, (3, n))
    order = rng.permutation(n)
    out = []
    for perm in (np.arange(n), order):
        st = [torch.tensor(np.ascontiguousarray(a[perm]), dtype=torch.float32, device="cuda")
              for a in (*pos, *vel)]
        r = s.leapfrog(st, dt=2e-3, nsteps=250, energy_every=50)
        out.append((torch.stack(st).cpu().numpy(), r["drift"].cpu().numpy(), perm))
    (a, da, _), (b, db, p) = out
    inv = np.argsort(p)
    assert np.array_equal(a, b[:, inv])
    assert np.array_equal(da, db[inv])


def test_cell_ordered_accumulate_matches_oracle(big):
    """From order_min (65536) particles on, accumulate sorts the particles by
    table cell and regroups the sums per cell run; it must meet the same bound
    against the oracle as the input-order kernels, and be reproducible."""
    s, slo, eo = big
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(100000, 51)
    dx, dy, dz, dm = f32(dx, dy, dz, dm)
    hx, hy, hz, hm = f32(hx, hy, hz, hm)
    rc, rs = O.eof_accumulate(eo, dx, dy, dz, dm)
    ref_h = O.sl_accumulate(slo, hx, hy, hz, hm)
    out = {}
    for omin in (65536, 1 << 40):           # cell-ordered, then input order
        s.set_tuning("order_min", omin)
        c, sn = (v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, dm))
        h = s.sl_accumulate(hx, hy, hz, hm).cpu().numpy()
        assert block_rel_eof(c, rc) <= COEF_TOL, omin
        assert block_rel_eof(sn[1:], rs[1:]) <= COEF_TOL, omin
        assert block_rel_sl(h, ref_h, slo.lmax) <= COEF_TOL, omin
        out[omin] = (c, sn, h)
    s.set_tuning("order_min", 65536)
    c2, s2 = (v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, dm))
    h2 = s.sl_accumulate(hx, hy, hz, hm).cpu().numpy()
    c1, s1, h1 = out[65536]
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2) and np.array_equal(h1, h2)
    hn = s.sl_accumulate(hx, hy, hz, hm, no_odd=True).cpu().numpy()
    assert block_rel_sl(hn, O.sl_accumulate(slo, hx, hy, hz, hm, no_odd=True), slo.lmax) <= COEF_TOL


def test_accumulate_disc_halo_equals_separate_calls(big):
    """The combined disc + halo accumulate (concurrent halves, the halo half
    with its own work buffers) returns exactly what the two single calls
    return, for cell-ordered and input-order sizes and empty halves, when the
    halo half is cut into as many segments as the single call
    (sl_acc_blocks_beside = sl_acc_blocks).  At cell-ordered sizes this also
    holds the two instantiations of the SL deposit (two waves per SIMD
    beside the EOF accumulate, three alone) to the same bits.  With the
    default, fewer segments beside the EOF half, the float32 run sums are
    regrouped: within COEF_TOL of the single call."""
    s, slo, eo = big
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(90000, 57)
    D = f32(dx, dy, dz, dm)
    H = f32(hx, hy, hz, hm)
    beside = s.get_tuning("sl_acc_blocks_beside")
    s.set_tuning("sl_acc_blocks_beside", s.get_tuning("sl_acc_blocks"))
    try:
        for nd, nh in ((90000, 90000), (90000, 5000), (3000, 90000), (0, 7000), (7000, 0)):
            d = [v[:nd] for v in D]
            h = [v[:nh] for v in H]
            c, sn, hc = (v.cpu().numpy() for v in s.accumulate_disc_halo(d, h))
            c1, s1 = (v.cpu().numpy() for v in s.eof_accumulate(*d))
            h1 = s.sl_accumulate(*h).cpu().numpy()
            assert np.array_equal(c, c1) and np.array_equal(sn, s1) and np.array_equal(hc, h1), (nd, nh)
        c, sn, hc = (v.cpu().numpy() for v in s.accumulate_disc_halo(D, H, no_odd=True))
        assert np.array_equal(hc, s.sl_accumulate(*H, no_odd=True).cpu().numpy())
    finally:
        s.set_tuning("sl_acc_blocks_beside", beside)
    rc, rs = O.eof_accumulate(eo, *D)
    assert block_rel_eof(c, rc) <= COEF_TOL
    c, sn, hc = (v.cpu().numpy() for v in s.accumulate_disc_halo(D, H))
    assert np.array_equal(c, s.eof_accumulate(*D)[0].cpu().numpy())
    assert block_rel_sl(hc, s.sl_accumulate(*H).cpu().numpy(), 6) <= COEF_TOL


def test_forces_disc_halo_equals_separate_calls(big):
    """bfe_forces_disc_halo (coefficients, field and forces at the same
    particles) equals accumulate_disc_halo + set_field + field_forces bit for
    bit, for cell-ordered and input-order sizes, empty halves (also an empty
    disc beside a cell-ordered halo), field options and the eight
    components; the forces meet the oracle bound."""
    s, slo, eo = big
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(90000, 61)
    D = f32(dx, dy, dz, dm)
    H = f32(hx, hy, hz, hm)
    cases = ((90000, 90000, {}), (90000, 5000, dict(no_odd=True, compat=True, rotpos=0.4)),
             (3000, 90000, dict(halofac=0.5, disk_m=4, halo_l=3, halo_n=12)), (0, 7000, {}),
             (7000, 0, {}), (0, 90000, {}))
    for nd, nh, kw in cases:
        d = [v[:nd] for v in D]
        h = [v[:nh] for v in H]
        got = s.forces_disc_halo(d, h, components=True, **kw)
        rot = kw.pop("rotpos", 0.0)
        c, sn, hc = s.accumulate_disc_halo(d, h, no_odd=kw.get("no_odd", False))
        for a, b in ((got["cos"], c), (got["sin"], sn), (got["coef"], hc)):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy()), (nd, nh)
        s.set_field(c, sn, hc, **kw)
        ref = s.field_forces(np.concatenate([d[0], h[0]]), np.concatenate([d[1], h[1]]),
                             np.concatenate([d[2], h[2]]), rotpos=rot, components=True)
        for k, r in zip(("fx", "fy", "fz", "pot", "comp"), ref):
            assert np.array_equal(got[k].cpu().numpy(), r.cpu().numpy()), (nd, nh, k)
    got = s.forces_disc_halo(D, H)
    x, y, z = (np.concatenate([a, b]) for a, b in zip(D[:3], H[:3]))
    F = O.FieldSpec(eo, got["cos"].cpu().numpy(), got["sin"].cpu().numpy(), slo,
                    got["coef"].cpu().numpy())
    idx = np.random.default_rng(0).choice(x.size, 4000, replace=False)
    ref = O.total_force(F, x[idx], y[idx], z[idx])
    res = [got[k].cpu().numpy().astype(np.float64)[idx] for k in ("fx", "fy", "fz", "pot")]
    assert np.max(_field_rel(res, ref)) <= FORCE_TOL


def test_cell_ordered_accumulate_every_lane_shape(big):
    """Each lane shape of the cell-ordered kernels (slots per lane, radial
    orders per slot; shapes that do not fit the tables fall back to the
    automatic one) meets the oracle bound; an index outside the list raises."""
    from exptool_amd._lib import BFEError
    s, slo, eo = big
    (dx, dy, dz, dm), (hx, hy, hz, hm) = disc_halo(80000, 53)
    dx, dy, dz, dm = f32(dx, dy, dz, dm)
    hx, hy, hz, hm = f32(hx, hy, hz, hm)
    rc, rs = O.eof_accumulate(eo, dx, dy, dz, dm)
    ref_h = O.sl_accumulate(slo, hx, hy, hz, hm)
    for k in range(-1, 8):
        s.set_tuning("eof_dep_shape", k)
        c, sn = (v.cpu().numpy() for v in s.eof_accumulate(dx, dy, dz, dm))
        assert block_rel_eof(c, rc) <= COEF_TOL, k
        assert block_rel_eof(sn[1:], rs[1:]) <= COEF_TOL, k
    for k in range(-1, 7):
        s.set_tuning("sl_dep_shape", k)
        h = s.sl_accumulate(hx, hy, hz, hm).cpu().numpy()
        assert block_rel_sl(h, ref_h, slo.lmax) <= COEF_TOL, k
    s.set_tuning("eof_dep_shape", -1)
    s.set_tuning("sl_dep_shape", -1)
    assert s.get_tuning("eof_dep_shape") == -1
    with pytest.raises(BFEError):
        s.set_tuning("eof_dep_shape", 8)
    with pytest.raises(BFEError):
        s.set_tuning("sl_dep_shape", -2)


@pytest.mark.parametrize("n", [3000, 100000])
def test_sl_accumulate_mfma_valu_and_bf16_modes(big, n):
    """SL accumulate on the MFMA units (three-piece bf16 split, default for
    lmax <= 7), on the vector ALUs (sl_acc_path = 2, cell-ordered from 65536
    particles) and in the two-piece bf16 mode (BFE_ACC_BF16, <= 1e-4)."""
    s, slo, _ = big
    _, (hx, hy, hz, hm) = disc_halo(n, 61)
    hx, hy, hz, hm = f32(hx, hy, hz, hm)
    for no_odd in (False, True):
        ref = O.sl_accumulate(slo, hx, hy, hz, hm, no_odd=no_odd)
        for path in (1, 2):
            s.set_tuning("sl_acc_path", path)
            got = s.sl_accumulate(hx, hy, hz, hm, no_odd=no_odd).cpu().numpy()
            assert block_rel_sl(got, ref, slo.lmax) <= COEF_TOL, (path, no_odd)
        s.set_tuning("sl_acc_path", 0)
        b = s.sl_accumulate(hx, hy, hz, hm, no_odd=no_odd, bf16=True).cpu().numpy()
        e = block_rel_sl(b, ref, slo.lmax)
        assert e <= 1e-4, e
    a1 = s.sl_accumulate(hx, hy, hz, hm).cpu().numpy()
    a2 = s.sl_accumulate(hx, hy, hz, hm).cpu().numpy()
    assert np.array_equal(a1, a2)
