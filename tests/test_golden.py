"""The oracle and the table readers against golden vectors written by the
reference itself (tests/golden/make_golden.py).  These pin the restatement of
every quirk the oracle reproduces in compat mode."""
import numpy as np
import pytest

from oracle import bfe_oracle as O
from exptool_amd.basis import eof as E


def sl_table(g):
    return O.SLTable(int(g["sl_lmax"]), int(g["sl_nmax"]), int(g["sl_numr"]), int(g["sl_cmap"]),
                     float(g["sl_rmin"]), float(g["sl_rmax"]), float(g["sl_scale"]), g["sl_ev"],
                     g["sl_ef"], g["sl_xi"], g["sl_p0"], g["sl_d0"])


def eof_table(g):
    p = g["eof_params"]
    return O.EOFTable(int(p[4]), int(p[5]), int(p[2]), int(p[3]), int(p[8]), p[0], p[1], p[6], p[7],
                      g["eof_potC"], g["eof_rforceC"], g["eof_zforceC"], g["eof_potS"],
                      g["eof_rforceS"], g["eof_zforceS"], g["eof_densC"], g["eof_densS"])


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_legendre(golden):
    for i, x in enumerate(golden["leg_x"]):
        assert rel(O.legendre_R(6, x)[:, :, 0], golden["leg_P"][i]) < 1e-13
        p, dp = O.dlegendre_R(6, x)
        assert rel(p[:, :, 0], golden["dleg_P"][i]) < 1e-13
        assert rel(dp[:, :, 0], golden["dleg_dP"][i]) < 1e-10


def test_factorial(golden):
    assert rel(O.factorial_return(6), golden["factorial6"]) < 1e-14


def test_sl_radial_functions(golden):
    t = sl_table(golden)
    rs = golden["sl_rs"]
    assert rel(O.get_halo_pot_matrix(t, rs), golden["sl_potmat"]) < 1e-13
    d, f, p = O.get_halo_dens_pot_force(t, rs, compat=True)
    assert rel(d, golden["sl_dens"]) < 1e-13
    assert rel(f, golden["sl_dpot"]) < 1e-12
    assert rel(p, golden["sl_pot"]) < 1e-13
    # the non-compat stencil differs only in the first cell
    _, fc, _ = O.get_halo_dens_pot_force(t, rs, compat=False)
    _, i, _, _, _ = O._sl_locate(t, rs)
    assert rel(fc[i > 0], golden["sl_dpot"][i > 0]) < 1e-12
    assert np.any(np.abs(fc[i == 0] - golden["sl_dpot"][i == 0]) > 0)


def test_sl_accumulate(golden):
    t = sl_table(golden)
    x, y, z, m = (golden[k] for k in ("px", "py", "pz", "pm"))
    assert rel(O.sl_accumulate(t, x, y, z, m), golden["sl_coef"]) < 1e-12
    assert rel(O.sl_accumulate(t, x, y, z, m, no_odd=True), golden["sl_coef_noodd"]) < 1e-12


def test_sl_all_eval_particles(golden):
    t = sl_table(golden)
    x, y, z = (golden[k] for k in ("px", "py", "pz"))
    c = golden["sl_coef"]
    with np.errstate(all="ignore"):
        got = O.sl_all_eval_particles(t, x, y, z, c, compat=True)
        got2 = O.sl_all_eval_particles(t, x, y, z, c, L1=1, L2=2, no_odd=True, compat=True)
    for a, b in zip(got, golden["sl_alleval"]):
        assert rel(a, b) < 1e-11
    for a, b in zip(got2, golden["sl_alleval_L12_noodd"]):
        assert rel(a, b) < 1e-11


@pytest.mark.parametrize("key,kw", [("sl_forceeval", {}), ("sl_forceeval_l2n3", dict(lmax=2, nmax=3)),
                                    ("sl_forceeval_noodd", dict(no_odd=True))])
def test_sl_force_eval(golden, key, kw):
    t = sl_table(golden)
    x, y, z = (golden[k][:120] for k in ("px", "py", "pz"))
    r3 = np.sqrt(x * x + y * y + z * z) + 1e-15
    got = O.sl_force_eval(t, r3, z / r3, np.arctan2(y, x), golden["sl_coef"], compat=True, **kw)
    ref = golden[key]
    for j in range(5):
        assert rel(got[j], ref[:, j]) < 1e-11, j


def test_eof_readers_and_grid(golden):
    p = golden["eof_params"]
    assert np.array_equal(p, golden["eof_params_old"])
    t = eof_table(golden)
    assert np.allclose([t.XMIN, t.XMAX, t.dX, t.YMIN, t.YMAX, t.dY], golden["eof_grid"], rtol=1e-15)


def test_return_bins(golden):
    t = eof_table(golden)
    x, y, z = (golden[k] for k in ("px", "py", "pz"))
    R = (x * x + y * y + 1e-10) ** 0.5
    X, Y, ix, iy = O.return_bins(R, z, t.XMIN, t.dX, t.YMIN, t.dY, t.numx, t.numy, t.ascale,
                                 t.hscale, t.cmap)
    assert np.array_equal(ix, golden["eof_ix"]) and np.array_equal(iy, golden["eof_iy"])
    assert np.allclose(X, golden["eof_X"], rtol=1e-14) and np.allclose(Y, golden["eof_Y"], rtol=1e-14)
    assert np.any(golden["eof_X"] > t.numx)   # the extrapolation case is exercised


def test_eof_accumulate(golden):
    t = eof_table(golden)
    x, y, z, m = (golden[k] for k in ("px", "py", "pz", "pm"))
    c, s = O.eof_accumulate(t, x, y, z, m)
    assert rel(c, golden["eof_cos"]) < 1e-13
    assert rel(s, golden["eof_sin"]) < 1e-13


@pytest.mark.parametrize("key,kw", [("eof_forceeval", {}), ("eof_forceeval_noodd", dict(no_odd=True)),
                                    ("eof_forceeval_m1n2", dict(MMAX=1, NMAX=2))])
def test_eof_force_eval(golden, key, kw):
    t = eof_table(golden)
    x, y, z = (golden[k][:120] for k in ("px", "py", "pz"))
    r2 = (x * x + y * y) ** 0.5 + 1e-15
    got = O.eof_force_eval(t, r2, z, np.arctan2(y, x), golden["eof_cos"], golden["eof_sin"],
                           compat=True, **kw)
    for j in range(5):
        assert rel(got[j], golden[key][:, j]) < 1e-12, j


def _field(golden):
    return O.FieldSpec(eof_table(golden), golden["eof_cos"], golden["eof_sin"], sl_table(golden),
                       golden["sl_coef"])


def test_fields_return_forces_cart(golden):
    if "fields_cart" not in golden.files:
        pytest.skip("fixture has no Fields vectors")
    F = _field(golden)
    x, y, z = (golden[k][:120] for k in ("px", "py", "pz"))
    rot = golden["fields_rotpos"]
    got = np.array([np.array(O.return_forces_cart(F, x[i], y[i], z[i], rot[i], compat=True))[:, 0]
                    for i in range(120)])
    ref = golden["fields_cart"]
    for j in range(8):
        assert rel(got[:, j], ref[:, j]) < 1e-11, j


def test_leapfrog(golden):
    if "leap_orbits" not in golden.files:
        pytest.skip("fixture has no orbit vectors")
    F = _field(golden)
    init = golden["leap_init"]
    orb = O.leapfrog(F, 200, float(golden["leap_dt"]), init[:, :3], init[:, 3:],
                     rotfreq=float(golden["leap_rotfreq"]), compat=True)
    ref = golden["leap_orbits"]
    for j, k in enumerate(("T", "X", "Y", "Z", "VX", "VY", "VZ", "P", "FX", "FY", "FZ")):
        got = orb[k] if k != "T" else np.broadcast_to(orb["T"], ref[:, 0].shape)
        assert rel(got, ref[:, j]) < 1e-10, k


def test_reader_parity(golden, tmp_path):
    """The product writers + readers round-trip the tables the reference's
    readers produced (the fixture's tables were read by the reference from
    files written by the same writers)."""
    from exptool_amd.utils import halo_methods as H
    g = golden
    lmax, nmax, numr = int(g["sl_lmax"]), int(g["sl_nmax"]), int(g["sl_numr"])
    H.write_cached_table(tmp_path / "s.cache", lmax, nmax, numr, int(g["sl_cmap"]),
                         float(g["sl_rmin"]), float(g["sl_rmax"]), float(g["sl_scale"]),
                         g["sl_ev"], g["sl_ef"])
    H.write_sph_model_table(tmp_path / "s.model", g["sl_r"], g["sl_d0"] / (4 * np.pi),
                            np.zeros(numr), g["sl_p0"])
    t = H.load_sl(tmp_path / "s.cache", tmp_path / "s.model")
    assert np.array_equal(t["ev"], g["sl_ev"]) and np.array_equal(t["ef"], g["sl_ef"])
    assert np.array_equal(t["xi"], g["sl_xi"])
    assert np.allclose(t["p0"], g["sl_p0"], rtol=1e-14, atol=0)
    assert np.allclose(t["d0"], g["sl_d0"], rtol=1e-14, atol=0)
    p = g["eof_params"]
    eof = dict(mmax=int(p[4]), norder=int(p[5]), numx=int(p[2]), numy=int(p[3]), cmap=int(p[8]),
               rmin=p[0], rmax=p[1], ascale=p[6], hscale=p[7], dens=int(p[9]))
    for k in ("potC", "rforceC", "zforceC", "densC", "potS", "rforceS", "zforceS", "densS"):
        eof[k] = g["eof_" + k]
    E.write_eof_cache(tmp_path / "e.cache", eof, newstyle=True)
    E.write_eof_cache(tmp_path / "o.cache", eof, newstyle=False)
    for f in ("e.cache", "o.cache"):
        arrs = E.parse_eof(tmp_path / f)
        for k, a in zip(("potC", "rforceC", "zforceC", "densC", "potS", "rforceS", "zforceS", "densS"), arrs):
            assert np.array_equal(a, g["eof_" + k]), (f, k)
        assert np.allclose(E.eof_params(tmp_path / f), g["eof_params"])


@pytest.mark.parametrize("key,kw", [("sl_alleval_pt", {}), ("sl_alleval_pt_n3", dict(nmax=3)),
                                    ("sl_alleval_pt_noodd", dict(no_odd=True))])
def test_sl_all_eval_point(golden, key, kw):
    """spheresl.all_eval (den1 seeded with l = 0, densfac 0.25/pi, cos m phi)."""
    t = sl_table(golden)
    x, y, z = (golden[k][:120] for k in ("px", "py", "pz"))
    r3 = np.sqrt(x * x + y * y + z * z) + 1e-15
    with np.errstate(all="ignore"):
        got = O.sl_all_eval(t, r3, z / r3, np.arctan2(y, x), golden["sl_coef"], **kw)
    ref = golden[key]
    for j in range(7):
        assert rel(got[j], ref[:, j]) < 1e-11, j


@pytest.mark.parametrize("key,kw", [("eof_acceval", {}), ("eof_acceval_noodd", dict(no_odd=True)),
                                    ("eof_acceval_m1", dict(MMAX=1))])
def test_eof_accumulated_eval(golden, key, kw):
    """eof.accumulated_eval: (p0, p, fr, fp, fz, d0, d) with density."""
    t = eof_table(golden)
    x, y, z = (golden[k][:120] for k in ("px", "py", "pz"))
    r2 = (x * x + y * y) ** 0.5 + 1e-15
    got = O.eof_accumulated_eval(t, r2, z, np.arctan2(y, x), golden["eof_cos"],
                                 golden["eof_sin"], **kw)
    for j in range(7):
        assert rel(got[j], golden[key][:, j]) < 1e-12, j


def test_eof_eval_particles_pinned(golden):
    """Every output of the batch evaluator the GPU EOF evaluator is compared
    with: against accumulated_eval at the radius accumulated_eval_particles
    computes (sqrt(x^2 + y^2 + 1e-10)), and its forces against force_eval."""
    t = eof_table(golden)
    x, y, z = (golden[k][:120] for k in ("px", "py", "pz"))
    p0, p, d0, d, fr, fp, fz, R = O.eof_eval_particles(t, x, y, z, golden["eof_cos"],
                                                       golden["eof_sin"], density=True)
    ref = golden["eof_acceval_part"]
    for j, v in enumerate((p0, p0 + p, fr, fp, fz, d0, d0 + d)):
        assert rel(v, ref[:, j]) < 1e-12, j
    assert rel(p, ref[:, 1] - ref[:, 0]) < 1e-10
    assert rel(d, ref[:, 6] - ref[:, 5]) < 1e-10
    # the two reference point evaluators agree where they overlap (same r)
    a, f = golden["eof_acceval"], golden["eof_forceeval"]
    for ja, jf in ((2, 0), (3, 1), (4, 2), (1, 3), (0, 4)):
        assert rel(a[:, ja], f[:, jf]) < 1e-12, (ja, jf)
