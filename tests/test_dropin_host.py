"""Host logic of the drop-in modules: the PSP reader on a data file the
reference's own tests hold (tests/initial001dump.dat, old-style info strings,
float records), PSP and SL-coefficient writers and readers, the apse stop and
the force_eval truncation rule.  No GPU calls."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PSP = os.path.join(HERE, "golden", "psp_initial001dump.dat")


def test_psp_reference_fixture():
    from exptool_amd.io import particle, psp_io
    I = psp_io.Input(PSP)
    assert I.time == 0.0 and list(I.header) == ["dark", "star"]
    assert I._float_len == 4
    d = I.header["dark"]
    assert d["nbodies"] == 30000 and d["id"] == "sphereSL"
    assert d["force"]["parameters"]["Lmax"] == 6 and d["force"]["parameters"]["nmax"] == 20
    assert d["parameters"]["indexing"] is False
    s = I.header["star"]
    assert s["force"]["parameters"]["mmax"] == 6 and s["force"]["parameters"]["ncylorder"] == 12
    # the two components tile the file exactly
    assert d["data_start"] == 16 + 24 + 1024 and s["data_end"] == os.path.getsize(PSP)
    star = psp_io.Input(PSP, comp="star").data
    assert set(star) == {"m", "x", "y", "z", "vx", "vy", "vz", "potE"}
    assert star["x"].dtype == np.float32 and star["x"].size == 30000
    assert np.all(star["m"] > 0) and np.all(np.isfinite(star["potE"]))
    L = particle.Input(PSP, comp="dark", legacy=True)
    assert L.xpos.size == 30000 and L.mass.dtype == np.float32
    with pytest.raises(IOError):
        psp_io.Input(PSP, comp="gas")


@pytest.mark.parametrize("float_len,index", [(4, False), (8, True)])
def test_psp_roundtrip(tmp_path, float_len, index):
    from exptool_amd.io import psp_io
    g = np.random.default_rng(1)
    comps = []
    for name, n in (("disc", 100), ("halo", 57)):
        d = {k: g.normal(size=n) for k in ("m", "x", "y", "z", "vx", "vy", "vz", "potE")}
        if index:
            d["id"] = np.arange(n, dtype=np.int64) + 7
        comps.append(({"name": name, "id": "x", "parameters": {"a": 1}}, d))
    f = tmp_path / "OUT.t.00001"
    psp_io.write_psp(f, 2.25, comps, float_len=float_len)
    for (h, d) in comps:
        I = psp_io.Input(str(f), comp=h["name"])
        assert I.time == 2.25 and I._float_len == float_len
        dt = np.float32 if float_len == 4 else np.float64
        for k, v in d.items():
            assert np.array_equal(I.data[k], v.astype(np.int64 if k == "id" else dt)), k


def test_sl_coefficient_file_roundtrip(tmp_path):
    from exptool_amd.basis import spheresl
    f = str(tmp_path / "coefs.sl")
    objs = []
    for t in (0.5, 1.0):
        o = spheresl.SL_Object()
        o.time, o.filename, o.comp, o.nbodies = t, "OUT.run.001", "dark", 1234
        o.sph_file, o.model_file, o.lmax, o.nmax = "sl.cache", "sl.model", 3, 5
        o.expcoef = np.random.default_rng(int(t * 10)).normal(size=(16, 5))
        spheresl.save_sl_coefficients(f, o)
        objs.append(o)
    assert os.path.getsize(f) == 4 + 2 * (324 + 8 * 16 * 5)
    last, d = spheresl.restore_sl_coefficients(f)
    assert sorted(d) == [0.5, 1.0]
    for o in objs:
        r = d[o.time]
        assert np.array_equal(r.expcoef, o.expcoef) and r.lmax == 3 and r.nmax == 5
        assert r.comp == b"dark" and r.nbodies == 1234
    assert last.time == 1.0


def test_force_eval_truncation_rule():
    from exptool_amd.basis.spheresl import _truncated
    c = np.ones((16, 6))
    assert np.array_equal(_truncated(c, 3, 5, 3, 6), c)          # nmax_table - 1: no cut
    assert np.array_equal(_truncated(c, 3, 6, 3, 6), c)
    t = _truncated(c, 1, 2, 3, 6)
    assert t[:4, :2].sum() == 8 and t.sum() == 8


def test_apse_stop_counts_apocentres():
    from exptool_amd.utils.integrate import _apse_stop
    t = np.linspace(0, 10 * np.pi, 2001)
    R = 1.0 + 0.2 * np.cos(t)          # apocentres at t = 2 pi k
    X, Y = R * np.cos(0.3 * t), R * np.sin(0.3 * t)
    n1 = _apse_stop(X, Y, 1)
    n2 = _apse_stop(X, Y, 2)
    # first interior maximum of R at t = 2 pi (step 400): counted at step 401
    assert n1 == 402 and n2 == 802
    assert _apse_stop(X, Y, 100) == t.size


def test_eof_variance_and_sn_hand_values():
    """eof.compute_variance / compute_sn (eof.py:2135-2226) on values worked
    by hand: masses (1, 1, 2) so w = 1/4, n = 3; cos = 8, cos2 = 32 give
    var = 32/16 - 4/3 = 2/3, b_Hall = 4 / ((2/3)/4 + 4 + 1e-10) = 24/25 and
    S/N = sqrt(64 / (2/3)) = sqrt(96); sin = -4, sin2 = 8 give
    var = 1/2 - 1/3 = 1/6, b_Hall = 1 / ((1/6)/4 + 1 + 1e-10) and
    S/N = sqrt(96)."""
    from exptool_amd.basis import eof
    from exptool_amd.io import particle
    P = particle.make_holder(np.zeros(3), np.zeros(3), np.zeros(3), np.array([1.0, 1.0, 2.0]))
    c, s = np.array([[8.0]]), np.array([[-4.0]])
    c2, s2 = np.array([[[32.0]]]), np.array([[[8.0]]])
    varC, varS, facC, facS = eof.compute_variance(P, c, s, c2, s2)
    np.testing.assert_allclose(varC, [[[2 / 3]]], rtol=1e-14)
    np.testing.assert_allclose(varS, [[[1 / 6]]], rtol=1e-13)
    np.testing.assert_allclose(facC, [[[4 / (1 / 6 + 4 + 1e-10)]]], rtol=1e-14)
    np.testing.assert_allclose(facS, [[[1 / (1 / 24 + 1 + 1e-10)]]], rtol=1e-14)
    snC, snS = eof.compute_sn(P, c, s, c2, s2)
    np.testing.assert_allclose(snC, [[[96 ** 0.5]]], rtol=1e-14)
    np.testing.assert_allclose(snS, [[[96 ** 0.5]]], rtol=1e-13)


BARFILE = os.path.join(HERE, "golden", "barfile001.dat")


def test_bar_file_fixture_and_pattern_speed(tmp_path):
    """barfile001.dat (the reference's own test data: time, bar position,
    pattern speed) pins BarDetermine.read_bar, the backward differences of
    frequency_and_derivative (pattern.py:362-402), unwrap_bar_position and
    find_barpattern / find_barangle at tabulated times."""
    from exptool_amd.analysis import pattern
    ref = np.loadtxt(BARFILE)
    B = pattern.BarDetermine(file=BARFILE)
    assert np.array_equal(B.time, ref[:, 0]) and np.array_equal(B.pos, ref[:, 1])
    B.frequency_and_derivative()
    assert np.max(np.abs(B.deriv - ref[:, 2])) <= 1e-6      # file values printed to ~12 digits
    assert B.deriv[0] == 0.0
    i = 1234
    assert pattern.find_barpattern(ref[i, 0] + 1e-4, B, smth_order=None) == B.deriv[i]
    assert np.array_equal(pattern.find_barpattern(ref[[5, 9], 0], B, smth_order=None), B.deriv[[5, 9]])
    assert abs(pattern.find_barangle(ref[i, 0], B) + ref[i, 1]) <= 1e-9
    assert pattern.find_barangle(ref[i, 0] + 1e-4, B, interpolate=False) == -ref[i, 1]
    fit = pattern.find_barpattern(ref[i, 0], B, smth_order=2)
    d0 = np.zeros(len(ref)); d0[1:] = np.diff(ref[:, 1]) / np.diff(ref[:, 0])
    assert fit == pytest.approx(np.polyval(np.polyfit(ref[:, 0], d0, 2), ref[i, 0]), rel=1e-9)
    # the stored position is unwrapped: wrapping it into the m = 2 phase range
    # and unwrapping again gives it back
    W = pattern.BarDetermine()
    W.time, W.pos = ref[:, 0].copy(), (ref[:, 1] + np.pi / 2) % np.pi - np.pi / 2
    W.unwrap_bar_position()
    assert np.max(np.abs(W.pos - ref[:, 1])) <= 1e-9
    W.frequency_and_derivative()
    W.print_bar(str(tmp_path / "bar.dat"))
    R = pattern.BarDetermine(file=str(tmp_path / "bar.dat"))
    assert np.array_equal(R.pos, W.pos) and np.array_equal(R.deriv, W.deriv)


def test_bar_transform_given_angle():
    from exptool_amd.analysis import pattern
    from exptool_amd.io import particle
    P = particle.make_holder(np.array([1.0, 0.0]), np.array([0.0, 2.0]), np.array([0.5, -0.5]),
                             np.array([1.0, 2.0]))
    T = pattern.BarTransform(P, bar_angle=np.pi / 2, rel_bar_angle=np.pi / 2)
    assert T.bar_angle == np.pi
    np.testing.assert_allclose(T.data['x'], [-1.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(T.data['y'], [0.0, -2.0], atol=1e-15)
    assert np.array_equal(T.data['z'], [0.5, -0.5]) and np.array_equal(T.data['m'], [1.0, 2.0])
