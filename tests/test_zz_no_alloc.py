"""The allocation-free EOF accumulate (bfe_eof_accumulate_noalloc) against
bfe_eof_accumulate, both called through the C ABI: two calls in a row, the
second with fewer particles, each checked against the allocating entry
point (bit for bit when that one also runs in input order, <= 1e-6 per m
block against its cell-ordered form) and against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import bfe_oracle as O
from tests.test_gpu_parity import block_rel_eof, disc_halo, f32

pytestmark = pytest.mark.gpu


def _declare(lib):
    fn = lib.bfe_eof_accumulate_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p] * 5 + [C.c_int64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    return fn


@pytest.fixture(scope="module")
def sysm():
    from exptool_amd.basis import tablegen as tg
    from exptool_amd.bfe import BFESystem
    from exptool_amd.utils import samplers as S
    d = S.exponential_disc(30000, seed=3)
    et = tg.make_eof_tables(d, mmax=6, norder=18, numx=256, numy=128, lmax_s=16, nmax_s=12)
    eo = O.EOFTable(et["mmax"], et["norder"], et["numx"], et["numy"], et["cmap"], et["rmin"],
                    et["rmax"], et["ascale"], et["hscale"], et["potC"], et["rforceC"],
                    et["zforceC"], et["potS"], et["rforceS"], et["zforceS"])
    return BFESystem(eof=et), eo


def test_noalloc_matches_allocating_entry_point(sysm):
    import torch
    from exptool_amd._lib import ptr, stream_ptr
    s, eo = sysm
    fn = _declare(s.ctx.lib)
    shape = (s.eof["mmax"] + 1, s.eof["norder"])
    for n, seed in ((300_000, 51), (70_001, 52)):
        print(f"noalloc accumulate: n={n} seed={seed}", flush=True)
        (x, y, z, m), _ = disc_halo(n, seed)
        x, y, z, m = f32(x, y, z, m)
        X, Y, Z, Mm = (torch.tensor(v, dtype=torch.float32, device="cuda") for v in (x, y, z, m))
        cos = torch.full(shape, np.nan, dtype=torch.float64, device="cuda")
        sin = torch.full(shape, np.nan, dtype=torch.float64, device="cuda")
        rc = fn(s.ctx.h, ptr(X), ptr(Y), ptr(Z), ptr(Mm), n, 0, ptr(cos), ptr(sin), stream_ptr())
        s.ctx.check(rc, "bfe_eof_accumulate_noalloc")
        got_c, got_s = cos.cpu().numpy(), sin.cpu().numpy()
        # the allocating entry point in input order: same kernel, same segments
        s.set_tuning("order_min", 1 << 40)
        ref_c, ref_s = (v.cpu().numpy() for v in s.eof_accumulate(X, Y, Z, Mm))
        s.set_tuning("order_min", 65536)
        assert np.array_equal(got_c, ref_c) and np.array_equal(got_s, ref_s)
        # its cell-ordered form and the oracle
        cc, cs = (v.cpu().numpy() for v in s.eof_accumulate(X, Y, Z, Mm))
        assert block_rel_eof(got_c, cc) <= 1e-6 and block_rel_eof(got_s, cs) <= 1e-6
        oc, os_ = O.eof_accumulate(eo, x, y, z, m)
        assert block_rel_eof(got_c, oc) <= 1e-6 and block_rel_eof(got_s, os_) <= 1e-6
        assert np.all(got_s[0] == 0)
    # an empty input gives zero coefficients
    e = torch.zeros(0, dtype=torch.float32, device="cuda")
    rc = fn(s.ctx.h, ptr(e), ptr(e), ptr(e), ptr(e), 0, 0, ptr(cos), ptr(sin), stream_ptr())
    s.ctx.check(rc, "bfe_eof_accumulate_noalloc")
    assert np.all(cos.cpu().numpy() == 0) and np.all(sin.cpu().numpy() == 0)
