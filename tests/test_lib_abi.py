"""CPU checks of the C ABI library: it is built, loads next to torch's HIP
runtime, and exports every entry point include/*.h declares."""
import ctypes
import glob
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(bfe_[a-z0-9_]+)\s*\(", src))
    return names


def test_header_declares_entry_points():
    d = declared()
    assert {"bfe_create", "bfe_destroy", "bfe_sl_accumulate", "bfe_eof_accumulate",
            "bfe_field_eval_cart", "bfe_leapfrog"} <= d


def test_library_exports_every_declared_symbol():
    from exptool_amd import _lib
    lib = _lib.load()
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= declared()


def test_single_hip_runtime_in_process():
    """The library must bind to torch's libamdhip64 (same SONAME), not load a second one."""
    import torch  # noqa: F401
    from exptool_amd import _lib
    _lib.load()
    maps = open("/proc/self/maps").read()
    libs = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    assert len(libs) == 1, libs


def test_create_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from exptool_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.bfe_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert b"device" in lib.bfe_last_error(None).lower() or len(lib.bfe_last_error(None)) > 0
