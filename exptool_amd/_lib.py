"""ctypes binding of libexptool_bfe.so (the C ABI declared in include/exptool_bfe.h).

torch is imported before the library is loaded: torch ships its own HIP
runtime with the same SONAME (libamdhip64.so.7), so the library binds to that
copy and device pointers from torch tensors are valid in it.  There is no CPU
fallback: if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np
import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EXPTOOL_BFE_LIB", os.path.join(_HERE, "libexptool_bfe.so"))

# flags (include/exptool_bfe.h)
COMPAT_SL_COS_L = 0x1
COMPAT_SL_EDGE = 0x2
COMPAT_SL_DENS = 0x4
COMPAT_EOF_MASK = 0x8
NO_ODD = 0x10
ACC_BF16 = 0x20
COMPAT_ALL = 0xF
COMPAT_RK4_FRAME = 0x40
EVAL_DENS = 0x80

EXPORTS = (
    "bfe_create", "bfe_destroy", "bfe_last_error", "bfe_sl_tables_set", "bfe_eof_tables_set",
    "bfe_sl_accumulate", "bfe_eof_accumulate", "bfe_sl_eval", "bfe_eof_eval", "bfe_field_set",
    "bfe_field_eval_cart", "bfe_leapfrog", "bfe_leapfrog_dt", "bfe_rk4", "bfe_mlp_set_params", "bfe_mlp_get_params",
    "bfe_mlp_forward", "bfe_mlp_gradient", "bfe_mlp_train_step", "bfe_mlp_leapfrog",
    "bfe_eof_accumulate_batch", "bfe_sl_accumulate_batch", "bfe_field_set_batch",
    "bfe_field_eval_batch", "bfe_accumulate_disc_halo", "bfe_forces_disc_halo",
    "bfe_set_tuning", "bfe_get_tuning", "bfe_kernel_times", "bfe_sl_eval_sph", "bfe_eof_eval_cyl",
    "bfe_mlp_set_base", "bfe_mlp_set_mode", "bfe_eof_accumulate_noalloc",
)
MLP_NPARAMS = 50564

_lib = None
_lock = threading.Lock()

_VP, _I32, _I64, _U32, _F64 = C.c_void_p, C.c_int, C.c_int64, C.c_uint32, C.c_double


def _declare(lib):
    sig = {
        "bfe_create": (_I32, [_I32, C.POINTER(_VP)]),
        "bfe_destroy": (_I32, [_VP]),
        "bfe_last_error": (C.c_char_p, [_VP]),
        "bfe_sl_tables_set": (_I32, [_VP, _I32, _I32, _I32, _I32, _F64, _VP, _VP, _VP, _VP, _VP]),
        "bfe_eof_tables_set": (_I32, [_VP, _I32, _I32, _I32, _I32, _I32, _F64, _F64, _F64, _F64, _F64, _F64,
                                     _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
        "bfe_sl_accumulate": (_I32, [_VP, _VP, _VP, _VP, _VP, _I64, _U32, _VP, _VP]),
        "bfe_eof_accumulate": (_I32, [_VP, _VP, _VP, _VP, _VP, _I64, _U32, _VP, _VP, _VP]),
        "bfe_eof_accumulate_noalloc": (_I32, [_VP, _VP, _VP, _VP, _VP, _I64, _U32, _VP, _VP, _VP]),
        "bfe_sl_eval": (_I32, [_VP, _VP, _VP, _VP, _I64, _VP, _I32, _I32, _U32] + [_VP] * 7 + [_VP]),
        "bfe_eof_eval": (_I32, [_VP, _VP, _VP, _VP, _I64, _VP, _VP, _I32, _I32, _U32] + [_VP] * 7 + [_VP]),
        "bfe_sl_eval_sph": (_I32, [_VP, _VP, _VP, _VP, _I64, _VP, _I32, _I32, _U32, _VP, _VP]),
        "bfe_eof_eval_cyl": (_I32, [_VP, _VP, _VP, _VP, _I64, _VP, _VP, _I32, _I32, _U32, _VP, _VP]),
        "bfe_field_set": (_I32, [_VP, _VP, _VP, _VP, _F64, _I32, _I32, _I32, _I32, _U32, _VP]),
        "bfe_field_eval_cart": (_I32, [_VP, _VP, _VP, _VP, _I64, _F64, _VP, _VP, _VP, _VP, _VP, _VP]),
        "bfe_leapfrog": (_I32, [_VP] + [_VP] * 6 + [_I64, _F64, _I64, _F64, _F64, _I32, _VP, _VP, _I32, _VP,
                                                  _I32, _VP]),
        "bfe_leapfrog_dt": (_I32, [_VP] + [_VP] * 6 + [_I64, _VP, _I64, _F64, _F64, _I32, _VP, _VP, _I32,
                                                     _VP, _I32, _VP]),
        "bfe_rk4": (_I32, [_VP] + [_VP] * 6 + [_I64, _F64, _I64, _F64, _U32, _I32, _VP, _VP]),
        "bfe_eof_accumulate_batch": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP, _I32, _U32, _VP, _VP, _VP]),
        "bfe_sl_accumulate_batch": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP, _I32, _U32, _VP, _VP]),
        "bfe_field_set_batch": (_I32, [_VP, _VP, _VP, _VP, _I32, _F64, _I32, _I32, _I32, _I32, _U32, _VP]),
        "bfe_field_eval_batch": (_I32, [_VP, _VP, _VP, _VP, _VP, _I32, _F64, _VP, _VP, _VP, _VP, _VP]),
        "bfe_accumulate_disc_halo": (_I32, [_VP, _VP, _VP, _VP, _VP, _I64, _VP, _VP, _VP, _VP, _I64,
                                            _U32, _VP, _VP, _VP, _VP]),
        "bfe_forces_disc_halo": (_I32, [_VP, _VP, _VP, _VP, _VP, _I64, _VP, _VP, _VP, _VP, _I64, _U32,
                                        _VP, _VP, _VP, _F64, _I32, _I32, _I32, _I32, _U32, _F64,
                                        _VP, _VP, _VP, _VP, _VP, _VP]),
        "bfe_set_tuning": (_I32, [_VP, C.c_char_p, _I64]),
        "bfe_get_tuning": (_I32, [_VP, C.c_char_p, C.POINTER(_I64)]),
        "bfe_kernel_times": (_I32, [_VP, C.c_char_p, C.POINTER(_F64), C.POINTER(_I64)]),
        "bfe_mlp_set_params": (_I32, [_VP, _VP, _I64, _VP]),
        "bfe_mlp_get_params": (_I32, [_VP, _VP, _I64]),
        "bfe_mlp_set_base": (_I32, [_VP, _VP, _I32]),
        "bfe_mlp_set_mode": (_I32, [_VP, _I32, _VP]),
        "bfe_mlp_forward": (_I32, [_VP, _VP, _VP, _VP, _I64, _VP, _VP]),
        "bfe_mlp_gradient": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP, _I64, _VP, _VP, _VP]),
        "bfe_mlp_train_step": (_I32, [_VP, _VP, _VP, _VP, _VP, _VP, _I64, C.c_float, C.c_float,
                                      C.c_float, C.c_float, _VP, _VP]),
        "bfe_mlp_leapfrog": (_I32, [_VP] + [_VP] * 6 + [_I64, _F64, _I64, _F64, _F64, _I32,
                                                      C.c_float, C.c_float, _I32, _VP, _VP, _VP,
                                                      _VP, _VP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load():
    """Load (once) and return the library; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built: run `python -m exptool_amd.csrc.build` "
                    "(there is no CPU fallback for the BFE kernels)")
            lib = C.CDLL(LIB_PATH)
            _declare(lib)
            _lib = lib
    return _lib


class BFEError(RuntimeError):
    pass


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def host_ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Context:
    """A bfe_ctx bound to one device (one per host thread)."""

    def __init__(self, device=0):
        self.lib = load()
        if not torch.cuda.is_available():
            raise BFEError("exptool_amd: no GPU visible (torch.cuda.is_available() is False)")
        self.device = int(device)
        torch.cuda.set_device(self.device)
        torch.cuda.init()
        h = C.c_void_p()
        rc = self.lib.bfe_create(self.device, C.byref(h))
        if rc != 0:
            raise BFEError(f"bfe_create failed ({rc}): {self.lib.bfe_last_error(None).decode()}")
        self.h = h
        self._keep = []

    def check(self, rc, what):
        if rc != 0:
            msg = self.lib.bfe_last_error(self.h).decode()
            raise BFEError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.bfe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def f32(a, device="cuda"):
    """Contiguous float32 device tensor from a numpy array or tensor."""
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)


def f64(a, device="cuda"):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
