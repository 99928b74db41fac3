"""ctypes binding of ``libchol_amd.so`` (the C ABI in include/chol_amd.h).

The library is loaded once, from this package directory only.  There is no
fallback: a missing or unloadable library raises ``ChollibError`` on first
use, so no caller can silently get a non-HIP result.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchol_amd.so")

MODE_PLAIN = 0
MODE_MIXED = 1
MODE_FP8 = 2

E_ARG = -1
E_WORKSPACE = -2
E_MODE = -3

# Every symbol include/chol_amd.h declares, with its ctypes signature.
_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_sz = ctypes.c_size_t
SIGNATURES = {
    "chol_abi_version": (ctypes.c_int, []),
    "chol_last_error": (ctypes.c_char_p, []),
    "chol_workspace_bytes": (_sz, [_i64, _i64, _i32]),
    "chol_potrf_f32": (ctypes.c_int, [_p, _p, _i64, _i64, _i64, _i64, _i32, _p, _p, _sz, _p]),
    "chol_set_tuning": (ctypes.c_int, [_p, _i64]),
    "chol_ragged_workspace_bytes": (_sz, [_p, _i64]),
    "chol_potrf_ragged_f32": (ctypes.c_int, [_p, _p, _p, _p, _i64, _i32, _p, _p, _sz, _p]),
    "chol_backward_workspace_bytes": (_sz, [_i64, _i64]),
    "chol_backward_f32": (ctypes.c_int, [_p, _p, _p, _i64, _i64, _p, _i32, _p, _sz, _p]),
    "chol_tri_inverse_workspace_bytes": (_sz, [_i64, _i64]),
    "chol_tri_inverse_f32": (ctypes.c_int, [_p, _p, _i64, _i64, _p, _sz, _p]),
    "chol_route": (ctypes.c_int, [_i64, _i64, _i32, _p]),
    "chol_plan_potrf": (ctypes.c_int, [_i64, _i64, _i64, _i32, _i32, _p, _i64, _p]),
    "chol_route_counters": (_i64, [_p, _i64]),
    "chol_potrf_f32_noalloc": (ctypes.c_int, [_p, _p, _i64, _i64, _i64, _i64, _i32, _p, _p]),
    "chol_trailing_update_f32": (ctypes.c_int, [_p, _p, _i64, _i64, _i64, _i64, _i64, _i32, _p]),
}


class ChollibError(RuntimeError):
    """Raised when the HIP library is missing or a call returns non-zero."""


_lib = None
_lock = threading.Lock()


def load():
    """Load and bind the library (once).  Raises ChollibError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ChollibError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m gpumode_cholesky_amd.build` (hipcc, gfx950)")
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as exc:
            raise ChollibError(f"cannot load {LIB_PATH}: {exc}") from exc
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        # The autotuner's route / blocking table, when one has been written
        # (once, at first load; set_table re-enters load() after _lib is set).
        from . import autotune
        autotune.load_table()
    return _lib


def last_error() -> str:
    msg = load().chol_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise ChollibError(f"{what} failed with code {rc}: {last_error()}")
