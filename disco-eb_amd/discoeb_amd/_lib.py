"""ctypes binding of ``libdiscoeb_hip.so`` (the gfx950 C ABI, include/discoeb_hip.h).

The product path has no CPU fallback: if the library or a GPU is missing,
every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

from .packing import DebConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiscoeb_hip.so")
_LIB = None

_dp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


class DebError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise DebError(f"{LIB_PATH} is missing: build it with `python disco-eb_amd/build.py` "
                           "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.deb_abi_version.restype = ctypes.c_int
        L.deb_last_error.restype = ctypes.c_char_p
        L.deb_nvar.argtypes = [ctypes.POINTER(DebConfig)]
        L.deb_slots.argtypes = [ctypes.POINTER(DebConfig)]
        L.deb_workspace_bytes.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_size_t)]
        L.deb_solve_forward.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, _dp, _dp, _i64, _dp, _dp, _dp,
                                        ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_size_t,
                                        _dp]
        L.deb_start_times.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, _dp, _dp, _i64, _dp, _dp, _dp,
                                      ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.deb_rhs.argtypes = [ctypes.POINTER(DebConfig), _dp, _dp, _i64, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                              _dp, _dp, _dp]
        L.deb_rodas5_step.argtypes = [ctypes.POINTER(DebConfig), _dp, _dp, _i64, _dp, _dp, _dp, ctypes.c_int, _dp,
                                      _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.deb_batched_workspace_bytes.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_size_t)]
        L.deb_solve_forward_batched.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, _dp, _dp, _i64, _dp, _dp,
                                                _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                                _dp, _dp, ctypes.c_size_t, _dp]
        L.deb_solve_forward_batched_noalloc.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, _dp, _dp, _i64,
                                                        _dp, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                                        ctypes.c_int, _dp, _dp, _dp, _dp]
        L.deb_tape_bytes.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_size_t)]
        L.deb_solve_forward_tape.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, _dp, _dp, _i64, _dp, _dp, _dp,
                                             ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _dp,
                                             ctypes.c_size_t, _dp]
        L.deb_adjoint_workspace_bytes.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
        L.deb_solve_adjoint.argtypes = [ctypes.POINTER(DebConfig), ctypes.c_int, _dp, _dp, _i64, _dp, _dp, _dp,
                                        ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                        _dp, ctypes.c_int, _dp, ctypes.c_size_t, _dp, ctypes.c_size_t, _dp]
        L.deb_rhs_vjp.argtypes = [ctypes.POINTER(DebConfig), _dp, _dp, _i64, _dp, _dp, _dp, ctypes.c_int, _dp, _dp,
                                  ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.deb_rodas5_step_vjp.argtypes = [ctypes.POINTER(DebConfig), _dp, _dp, _i64, _dp, _dp, _dp, ctypes.c_int,
                                          _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                          _dp]
        _i = ctypes.c_int
        L.deb_observables_scratch_bytes.restype = ctypes.c_size_t
        L.deb_observables_scratch_bytes.argtypes = [_i, _i, _i]
        L.deb_power.argtypes = [_i, _i, _dp, _dp, _i, _i, _dp, _dp, _dp, _dp, _dp]
        L.deb_power_smoothed.argtypes = [_i, _i, _dp, _dp, _i, _i, _dp, _dp, _dp, _i, _i, _dp, ctypes.c_size_t, _dp,
                                         _dp]
        L.deb_power_kaiser.argtypes = [_i, _i, _i, _dp, _dp, _i, _dp, _dp, _dp, _dp, _dp, _i, _i, _dp,
                                       ctypes.c_size_t, _dp, _dp]
        L.deb_power_multipoles.argtypes = [_i, _i, _dp, _dp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.deb_xi_from_power.argtypes = [_i, _i, _dp, _dp, _i, _dp, _dp, _dp]
        if L.deb_abi_version() != 3:
            raise DebError("libdiscoeb_hip.so ABI version mismatch")
        _LIB = L
    return _LIB


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().deb_last_error().decode(errors="replace")
        raise DebError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> ctypes.c_void_p:
    """Address of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return ctypes.c_void_p(0)
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)
