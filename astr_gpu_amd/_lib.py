"""ctypes binding of the C ABI declared in include/astr_hip.h.

The shared library is built in-tree (astr_gpu_amd/libastr_hip.so) by
``astr_gpu_amd.build.build_library``.  There is no fallback: if the library
is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libastr_hip.so")

ASTR_OK = 0
ERRORS = {1: "ASTR_E_ARG", 2: "ASTR_E_UNSUPPORTED", 3: "ASTR_E_HIP",
          4: "ASTR_E_NOMEM", 5: "ASTR_E_STATE"}

# every symbol the header declares, with (restype, argtypes)
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int32


class AstrParams(ctypes.Structure):
    """Mirror of struct astr_params (include/astr_hip.h)."""
    _fields_ = [
        ("im", ctypes.c_int32), ("jm", ctypes.c_int32), ("km", ctypes.c_int32),
        ("hm", ctypes.c_int32), ("numq", ctypes.c_int32),
        ("npdci", ctypes.c_int32), ("npdcj", ctypes.c_int32), ("npdck", ctypes.c_int32),
        ("variant", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("alfa_filter", ctypes.c_double), ("reynolds", ctypes.c_double),
        ("mach", ctypes.c_double), ("prandtl", ctypes.c_double),
        ("gamma", ctypes.c_double), ("ref_tem", ctypes.c_double),
        ("sutherland", ctypes.c_double), ("beta_halo", ctypes.c_double),
        ("beta_bound", ctypes.c_double),
        ("convection", ctypes.c_int32), ("deriv_explicit", ctypes.c_int32),
        ("recon_schem", ctypes.c_int32), ("lchardecomp", ctypes.c_int32),
        ("lfilter", ctypes.c_int32), ("diffterm", ctypes.c_int32),
        ("shkcrt", ctypes.c_double), ("bfacmpld", ctypes.c_double),
    ]


SIGNATURES = {
    "astr_default_params": (None, [ctypes.POINTER(AstrParams), _I]),
    "astr_create": (ctypes.c_int, [ctypes.POINTER(AstrParams), _I, ctypes.POINTER(_P)]),
    "astr_destroy": (ctypes.c_int, [_P]),
    "astr_upload_geom": (ctypes.c_int, [_P, _D, _D]),
    "astr_upload_state": (ctypes.c_int, [_P, _D, _D, _D, _D, _D]),
    "astr_download_state": (ctypes.c_int, [_P, _D, _D, _D, _D, _D]),
    "astr_filterq": (ctypes.c_int, [_P]),
    "astr_qswap": (ctypes.c_int, [_P]),
    "astr_rhscal": (ctypes.c_int, [_P]),
    "astr_rk3_update": (ctypes.c_int, [_P, _I, ctypes.c_double]),
    "astr_stats": (ctypes.c_int, [_P, _D]),
    "astr_stats_device": (ctypes.c_int, [_P, _P]),
    "astr_rk3_step": (ctypes.c_int, [_P, ctypes.c_double, _D]),
    "astr_rk3_steps": (ctypes.c_int, [_P, _I, ctypes.c_double, _D]),
    "astr_download_rhs": (ctypes.c_int, [_P, _D, _D, _D]),
    "astr_download_shock": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint8)]),
    "astr_deriv_field": (ctypes.c_int, [_P, _I, _D, _D]),
    "astr_filter_field": (ctypes.c_int, [_P, _I, _D, _D]),
    "astr_flux_field": (ctypes.c_int, [_P, _I, _I, _I, _D, _D]),
    "astr_set_timing": (ctypes.c_int, [_P, _I]),
    "astr_synchronize": (ctypes.c_int, [_P]),
    "astr_set_spacing": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "astr_set_bc": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "astr_set_lengths": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "astr_set_twall": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "astr_set_xslip": (ctypes.c_int, [_P, ctypes.c_double]),
    "astr_line_adjoint": (ctypes.c_int, [_P, _I, _I, _I, _D, _D]),
    "astr_adjoint_enstrophy": (ctypes.c_int, [_P, _I, ctypes.c_double, _I, _D, _D, _D, _D]),
    "astr_adjoint_record": (ctypes.c_int, [_P, _I, ctypes.c_double, _I, _D]),
    "astr_qmax": (ctypes.c_int, [_P, _D]),
    "astr_adjoint_reverse": (ctypes.c_int, [_P, _D, _D, _D]),
    "astr_boucon": (ctypes.c_int, [_P]),
    "astr_set_time": (ctypes.c_int, [_P, ctypes.c_double]),
    "astr_get_time": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "astr_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), _I, ctypes.POINTER(_P)]),
    "astr_batch_destroy": (ctypes.c_int, [_P]),
    "astr_batch_rk3_steps": (ctypes.c_int, [_P, _I, ctypes.c_double]),
    "astr_batch_set_mode": (ctypes.c_int, [_P, _I]),
    "astr_batch_autotune": (ctypes.c_int, [_P, ctypes.c_double]),
    "astr_batch_get_tuning": (ctypes.c_int, [_P, _D, _I, ctypes.POINTER(ctypes.c_int32)]),
    "astr_batch_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "astr_bench_copy": (ctypes.c_int, [_P, _I, _D, ctypes.POINTER(ctypes.c_int64)]),
    "astr_get_timing": (ctypes.c_int, [_P, _D]),
    "astr_last_error": (ctypes.c_char_p, []),
}

_lib = None


class AstrError(RuntimeError):
    pass


def load(path: str = LIBPATH) -> ctypes.CDLL:
    """Load the HIP library (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AstrError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != ASTR_OK:
        msg = load().astr_last_error()
        raise AstrError(f"{what} failed: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")


def dptr(a):
    """double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_D)
