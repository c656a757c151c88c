"""ctypes binding of libbecgpu.so (declared in include/becgpu.h).

The product path has no CPU fallback: if the HIP library is missing or fails
to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libbecgpu.so")

BEC_OK, BEC_EINVAL, BEC_EDEVICE, BEC_ENOMEM, BEC_EUNSUP = 0, -1, -2, -3, -4
BEC_FP64, BEC_FP32 = 0, 1
(E_KIN, E_TRAP, E_ZEEMAN, E_C0, E_C1, E_DDI, E_LHY, E_TENSOR, E_NORM, E_MZ, E_RAMAN,
 E_TOTAL, E_COUNT) = range(13)
E_NAMES = ["kin", "trap", "zeeman", "c0", "c1", "ddi", "lhy", "tensor", "norm", "mz", "raman",
           "total"]
CHUNK_E, CHUNK_DPSI, CHUNK_NORM, CHUNK_PHIMAX, CHUNK_STATUS, CHUNK_COUNT = range(6)
K_NAMES = ["xpass", "fwd", "kin", "inv", "ke", "ddi_fft", "ddi_conv", "reduce", "sites", "energy",
           "xsite"]


class BecConfig(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("n", C.c_int * 3), ("L", C.c_double * 3), ("F", C.c_int),
        ("n_instances", C.c_int), ("precision", C.c_int), ("imaginary_time", C.c_int),
        ("normalize_every", C.c_int), ("enable_ddi", C.c_int), ("ddi_secular", C.c_int),
        ("ddi_padding", C.c_int), ("tensor", C.c_int),
        ("x_tab", C.POINTER(C.c_double)), ("k_tab", C.POINTER(C.c_double)),
    ]


class BecInstParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("c0", C.c_double), ("c1", C.c_double), ("c_lhy", C.c_double),
        ("c2", C.c_double), ("c_dd", C.c_double), ("p", C.c_double), ("q", C.c_double), ("omega", C.c_double * 3),
    ]


class BecError(RuntimeError):
    pass


_lib = None

# name -> (argtypes, restype); every symbol include/becgpu.h declares
SIGNATURES = {
    "bec_ws_create": ([C.POINTER(BecConfig), C.POINTER(C.c_void_p)], C.c_int),
    "bec_ws_destroy": ([C.c_void_p], None),
    "bec_set_params": ([C.c_void_p, C.c_int, C.POINTER(BecInstParams)], C.c_int),
    "bec_set_tensor": ([C.c_void_p, C.c_int, C.c_void_p, C.c_int], C.c_int),
    "bec_set_potential": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bec_set_psi": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bec_get_psi": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bec_set_active": ([C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "bec_split_step": ([C.c_void_p, C.c_int], C.c_int),
    "bec_leapfrog_segment": ([C.c_void_p, C.c_int], C.c_int),
    "bec_half_potential": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "bec_kinetic": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "bec_strang_core": ([C.c_void_p, C.c_double], C.c_int),
    "bec_yoshida_core": ([C.c_void_p, C.c_double], C.c_int),
    "bec_strang_core_from": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "bec_yoshida_core_from": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "bec_normalize": ([C.c_void_p], C.c_int),
    "bec_energy": ([C.c_void_p, C.c_void_p], C.c_int),
    "bec_copy_slot": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "bec_swap_slots": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "bec_diff": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p], C.c_int),
    "bec_itp_guard": ([C.c_void_p, C.c_void_p], C.c_int),
    "bec_spin_density": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bec_itp_chunk": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bec_profile_enable": ([C.c_void_p, C.c_int], C.c_int),
    "bec_profile_read": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "bec_profile_work": ([C.c_void_p, C.c_void_p], C.c_int),
    "bec_set_loss": ([C.c_void_p, C.c_int, C.c_void_p, C.c_int], C.c_int),
    "bec_set_raman": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bec_loss_step": ([C.c_void_p, C.c_double], C.c_int),
    "bec_normalize_constrained": ([C.c_void_p, C.c_void_p], C.c_int),
    "bec_synchronize": ([C.c_void_p], C.c_int),
    "bec_line_plan": ([C.c_int], C.c_int),
    "bec_set_xsite": ([C.c_void_p, C.c_int], C.c_int),
    "bec_query": ([C.c_void_p], C.c_int),
    "bec_autotune": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "bec_split_step_graph": ([C.c_void_p, C.c_int], C.c_int),
    "bec_fft_derivative": ([C.POINTER(BecConfig), C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bec_fft_gradient": ([C.POINTER(BecConfig), C.c_void_p, C.c_void_p], C.c_int),
    "bec_fft_gradient_noalloc": ([C.POINTER(BecConfig), C.c_void_p, C.c_void_p], C.c_int),
    "bec_group_create": ([C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "bec_group_split_step": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "bec_group_grouped": ([C.c_void_p], C.c_int),
    "bec_group_destroy": ([C.c_void_p], None),
    "bec_last_error": ([], C.c_char_p),
    "bec_workspace_bytes": ([C.POINTER(BecConfig)], C.c_longlong),
}


def lib():
    """Load the HIP library (once).  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BecError(f"{LIB_PATH} not found: build it with becgpu.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(code: int, what: str = ""):
    if code != BEC_OK:
        msg = lib().bec_last_error().decode(errors="replace")
        if code in (BEC_EINVAL,):
            raise ValueError(f"{what}: {msg}")
        if code == BEC_EUNSUP:
            raise NotImplementedError(f"{what}: {msg}")
        if code == BEC_ENOMEM:
            raise MemoryError(f"{what}: {msg}")
        raise BecError(f"{what} failed ({code}): {msg}")
