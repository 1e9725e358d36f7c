"""ctypes declarations of include/nssdmrg.h (the C ABI of libnssdmrg.so).

The library is loaded from this package directory only.  There is no CPU
fallback: if the HIP library is absent, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnssdmrg.so")

NSS_OK = 0
NSS_FP64, NSS_FP32_MIXED = 0, 1
NSS_ROW_MAJOR, NSS_COL_MAJOR = 0, 1
NSS_MODEL_SPIN12, NSS_MODEL_SPIN1, NSS_MODEL_HUBBARD = 0, 1, 2

ERRORS = {1: "NSS_E_ARG", 2: "NSS_E_HIP", 3: "NSS_E_STATE", 4: "NSS_E_NOCONV",
          5: "NSS_E_NUMERIC", 6: "NSS_E_ALLOC"}


class NssError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Model(C.Structure):
    _fields_ = [("model", C.c_int32), ("hfmode", C.c_int32), ("hij", C.c_double * 4),
                ("uloc", C.c_double), ("hvec", C.c_double * 2), ("eloc", C.c_double * 2)]


class Params(C.Structure):
    _fields_ = [("ldmrg", C.c_int32), ("mdmrg", C.c_int32), ("edmrg", C.c_double),
                ("nsweep", C.c_int32), ("msweep", C.POINTER(C.c_int32)),
                ("esweep", C.POINTER(C.c_double)), ("target_qn", C.c_double * 2),
                ("neigen", C.c_int32), ("lanc_dim_threshold", C.c_int32),
                ("lanc_kmax", C.c_int32), ("lanc_maxiter", C.c_int32), ("lanc_tol", C.c_double),
                ("seed", C.c_uint64), ("v0_mode", C.c_int32), ("deg_thr", C.c_double),
                ("precision", C.c_int32), ("verbose", C.c_int32), ("predict", C.c_int32),
                ("pbc", C.c_int32), ("qn_type", C.c_int32), ("qn_trunc_error", C.c_double),
                ("qn_trunc_dim", C.c_int32)]


class StepInfo(C.Structure):
    _fields_ = [("type", C.c_int32), ("label", C.c_int32), ("sweep", C.c_int32),
                ("left_len", C.c_int32), ("right_len", C.c_int32), ("L", C.c_int32),
                ("neigen", C.c_int32), ("energies", C.c_double * 4),
                ("entropy_left", C.c_double), ("entropy_right", C.c_double),
                ("m_left", C.c_int32), ("m_right", C.c_int32), ("renormalized", C.c_int32),
                ("trunc_left", C.c_double), ("trunc_right", C.c_double),
                ("cut_left", C.c_double), ("cut_right", C.c_double), ("m_sb", C.c_int64), ("nsect", C.c_int32), ("nterms", C.c_int32),
                ("numop", C.c_int32), ("numop_fp32", C.c_int32), ("hv_flops", C.c_double),
                ("t_step", C.c_double), ("t_diag", C.c_double), ("t_hv", C.c_double),
                ("t_rdm", C.c_double), ("t_renorm", C.c_double), ("t_enlarge", C.c_double),
                ("link", C.c_int32), ("sn_left", C.c_double), ("sc_left", C.c_double),
                ("sn_right", C.c_double), ("sc_right", C.c_double)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_DPP = C.POINTER(_D)

SIGNATURES = {
    "nss_last_error": (C.c_char_p, []),
    "nss_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "nss_ctx_destroy": (C.c_int, [_P]),
    "nss_device_ok": (C.c_int, [C.c_int]),
    "nss_sb_set_layout": (C.c_int, [_P, C.c_int, _I32, _I32, _I64]),
    "nss_sb_upload_blocks": (C.c_int, [_P, _DPP, _DPP, C.c_int, _I32, _I32, _DPP, _DPP, C.c_int, C.c_int]),
    "nss_sb_hv": (C.c_int, [_P, _D, _D, C.c_int64]),
    "nss_sb_hv_device": (C.c_int, [_P, _P, _P, C.c_int64]),
    "nss_sb_hv_device_noalloc": (C.c_int, [_P, _P, _P, C.c_int64]),
    "nss_ctx_hv_stats": (C.c_int, [_P, C.c_int, _D]),
    "nss_ctx_timers": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "nss_sb_hv_fp32": (C.c_int, [_P, _D, _D, C.c_int64]),
    "nss_sb_flops": (C.c_int, [_P, _D]),
    "nss_lanczos": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int,
                              _D, _D, C.POINTER(C.c_int)]),
    "nss_rdm_eigh": (C.c_int, [_P, _D, C.c_int, _D, _D]),
    "nss_kept_count": (C.c_int, [_D, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                 C.POINTER(C.c_int), _D]),
    "nss_rotate_blocks": (C.c_int, [_P, C.c_int, _I32, _I32, _I32, _I32, _DPP, _DPP, _DPP, _DPP]),
    "nss_host_sym_eig": (C.c_int, [C.c_int, _D, _D]),
    "nss_probe_mfma": (C.c_int, [C.c_int, _D, _D, _D]),
    "nss_probe_copy": (C.c_int, [C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "nss_gemm_tune_table": (C.c_int, [C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
    "nss_dmrg_measure": (C.c_int, [C.c_void_p, C.c_int, _D, C.c_int, C.POINTER(C.c_int32), _D]),
    "nss_dmrg_load_blocks": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "nss_dmrg_save_blocks": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "nss_host_tridiag_eig": (C.c_int, [C.c_int, _D, _D, _D, _D]),
    "nss_host_tridiag_lowest": (C.c_int, [C.c_int, _D, _D, _D, _D]),
    "nss_dmrg_create": (C.c_int, [_P, C.POINTER(Model), C.POINTER(Params), C.POINTER(_P)]),
    "nss_dmrg_run": (C.c_int, [_P, C.c_int]),
    "nss_dmrg_finite_begin": (C.c_int, [_P]),
    "nss_dmrg_sweep": (C.c_int, [_P, C.c_int]),
    "nss_dmrg_finite_step": (C.c_int, [_P, _I32]),
    "nss_dmrg_batch_run": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.c_int]),
    "nss_dmrg_batch_finite_begin": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "nss_dmrg_batch_sweep": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.c_int]),
    "nss_dmrg_nsteps": (C.c_int, [_P]),
    "nss_dmrg_step_info": (C.c_int, [_P, C.c_int, C.POINTER(StepInfo)]),
    "nss_dmrg_kept_qn": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_int, _D, _I32, _I32]),
    "nss_dmrg_lambda": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _D, _I32, C.c_int64, _D, _I32, _I64]),
    "nss_dmrg_last_sb_dims": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), _I64]),
    "nss_dmrg_last_sb_export": (C.c_int, [_P, _I32, _I32, _I32, _I32, _D, _D, _D, _D, _D]),
    "nss_dmrg_destroy": (C.c_int, [_P]),
}

_lib = None


def lib():
    """Load libnssdmrg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m nssdmrg_amd.build_lib` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != NSS_OK:
        raise NssError(rc, lib().nss_last_error().decode())


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I32)


def lptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_I64)


def ptr_array(arrs):
    """double** from a list of contiguous float64 arrays (kept alive by caller)."""
    out = (_D * max(1, len(arrs)))()
    for i, a in enumerate(arrs):
        out[i] = dptr(a)
    return out
