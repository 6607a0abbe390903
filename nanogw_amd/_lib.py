"""ctypes binding of libngw.so (the C ABI declared in include/ngw.h).

The library is loaded from the package directory only; there is no fallback:
if it is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libngw.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ngw.h")

NGW_OK = 0
STATUS = {-1: "NGW_ERR_ARG", -2: "NGW_ERR_STATE", -3: "NGW_ERR_HIP", -4: "NGW_ERR_NOMEM",
          -5: "NGW_ERR_RANGE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

_SIGS = {
    "ngw_version": ([], C.c_char_p),
    "ngw_last_error": ([], C.c_char_p),
    "ngw_create": ([C.POINTER(_vp), C.c_int], C.c_int),
    "ngw_destroy": ([_vp], C.c_int),
    "ngw_synchronize": ([_vp], C.c_int),
    "ngw_set_isdf": ([_vp, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp], C.c_int),
    "ngw_set_precision": ([_vp, C.c_int], C.c_int),
    "ngw_set_reorthogonalization": ([_vp, C.c_int], C.c_int),
    "ngw_qp_newton_dev_noalloc": ([_vp, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p], C.c_int),
    "ngw_set_explicit_c": ([_vp, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp], C.c_int),
    "ngw_set_isdf_blocks": ([_vp, C.c_int, C.c_int, _ip, _ip, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp],
                            C.c_int),
    "ngw_matvec_isdf_block": ([_vp, _dp, _dp, C.c_int, C.c_int, C.c_int], C.c_int),
    "ngw_polynomial_matvec_isdf_block": ([_vp, _dp, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int],
                                         C.c_int),
    "ngw_lanczos_spectra_isdf_block": ([_vp, _dp, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_int,
                                        _dp, C.c_int, C.c_int, _dp, _dp], C.c_int),
    "ngw_sigma_lanczos": ([_vp, C.c_int, _ip, C.c_int, C.c_double, C.c_double, C.c_int, _dp, C.c_int,
                           C.c_int, C.c_int, _dp], C.c_int),
    "ngw_sigma_lanczos_bands": ([_vp, C.c_int, _ip, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                 _dp, C.c_int, C.c_int, C.c_int, _dp], C.c_int),
    "ngw_sigma_x_isdf": ([_vp, C.c_int, _ip, _dp], C.c_int),
    "ngw_qp_newton": ([_vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                       _ip], C.c_int),
    "ngw_polyfit_sqrt": ([C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp], C.c_int),
    "ngw_isdf_coulomb_kernel": ([_vp, C.c_int, C.c_int, _dp, C.c_long, _ip, _ip, C.c_double,
                                 C.c_double, C.c_double, _dp, C.c_int], C.c_int),
    "ngw_isdf_coulomb_kernel_dev": ([_vp, C.c_int, C.c_int, _vp, C.c_long, _ip, _ip, C.c_double,
                                     C.c_double, C.c_double, _dp, C.c_int], C.c_int),
    "ngw_isdf_zeta": ([_vp, C.c_int, C.c_int, C.c_int, _dp, C.c_long, C.c_int, _ip, C.c_double, _dp, C.c_long,
                       _ip], C.c_int),
    "ngw_isdf_zeta_dev": ([_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_long, C.c_int, _ip, C.c_double, _vp,
                           C.c_long, _ip], C.c_int),
    "ngw_cvt_points": ([_vp, C.c_int, _ip, _dp, C.c_int, _dp, C.c_double, C.c_double, C.c_int, _ip, _ip],
                       C.c_int),
    "ngw_matvec_isdf_block_dev": ([_vp, _vp, _vp, C.c_long, C.c_int, C.c_int], C.c_int),
    "ngw_profile_enable": ([_vp, C.c_int], C.c_int),
    "ngw_profile_reset": ([_vp], C.c_int),
    "ngw_profile_read": ([_vp, C.c_int, _dp, C.POINTER(C.c_long)], C.c_int),
    "ngw_autotune": ([_vp, C.c_int, C.c_int, _ip, _ip], C.c_int),
    "ngw_set_tile_config": ([_vp, C.c_int, C.c_int], C.c_int),
    "ngw_gemm_lowp_dev": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_long, _vp, C.c_long, _vp,
                           C.c_long], C.c_int),
    "ngw_batch_create": ([_vp, C.POINTER(_vp)], C.c_int),
    "ngw_batch_destroy": ([_vp], C.c_int),
    "ngw_batch_add_system": ([_vp, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _ip], C.c_int),
    "ngw_batch_remove_system": ([_vp, C.c_int], C.c_int),
    "ngw_batch_request_bytes": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_long)], C.c_int),
    "ngw_batch_submit": ([_vp, C.c_int, C.c_int, _ip, C.c_int, C.c_double, C.c_double, C.c_int, _dp, C.c_int,
                          C.c_int, _ip], C.c_int),
    "ngw_batch_step": ([_vp, C.c_int, C.c_int, _ip], C.c_int),
    "ngw_batch_active": ([_vp, _ip], C.c_int),
    "ngw_batch_result": ([_vp, C.c_int, _dp, _ip], C.c_int),
    "ngw_batch_release": ([_vp, C.c_int], C.c_int),
    "ngw_batch_hold": ([_vp, C.c_int, C.c_int], C.c_int),
    "ngw_batch_stats": ([_vp, C.POINTER(C.c_long), C.POINTER(C.c_long)], C.c_int),
    "ngw_gemm_f64_dev": ([_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_long, _vp, C.c_long, _vp, C.c_long],
                         C.c_int),
}

_LIB = None


class NgwError(RuntimeError):
    pass


def load():
    """Load libngw.so from the package directory (raises if absent)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NgwError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def declared_symbols():
    """Function names declared in include/ngw.h."""
    import re

    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ngw_\w+)\s*\(", text, re.M)))


def check(status: int, what: str = ""):
    if status != NGW_OK:
        msg = load().ngw_last_error().decode(errors="replace")
        raise NgwError(f"{what}: {STATUS.get(status, status)}: {msg}")
