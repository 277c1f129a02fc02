"""ctypes binding of libmicpmd.so (the C ABI declared in include/micpmd.h).

The shared library is built in-tree by ``cpmd_amd.build``; loading fails loudly
when it is missing -- there is no CPU fallback behind these entry points.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmicpmd.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

MC_OK = 0
ERRORS = {1: "MC_ERR_ARG", 2: "MC_ERR_HIP", 3: "MC_ERR_NOTPOSDEF", 4: "MC_ERR_NOCONV",
          5: "MC_ERR_UNSUPPORTED", 6: "MC_ERR_INTERNAL"}

KID = {"zinv": 0, "yinv": 1, "xrho": 2, "xvpsi": 3, "yfwd": 4, "zfwd": 5, "misc": 6, "gemm": 7, "vofrho": 8,
       "gemm_tn": 9, "tn_planes": 10}


class McError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class mc_layout(C.Structure):
    _fields_ = [("nr1", C.c_int), ("nr2", C.c_int), ("nr3", C.c_int),
                ("kr1", C.c_int), ("kr2", C.c_int), ("kr3", C.c_int),
                ("ngw", C.c_int), ("inyh", _ip), ("hg", _dp),
                ("tpiba2", C.c_double), ("omega", C.c_double), ("geq0", C.c_int)]


class mc_restart_header(C.Structure):
    _fields_ = [("ibrav", C.c_int), ("indpg", C.c_int), ("celldm", C.c_double * 6),
                ("nsp", C.c_int), ("nat", C.c_int), ("ecut", C.c_double), ("cdual", C.c_double),
                ("dual", C.c_int), ("nel", C.c_int), ("nr1s", C.c_int), ("nr2s", C.c_int), ("nr3s", C.c_int),
                ("nstate", C.c_int), ("ngwks", C.c_int), ("ngwls", C.c_int), ("nhgs", C.c_int),
                ("nhgls", C.c_int), ("sections", C.c_int)]


_rhp = C.POINTER(mc_restart_header)

# name -> (restype, argtypes)
SIGNATURES = {
    "mc_last_error": (C.c_char_p, []),
    "mc_version": (C.c_int, []),
    "mc_device_count": (C.c_int, [_ip]),
    "mc_layout_stats": (C.c_int, [C.POINTER(mc_layout), _ip, _ip, _ip]),
    "mc_ctx_create": (C.c_int, [C.POINTER(mc_layout), C.POINTER(_vp)]),
    "mc_ctx_destroy": (C.c_int, [_vp]),
    "mc_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "mc_ctx_synchronize": (C.c_int, [_vp]),
    "mc_ctx_set_batch": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mc_ctx_set_column_reuse": (C.c_int, [_vp, C.c_int]),
    "mc_wfn_ld": (C.c_size_t, [_vp]),
    "mc_set_precision": (C.c_int, [_vp, C.c_int]),
    "mc_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(_vp)]),
    "mc_dev_free": (C.c_int, [_vp]),
    "mc_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_size_t]),
    "mc_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_size_t]),
    "mc_memset_d": (C.c_int, [_vp, C.c_int, C.c_size_t]),
    "mc_set_mem_limit": (C.c_int, [C.c_size_t]),
    "mc_mem_stats": (C.c_int, [C.POINTER(C.c_size_t)] * 3),
    "mc_mem_reset_peak": (C.c_int, []),
    "mc_probe_mfma": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "mc_wfn_zero": (C.c_int, [_vp, _vp, C.c_int]),
    "mc_wfn_axpy": (C.c_int, [_vp, C.c_double, _vp, _vp, C.c_int]),
    "mc_wfn_axpby": (C.c_int, [_vp, C.c_double, _vp, C.c_double, _vp, C.c_int]),
    "mc_cp_posupa": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double]),
    "mc_cp_velocity": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double]),
    "mc_posupa_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                C.POINTER(C.c_int)]),
    "mc_rortv": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp, C.c_int]),
    "mc_rortv_dev": (C.c_int, [_vp, _vp, _vp, _dp, C.c_int]),
    "mc_rekine_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_double, _dp]),
    "mc_cp_velupa_rortv_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, _dp, _dp]),
    "mc_hesele": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double]),
    "mc_csize_dev": (C.c_int, [_vp, _vp, _dp, C.c_int, _dp, _dp]),
    "mc_sd_update_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double]),
    "mc_odiis_dev": (C.c_int, [_vp, _vp, _vp, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                               C.POINTER(C.c_int)]),
    "mc_odiis_release": (C.c_int, [_vp]),
    "mc_wfn_dotp": (C.c_int, [_vp, _vp, _vp, C.c_int, _dp]),
    "mc_wfn_upload": (C.c_int, [_vp, _dp, C.c_int, C.c_int, _vp]),
    "mc_wfn_download": (C.c_int, [_vp, _vp, C.c_int, _dp, C.c_int]),
    "mc_rhoofr": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp]),
    "mc_rhoofr_dev": (C.c_int, [_vp, _vp, _dp, C.c_int, _vp, _dp, _dp, _dp]),
    "mc_rhoofr_raw_charge": (C.c_int, [_vp, _dp, _dp]),
    "mc_vpsi": (C.c_int, [_vp, _dp, C.c_int, _dp, _dp, _dp, C.c_int]),
    "mc_vpsi_dev": (C.c_int, [_vp, _vp, _vp, _dp, _vp, C.c_int]),
    "mc_vpsi_dev_noalloc": (C.c_int, [_vp, _vp, _vp, _dp, _vp, C.c_int]),
    "mc_ovlap": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, C.c_int]),
    "mc_ovlap_dev": (C.c_int, [_vp, C.c_int, _dp, _vp, _vp]),
    "mc_rgs": (C.c_int, [_vp, _dp, C.c_int, C.c_int, _dp]),
    "mc_rgs_dev": (C.c_int, [_vp, _vp, C.c_int, _dp]),
    "mc_lowdin": (C.c_int, [_vp, _dp, C.c_int, C.c_int]),
    "mc_lowdin_dev": (C.c_int, [_vp, _vp, C.c_int, _ip]),
    "mc_rotate": (C.c_int, [_vp, C.c_double, _dp, C.c_double, _dp, C.c_int, _dp, C.c_int]),
    "mc_rotate_dev": (C.c_int, [_vp, C.c_double, _vp, C.c_double, _vp, _dp, C.c_int]),
    "mc_lagrange": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp, C.c_int]),
    "mc_lagrange_dev": (C.c_int, [_vp, _vp, _vp, _dp, C.c_int]),
    "mc_rortog": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_double, C.c_int, _ip]),
    "mc_rortog_dev": (C.c_int, [_vp, _vp, _vp, _dp, C.c_int, C.c_int, C.c_double, C.c_int, _ip]),
    "mc_set_projectors": (C.c_int, [_vp, C.c_int, _ip, C.c_int, _ip, _ip, _dp, _dp, C.c_int]),
    "mc_set_eigr": (C.c_int, [_vp, _dp, C.c_int, C.c_int]),
    "mc_rnlsm1": (C.c_int, [_vp, _dp, C.c_int, C.c_int, _dp]),
    "mc_rnlsm1_dev": (C.c_int, [_vp, _vp, C.c_int, _dp]),
    "mc_set_gk": (C.c_int, [_vp, _dp]),
    "mc_rnlsm2": (C.c_int, [_vp, _dp, C.c_int, C.c_int, _dp]),
    "mc_rnlsm2_dev": (C.c_int, [_vp, _vp, C.c_int, _dp]),
    "mc_rnlsm_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _dp, _dp]),
    "mc_rnlfor": (C.c_int, [_vp, _dp, C.c_int, _dp, _dp, _dp]),
    "mc_rnlrh": (C.c_int, [_vp, _dp, C.c_int, _dp]),
    "mc_set_density_gset": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp]),
    "mc_set_local_pp": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _dp, _dp]),
    "mc_set_ions": (C.c_int, [_vp, _dp, _dp]),
    "mc_vofrho": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "mc_vofrho_dev": (C.c_int, [_vp, _vp, _vp, _dp, _dp]),
    "mc_ewald": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "mc_phfac": (C.c_int, [_vp]),
    "mc_fnonloc": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp]),
    "mc_fnonloc_dev": (C.c_int, [_vp, _vp, _dp, C.c_int, _dp]),
    "mc_forces_c2_dev": (C.c_int, [_vp, _vp, _vp, _dp, _vp, C.c_int]),
    "mc_restart_write": (C.c_int, [C.c_char_p, _rhp, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]),
    "mc_restart_read_header": (C.c_int, [C.c_char_p, _rhp]),
    "mc_restart_read": (C.c_int, [C.c_char_p, _rhp, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]),
    "mc_restart_write_dev": (C.c_int, [_vp, C.c_char_p, _rhp, _ip, _dp, _dp, _dp, _dp, _vp, _vp]),
    "mc_restart_read_dev": (C.c_int, [_vp, C.c_char_p, _rhp, _ip, _dp, _dp, _dp, _dp, _vp, _vp]),
    "mc_timer_enable": (C.c_int, [_vp, C.c_int]),
    "mc_timer_reset": (C.c_int, [_vp]),
    "mc_timer_select": (C.c_int, [_vp, C.c_ulonglong]),
    "mc_timer_get": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_longlong)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m cpmd_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != MC_OK:
        raise McError(status, lib().mc_last_error().decode())
