"""ctypes binding of libscalit_hip.so (declarations: include/scalit.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscalit_hip.so")

SCALIT_OK = 0
SCALIT_EINVAL = -1001
SCALIT_EHIP = -1002
SCALIT_ENOMEM = -1003
SCALIT_ESTATE = -1004
SCALIT_EIO = -1005
TA0, TA1 = 0, 1
TAAP0, TMAP0 = 2, 3
TAAP, TAAPP, TAAPR = 100, 101, 102
TMAP, TMAPP, TMAPR = 200, 201, 202
AP, APP, APR = 0, 1, 2
TOSB, TOSBD1, TOSBD2, TOSBW = 0, 1, 2, 3
FP64, FP32, BF16 = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_l = C.c_long
_i = C.c_int
_d = C.c_double

FOP = C.CFUNCTYPE(None, _ip, _dp, _dp)


class QmrJob(C.Structure):
    """scalit_qmr_job of include/scalit.h."""
    _fields_ = [("ctx", C.c_void_p), ("nvec", C.c_int), ("b", C.c_void_p), ("x", C.c_void_p), ("ld", C.c_long),
                ("shifts", _dp), ("ret", _ip), ("res", _dp), ("status", C.c_int)]


class PistJob(C.Structure):
    """scalit_pist_job of include/scalit.h."""
    _fields_ = [("ctx", C.c_void_p), ("nwin", C.c_int), ("e0", _dp), ("x", C.c_void_p), ("ld", C.c_long),
                ("eig", _dp), ("ret", _ip), ("res", _dp), ("stats", C.c_int * 4), ("status", C.c_int)]


FLINSOLV = C.CFUNCTYPE(C.c_int, _ip, _dp, _dp, _dp)

# name -> (restype, argtypes)
SIGNATURES = {
    "scalit_create": (_i, [C.POINTER(_vp), _i, _ip, _ip, _i]),
    "scalit_destroy": (_i, [_vp]),
    "scalit_get_len": (_i, [_vp, C.POINTER(_l)]),
    "scalit_set_h0": (_i, [_vp, _i, _dp]),
    "scalit_set_dep": (_i, [_vp, _i, _dp]),
    "scalit_set_res": (_i, [_vp, _dp]),
    "scalit_set_outh": (_i, [_vp, _dp]),
    "scalit_set_shift": (_i, [_vp, _d, _d, _d]),
    "scalit_block_diag": (_i, [_vp, _i, _d]),
    "scalit_set_precond": (_i, [_vp, C.POINTER(_dp), _dp]),
    "scalit_get_precond": (_i, [_vp, C.POINTER(_dp), _dp]),
    "scalit_hx": (_i, [_vp, _i, _i, _vp, _l, _vp, _l, _vp, _vp]),
    "scalit_hx_lp": (_i, [_vp, _i, _i, _i, _vp, _l, _vp, _l, _vp, _vp]),
    "scalit_hx_adjoint": (_i, [_vp, _i, _i, _vp, _vp, _l, _vp, _vp, _vp]),
    "scalit_px": (_i, [_vp, _i, _i, _i, _vp, _l, _vp, _l, _vp, _vp]),
    "scalit_px_lp": (_i, [_vp, _i, _i, _i, _vp, _l, _vp, _l, _vp, _vp]),
    "scalit_qmr": (_i, [_vp, _i, _i, _d, _i, _vp, _vp, _l, _dp, _ip, _dp, _vp]),
    "scalit_qmr_mixed": (_i, [_vp, _i, _i, _d, _i, _d, _i, _vp, _vp, _l, _dp, _ip, _dp, _vp]),
    "scalit_pist_conv": (_i, [_vp, _d, _d, _i, _vp, _i, _i, _i, _i, _i, _i, _i, _d, _i, _dp, _dp, _ip]),
    "scalit_pist_conv_batched": (_i, [_vp, _i, _dp, _d, _i, _vp, _l, _i, _i, _i, _i, _i, _i, _i, _d, _i, _dp, _ip,
                                      _dp, _ip]),
    "scalit_pist_vectors": (_i, [_vp, _i, _vp, _l]),
    "scalit_pist_vectors_all": (_i, [_vp, _vp, _l, _i]),
    "scalit_lan_conv": (_i, [_vp, _d, _d, _i, _vp, _i, _i, _i, _i, _i, _dp, _dp]),
    "scalit_eig_gradients": (_i, [_vp, _i, _vp, _l, _i, _vp, _l, _dp]),
    "scalit_eig_gradients_noalloc": (_i, [_vp, _i, _vp, _l, _i, _vp, _l, _dp]),
    "scalit_podvr": (_i, [_i, _dp, _dp, _dp, _i, _dp, _dp, _dp]),
    "scalit_set_pist_orth": (_i, [_vp, _i]),
    "scalit_set_osb": (_i, [_vp, _i]),
    "scalit_set_osbw": (_i, [_vp, _i, _d, _d, _i, _vp, _vp]),
    "scalit_set_ap": (_i, [_vp, _i, _vp]),
    "scalit_hx_cx": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "scalit_px_cx": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "scalit_qmr_cx": (_i, [_vp, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp]),
    "scalit_pist_conv_sx": (_i, [_vp, _d, _d, _i, _vp, _i, _i, _i, _i, _i, _i, _i, _d, _i, _vp, _vp, _vp, _vp]),
    "scalit_small_eig_cx": (_i, [_i, _vp, _i, _vp]),
    "scalit_set_h0_cx": (_i, [_vp, _i, _vp, _vp]),
    "scalit_set_dep_cx": (_i, [_vp, _i, _vp, _vp]),
    "scalit_set_outh_cx": (_i, [_vp, _vp, _vp]),
    "scalit_set_scx": (_i, [_vp, _i]),
    "scalit_pist_conv_cx": (_i, [_vp, _d, _d, _i, _vp, _i, _i, _i, _i, _i, _i, _i, _d, _i, _vp, _vp, _vp, _vp]),
    "scalit_crp": (_i, [_vp, _i, _vp, _i, _vp, _i, _i, _d, _i, _vp, _vp]),
    "scalit_qmr_multi": (_i, [_i, _vp, _i, _i, _d]),
    "scalit_set_qmr_persistent": (_i, [_vp, _i]),
    "scalit_load_data_grid": (_i, [C.c_char_p, _l, _i, _dp]),
    "scalit_save_data_grid": (_i, [C.c_char_p, _l, _i, _dp]),
    "scalit_combine_2d": (_i, [_i, _dp, _dp, _i, _dp, _dp, _i, _ip, _ip, _dp]),
    "scalit_bind": (_i, [_vp]),
    "scalit_set_qmr_params": (None, [_i, _i, _d]),
    "hx_": (None, [_ip, _dp, _dp]),
    "hx0_": (None, [_ip, _dp, _dp]),
    "ehx_": (None, [_ip, _dp, _dp]),
    "px_": (None, [_ip, _dp, _dp]),
    "epx_": (None, [_ip, _dp, _dp]),
    "osb_qmr_": (_i, [_ip, _dp, _dp, _dp]),
    "qmr_": (_i, [_ip, _ip, _dp, _ip, _dp, FOP, FOP, _dp, _dp]),
    "pist_conv_": (_i, [_dp, _dp, _ip, _ip, _dp, _ip, _ip, _ip, _ip, _ip, FOP, FLINSOLV, _dp, _dp]),
    "lan_conv_": (_i, [_dp, _dp, _ip, _ip, _dp, _ip, _ip, _ip, _ip, _ip, FOP, _dp, _dp]),
    "mgs_orth_": (_i, [_ip, _dp, _ip, _dp]),
    "scalit_debug_force_generic": (_i, [_i]),
    "scalit_profile_hx_pass": (_i, [_vp, _i, _i, _vp, _l, _vp, _l, _vp]),
    "scalit_debug_mfma_probe": (_i, [_i, _dp]),
    "scalit_debug_mfma_peak": (_i, [_i, _i, _i, C.POINTER(C.c_float)]),
    "scalit_debug_hx_variant": (_i, [_vp, _i, _i, _ip, _ip]),
    "scalit_debug_sym_eig": (_i, [_i, _dp, _i, _dp, _dp, _i]),
    "scalit_debug_sym_eigvals_batch": (_i, [_i, _ip, _dp, _i, _dp]),
    "scalit_save_data": (_i, [C.c_char_p, _l, _dp, _i]),
    "scalit_load_data": (_i, [C.c_char_p, _l, _dp, _i]),
    "scalit_save_data_dir": (_i, [C.c_char_p, _l, _dp]),
    "scalit_load_data_dir": (_i, [C.c_char_p, _l, _dp]),
    "scalit_layout": (_i, [_i, _ip, _ip, _i, C.POINTER(_l), C.POINTER(_l), C.POINTER(_l), C.POINTER(_l)]),
    "scalit_pist_conv_multi": (_i, [_i, C.c_void_p, _d, _i, _i, _i, _i, _i, _i, _i, _i, _d, _i]),
}

_lib = None


class ScalitError(RuntimeError):
    pass


def lib():
    """The HIP library; raises when it is missing (build with `python -m scalit_amd.build`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ScalitError(f"libscalit_hip.so not built: {LIB_PATH} (run `python -m scalit_amd.build`)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "scalit"):
    if rc != SCALIT_OK:
        names = {SCALIT_EINVAL: "EINVAL", SCALIT_EHIP: "EHIP", SCALIT_ENOMEM: "ENOMEM", SCALIT_ESTATE: "ESTATE",
                 SCALIT_EIO: "EIO"}
        raise ScalitError(f"{what} failed: {names.get(rc, rc)}")
