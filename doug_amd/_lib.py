"""ctypes binding of libdoug_hip.so (the C-ABI in include/doug.h).

The library is the only compute path: if it is missing, or no GPU is
visible, every call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdoug_hip.so")

DOUG_OK = 0
ERRORS = {1: "DOUG_E_ARG", 2: "DOUG_E_STATE", 3: "DOUG_E_HIP", 4: "DOUG_E_NUMERIC",
          5: "DOUG_E_NOMEM", 6: "DOUG_E_UNSUPPORTED"}

_i32 = C.c_int32
_i64 = C.c_int64
_dbl = C.c_double
_vp = C.c_void_p


class DougParams(C.Structure):
    """doug_params: DOUG.ctl keys on the path (-1 = not set)."""

    _fields_ = [
        ("radius1", _i32), ("radius2", _i32), ("minasize1", _i32), ("minasize2", _i32),
        ("maxasize1", _i32), ("maxasize2", _i32),
        ("strong1", _dbl), ("strong2", _dbl),
        ("smoothers", _i32), ("overlap", _i32), ("levels", _i32), ("method", _i32),
        ("fine_method", _i32), ("symmstruct", _i32), ("symmnumeric", _i32),
        ("omega", _dbl),
        ("num_iters", _i32), ("reserved", _i32),
    ]


class DougInfo(C.Structure):
    _fields_ = [
        ("n", _i32), ("nagr_fine", _i32), ("nagr_coarse", _i32), ("nsub", _i32),
        ("ncoarse", _i32), ("overlap", _i32), ("smoothers", _i32),
        ("nnz", _i64), ("sub_total", _i64), ("factor_bytes", _i64), ("coarse_factor_bytes", _i64),
        ("t_aggregate", _dbl), ("t_coarse_build", _dbl), ("t_subdomains", _dbl),
        ("t_upload", _dbl), ("t_factor", _dbl), ("t_coarse_factor", _dbl), ("t_total", _dbl),
        ("factor_tiles", _i64), ("factor_flops", _dbl), ("temp_bytes", _i64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DougTuning(C.Structure):
    _fields_ = [
        ("spmv_shape", _i32), ("spmv_rows", _i32), ("spmv_xcd_order", _i32), ("trsv_order", _i32),
        ("spmv_ms", _dbl * 6), ("trsv_ms", _dbl * 2),
        ("trsm_tiles_per_step", _i32), ("trsm_col_groups", _i32), ("trsm_ms", _dbl * 2),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["spmv_ms"] = list(self.spmv_ms)
        d["trsv_ms"] = list(self.trsv_ms)
        d["trsm_ms"] = list(self.trsm_ms)
        return d


# every symbol include/doug.h declares, with its ctypes signature
SIGNATURES = {
    "doug_ctx_create": (_i32, [C.c_int, C.POINTER(_vp)]),
    "doug_ctx_destroy": (_i32, [_vp]),
    "doug_last_error": (C.c_char_p, [_vp]),
    "doug_params_default": (None, [C.POINTER(DougParams)]),
    "doug_set_params": (_i32, [_vp, C.POINTER(DougParams)]),
    "doug_set_matrix_coo": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _i32]),
    "doug_set_partition": (_i32, [_vp, _i32, _vp, _vp, _i32]),
    "doug_set_num_subdomains": (_i32, [_vp, _i32]),
    "doug_set_aggregates": (_i32, [_vp, _i32, _vp]),
    "doug_set_strong_mask": (_i32, [_vp, _i64, _vp]),
    "doug_set_restriction_coo": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _i32]),
    "doug_set_subdomains": (_i32, [_vp, _i32, _vp, _vp, _i32]),
    "doug_set_ranks": (_i32, [_vp, _i32, _vp, _i32]),
    "doug_get_rank_map": (_i32, [_vp, _i32, C.POINTER(_i32), C.POINTER(_i32), _vp]),
    "doug_partition_graph": (_i32, [_i32, _vp, _vp, _vp, _i32, _i32, _vp]),
    "doug_setup": (_i32, [_vp, _i32]),
    "doug_setup_host": (_i32, [_vp]),
    "doug_get_info": (_i32, [_vp, C.POINTER(DougInfo)]),
    "doug_get_fine_aggregates": (_i32, [_vp, _vp, _vp, _vp]),
    "doug_get_coarse_aggregates": (_i32, [_vp, _vp, _vp, _vp]),
    "doug_get_subdomains": (_i32, [_vp, _vp, _vp]),
    "doug_get_restriction": (_i32, [_vp, _vp, _vp, _vp]),
    "doug_get_coarse_matrix": (_i32, [_vp, _vp, _vp, _vp]),
    "doug_restriction_nnz": (_i64, [_vp]),
    "doug_coarse_matrix_nnz": (_i64, [_vp]),
    "doug_spmv": (_i32, [_vp, _vp, _vp]),
    "doug_spmv_device": (_i32, [_vp, _vp, _vp]),
    "doug_prec_apply": (_i32, [_vp, _vp, _vp]),
    "doug_prec_apply_device": (_i32, [_vp, _vp, _vp]),
    "doug_prec_apply_device_noalloc": (_i32, [_vp, _vp, _vp]),
    "doug_pcg": (_i32, [_vp, _vp, _vp, _dbl, _i32, C.POINTER(_i32), C.POINTER(_dbl), _vp]),
    "doug_pcg_device": (_i32, [_vp, _vp, _vp, _dbl, _i32, C.POINTER(_i32), C.POINTER(_dbl), _vp]),
    "doug_time_operator": (_i32, [_vp, _i32, _i32, C.POINTER(_dbl)]),
    "doug_prepare_batch": (_i32, [_vp]),
    "doug_value_grad_device": (_i32, [_vp, _vp, _vp, _vp]),
    "doug_set_option": (_i32, [_vp, _i32, _i64]),
    "doug_autotune": (_i32, [_vp, _i32, C.POINTER(DougTuning)]),
    "doug_get_tuning": (_i32, [_vp, C.POINTER(DougTuning)]),
    "doug_prec_apply_batch_device": (_i32, [_vp, _i32, _vp, _vp]),
    "doug_pcg_batch": (_i32, [_vp, _i32, _vp, _vp, _dbl, _i32, _vp, _vp]),
    "doug_pcg_batch_device": (_i32, [_vp, _i32, _vp, _vp, _dbl, _i32, _vp, _vp]),
}

_lib = None


def load():
    """Load libdoug_hip.so and attach signatures; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
                " or `make -C doug_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)
