"""ctypes binding of libtbsla_hip.so (the C ABI in include/tbsla_hip.h and
include/tbsla_cpp_c.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing, importing the binding raises, naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtbsla_hip.so")

COO, SCOO, CSR, ELL, DENSE = range(5)
F64, F32, BF16, PATTERN = range(4)
PR_PATTERN_F32 = 4  # PageRank only: value-free FP32 deltas on an FP64 anchor


class TbslaError(RuntimeError):
    """A libtbsla_hip call returned an error status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class PageRankParams(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("beta", C.c_double), ("max_iteration", C.c_int),
                ("precision", C.c_int), ("rule", C.c_int), ("check_every", C.c_int),
                ("anchor_tau", C.c_double), ("nodes", C.POINTER(C.c_int)), ("n_nodes", C.c_int),
                ("gather", C.c_int)]


class PageRankReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("anchors", C.c_int), ("last_error", C.c_double),
                ("bound", C.c_double), ("seconds", C.c_double)]


class AdjointParams(C.Structure):
    _fields_ = [("beta", C.c_double), ("epsilon", C.c_double), ("max_iteration", C.c_int),
                ("precision", C.c_int), ("check_every", C.c_int)]


class AdjointReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("last_change", C.c_double), ("s", C.c_double),
                ("seconds", C.c_double)]


class BatchParams(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("beta", C.c_double), ("max_iteration", C.c_int), ("rule", C.c_int),
                ("check_every", C.c_int), ("graph_capture", C.c_int), ("autotune", C.c_int)]


class BatchReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("seconds", C.c_double), ("tune_seconds", C.c_double),
                ("tune_iterations", C.c_int), ("tune_cached", C.c_int),
                ("bytes_per_iteration", C.c_double), ("wg_per_cu", C.c_int), ("unroll", C.c_int),
                ("xcd_map", C.c_int), ("tile_nnz", C.c_int), ("spmm_rows", C.c_int)]


# Every entry point the headers declare, with its ctypes signature.
_vp, _i, _u, _d, _sz, _l = C.c_void_p, C.c_int, C.c_uint, C.c_double, C.c_size_t, C.c_long
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
SIGNATURES = {
    "tbsla_abi_version": (_i, []),
    "tbsla_last_error": (C.c_char_p, []),
    "tbsla_matrix_create": (_vp, [_i]),
    "tbsla_matrix_destroy": (None, [_vp]),
    "tbsla_matrix_fill_cdiag": (_i, [_vp, _i, _i, _i, _i, _i, _i, _i]),
    "tbsla_matrix_fill_cqmat": (_i, [_vp, _i, _i, _i, _d, _u, _i, _i, _i, _i, C.c_uint64]),
    "tbsla_matrix_fill_cqmat_gpu": (_i, [_vp, _i, _i, _i, _d, _u, _i, _i, _i, _i, C.c_uint64, _i]),
    "tbsla_matrix_fill_cqmat_stochastic": (_i, [_vp, _i, _i, _i, _d, _u, _i, _i, _i, _i]),
    "tbsla_matrix_normalize_columns": (_i, [_vp]),
    "tbsla_matrix_convert": (_vp, [_vp, _i]),
    "tbsla_matrix_convert_gpu": (_vp, [_vp, _i]),
    "tbsla_matrix_from_arrays": (_vp, [_i, _i, _i, _dp, _sz, _ip, _sz, _ip, _sz]),
    "tbsla_matrix_info": (_i, [_vp, C.POINTER(_l)]),
    "tbsla_matrix_array": (_i, [_vp, _i, C.POINTER(_vp), C.POINTER(_sz)]),
    "tbsla_matrix_write": (_i, [_vp, C.c_char_p]),
    "tbsla_matrix_read": (_i, [_vp, C.c_char_p, _l]),
    "tbsla_matrix_read_mm": (_i, [_vp, C.c_char_p]),
    "tbsla_matrix_count_stochastic": (_i, [_vp]),
    "tbsla_matrix_spmv": (_i, [_vp, _dp, _sz, _i, _dp, _sz]),
    "tbsla_matrix_Ax": (_i, [_vp, _dp, _sz, _dp, _sz, _i]),
    "tbsla_matrix_a_axpx": (_i, [_vp, _dp, _sz, _dp, _sz]),
    "tbsla_matrix_page_rank": (_i, [_vp, _d, _d, _i, _dp, _sz, _ip]),
    "tbsla_matrix_personalized_page_rank": (_i, [_vp, _d, _d, _i, _ip, _i, _dp, _sz, _ip]),
    "tbsla_hip_prepare": (_i, [_vp, _i]),
    "tbsla_hip_mirror_info": (_i, [_vp, _i, C.POINTER(_l)]),
    "tbsla_hip_pb_layout_check": (_i, [_vp]),
    "tbsla_hip_Ax": (_i, [_vp, _i, _vp, _sz, _i, _vp, _sz, _i, _i, _vp]),
    "tbsla_hip_fill": (_i, [_vp, _i, _sz, _d, _vp]),
    "tbsla_hip_page_rank": (_i, [_vp, C.POINTER(PageRankParams), _dp, _sz, _vp, C.POINTER(PageRankReport)]),
    "tbsla_hip_page_rank_ex": (_i, [_vp, C.POINTER(PageRankParams), _vp, _sz, _vp, _sz, _vp, _sz, _vp,
                                    C.POINTER(PageRankReport)]),
    "tbsla_hip_page_rank_noalloc": (_i, [_vp, C.POINTER(PageRankParams), _dp, _sz, _vp, C.POINTER(PageRankReport)]),
    "tbsla_hip_page_rank_adjoint": (_i, [_vp, C.POINTER(AdjointParams), _vp, _sz, _vp, _sz, _vp, _vp, _sz, _vp,
                                         _vp, _vp, C.POINTER(AdjointReport)]),
    "tbsla_batch_create": (_vp, [_i]),
    "tbsla_batch_destroy": (None, [_vp]),
    "tbsla_batch_add_graph": (_i, [_vp, _vp, _ip, _ip, _i]),
    "tbsla_batch_prepare": (_i, [_vp]),
    "tbsla_batch_n_rhs": (_i, [_vp]),
    "tbsla_batch_layout": (_i, [_vp, C.POINTER(_l), _sz]),
    "tbsla_batch_run": (_i, [_vp, C.POINTER(BatchParams), _vp, _sz, _ip, _sz, _vp, C.POINTER(BatchReport)]),
    # reference C ABI (interface/c): declared for the export check
    "C_CPP_Vector_create": (_vp, []),
    "C_CPP_Vector_destroy": (None, [_vp]),
    "C_CPP_Vector_set": (None, [_vp, _vp]),
    "C_CPP_Vector_copy": (None, [_vp, _vp]),
    "C_CPP_Vector_print": (None, [_vp]),
    "C_CPP_Vector_fill": (None, [_vp, _i, _i]),
    "C_CPP_Vector_read": (C.c_bool, [_vp, C.c_char_p, _i]),
    "C_CPP_Vector_write": (C.c_bool, [_vp, C.c_char_p]),
    "C_CPP_Vector_add": (C.c_bool, [_vp, _vp]),
    "C_CPP_Vector_add_incr": (C.c_bool, [_vp, _vp, _i]),
    "C_CPP_Vector_gather": (C.c_bool, [_vp, _vp]),
}
for _F in ("COO", "SCOO", "CSR", "ELL", "DENSE"):
    SIGNATURES.update({
        f"C_CPP_Matrix{_F}_create": (_vp, []),
        f"C_CPP_Matrix{_F}_destroy": (None, [_vp]),
        f"C_CPP_Matrix{_F}_fill_cdiag": (None, [_vp, _i, _i, _i, _i, _i, _i, _i]),
        f"C_CPP_Matrix{_F}_fill_cqmat": (None, [_vp, _i, _i, _i, _d, _u, _i, _i, _i, _i]),
        f"C_CPP_Matrix{_F}_read": (C.c_bool, [_vp, C.c_char_p, _i]),
        f"C_CPP_Matrix{_F}_write": (C.c_bool, [_vp, C.c_char_p]),
        f"C_CPP_Matrix{_F}_print": (None, [_vp]),
        f"C_CPP_Matrix{_F}_spmv": (_vp, [_vp, _vp]),
    })


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tbsla_amd.build` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc: int) -> None:
    if rc != 0:
        raise TbslaError(rc, lib.tbsla_last_error().decode(errors="replace"))


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)
