"""Python mirror of the TBSLA matrix API (tbsla::cpp::Matrix{COO,SCOO,CSR,ELL,DENSE}).

Same method names, argument order and semantics as the reference classes
(src/cpp/include/tbsla/cpp/Matrix.hpp:9-47 and the per-format headers):

* ``fill_cdiag`` / ``fill_cqmat`` / ``fill_cqmat_stochastic`` / ``normalize_columns``
  build the host arrays, bit-identical to the reference generators;
* ``spmv(v)`` returns a zero-initialised vector of length ``n_row``;
  ``Ax(r, v)`` assigns rows for CSR/ELL and accumulates for COO/SCOO/DENSE;
* ``page_rank(epsilon, beta, max_iteration)`` and
  ``personalized_page_rank(epsilon, beta, max_iteration, nodes)`` return
  ``(vector, iterations)`` -- the reference returns the count through an
  ``int&`` argument.

Products run on the GPU through libtbsla_hip.so.  The ``*_device`` methods take
device buffers (torch tensors) and a HIP stream for zero-copy use.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr, lib

_INFO_KEYS = ["format", "n_row", "n_col", "f_row", "f_col", "ln_row", "ln_col", "pr", "pc", "NR", "NC",
              "nnz", "max_col", "n_values", "n_idx0", "n_idx1"]
_DTYPES = {"f64": _lib.F64, "float64": _lib.F64, "f32": _lib.F32, "float32": _lib.F32,
           "pattern_f32": _lib.PR_PATTERN_F32,
           "bf16": _lib.BF16, "bfloat16": _lib.BF16, "pattern": _lib.PATTERN}


def dtype_code(d) -> int:
    if isinstance(d, int):
        return d
    return _DTYPES[str(d).replace("torch.", "")]


def _tensor_ptr(t) -> int:
    return t.data_ptr() if hasattr(t, "data_ptr") else int(t)


def _tensor_dtype(t) -> int:
    import torch
    if t.dtype == torch.float64:
        return _lib.F64
    if t.dtype == torch.float32:
        return _lib.F32
    raise TypeError(f"unsupported vector dtype {t.dtype}")


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


class Matrix:
    FORMAT: int = -1
    NAME = ""

    def __init__(self, _handle=None):
        if _handle is None:
            _handle = lib.tbsla_matrix_create(self.FORMAT)
            if not _handle:
                check(-1)
        self._h = _handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:  # module globals are cleared at interpreter exit
            lib.tbsla_matrix_destroy(h)
            self._h = None

    # ---- generators (Matrix.hpp:37-41) -----------------------------------
    def fill_cdiag(self, n_row, n_col, cdiag, pr=0, pc=0, NR=1, NC=1):
        check(lib.tbsla_matrix_fill_cdiag(self._h, n_row, n_col, cdiag, pr, pc, NR, NC))
        return self

    def fill_cqmat(self, n_row, n_col, c, q, seed_mult=1, pr=0, pc=0, NR=1, NC=1, value_seed=0, device=False):
        """``value_seed`` != 0 applies the project value overlay
        value = 1 - u01(value_seed, i), i the entry's global generation index.
        ``device=True`` generates the entries on the GPU (COO, SCOO, CSR and
        ELL; DENSE is refused); the arrays are the same."""
        if device:
            check(lib.tbsla_matrix_fill_cqmat_gpu(self._h, n_row, n_col, c, float(q), int(seed_mult) & 0xFFFFFFFF,
                                                  pr, pc, NR, NC, int(value_seed), 0))
            return self
        check(lib.tbsla_matrix_fill_cqmat(self._h, n_row, n_col, c, float(q), int(seed_mult) & 0xFFFFFFFF,
                                          pr, pc, NR, NC, int(value_seed)))
        return self

    def fill_cqmat_stochastic(self, n_row, n_col, c, q, seed_mult=1, pr=0, pc=0, NR=1, NC=1, device=False):
        """``device=True`` generates the entries and the column counts on the
        GPU (COO, SCOO, CSR and ELL; DENSE is refused); the arrays are the
        same."""
        if device:
            check(lib.tbsla_matrix_fill_cqmat_gpu(self._h, n_row, n_col, c, float(q), int(seed_mult) & 0xFFFFFFFF,
                                                  pr, pc, NR, NC, 0, 1))
            return self
        check(lib.tbsla_matrix_fill_cqmat_stochastic(self._h, n_row, n_col, c, float(q),
                                                     int(seed_mult) & 0xFFFFFFFF, pr, pc, NR, NC))
        return self

    def normalize_columns(self):
        check(lib.tbsla_matrix_normalize_columns(self._h))
        return self

    # ---- getters -----------------------------------------------------------
    def info(self) -> dict:
        buf = (C.c_long * 16)()
        check(lib.tbsla_matrix_info(self._h, buf))
        return dict(zip(_INFO_KEYS, list(buf)))

    def get_n_row(self): return self.info()["n_row"]
    def get_n_col(self): return self.info()["n_col"]
    def get_f_row(self): return self.info()["f_row"]
    def get_f_col(self): return self.info()["f_col"]
    def get_ln_row(self): return self.info()["ln_row"]
    def get_ln_col(self): return self.info()["ln_col"]
    def get_pr(self): return self.info()["pr"]
    def get_pc(self): return self.info()["pc"]
    def get_NR(self): return self.info()["NR"]
    def get_NC(self): return self.info()["NC"]
    def get_nnz(self): return self.info()["nnz"]

    def count_stochastic(self) -> bool:
        return bool(lib.tbsla_matrix_count_stochastic(self._h))

    def _array(self, which, dtype):
        p = C.c_void_p()
        n = C.c_size_t()
        check(lib.tbsla_matrix_array(self._h, which, C.byref(p), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, dtype=dtype)
        ct = C.c_double if dtype == np.float64 else C.c_int
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n.value,)).copy()

    def arrays(self) -> dict:
        """Copies of the host arrays, keyed like the reference members."""
        out = {"values": self._array(0, np.float64)}
        for which, name in ((1, self._IDX0), (2, self._IDX1)):
            if name:
                out[name] = self._array(which, np.int32)
        return out

    _IDX0: Optional[str] = None
    _IDX1: Optional[str] = None

    # ---- binary I/O (MatrixCSR.cpp:179-232) -----------------------------------
    def write(self, path: str):
        check(lib.tbsla_matrix_write(self._h, path.encode()))

    def read(self, path: str, seek: int = 0):
        check(lib.tbsla_matrix_read(self._h, path.encode(), seek))
        return self

    # ---- products (GPU) -------------------------------------------------------
    def spmv(self, v, vect_incr: int = 0) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.get_n_row(), dtype=np.float64)
        check(lib.tbsla_matrix_spmv(self._h, dptr(v), v.size, vect_incr, dptr(out), out.size))
        return out

    def Ax(self, r, v, vect_incr: int = 0) -> np.ndarray:
        """Returns the updated copy of ``r`` (the reference updates r in place)."""
        r = np.array(r, dtype=np.float64, copy=True)
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(lib.tbsla_matrix_Ax(self._h, dptr(r), r.size, dptr(v), v.size, vect_incr))
        return r

    def a_axpx_(self, v) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.get_n_row(), dtype=np.float64)
        check(lib.tbsla_matrix_a_axpx(self._h, dptr(v), v.size, dptr(out), out.size))
        return out

    def page_rank(self, epsilon, beta, max_iteration):
        out = np.zeros(self.get_n_col(), dtype=np.float64)
        it = C.c_int()
        check(lib.tbsla_matrix_page_rank(self._h, epsilon, beta, max_iteration, dptr(out), out.size,
                                         C.byref(it)))
        return out, it.value

    def personalized_page_rank(self, epsilon, beta, max_iteration, personalized_nodes):
        nodes = np.ascontiguousarray(personalized_nodes, dtype=np.int32)
        out = np.zeros(self.get_n_col(), dtype=np.float64)
        it = C.c_int()
        check(lib.tbsla_matrix_personalized_page_rank(self._h, epsilon, beta, max_iteration, iptr(nodes),
                                                      nodes.size, dptr(out), out.size, C.byref(it)))
        return out, it.value

    # ---- device-resident use ----------------------------------------------------
    def prepare(self, value_dtype="f64"):
        check(lib.tbsla_hip_prepare(self._h, dtype_code(value_dtype)))
        return self

    def mirror_info(self, value_dtype="f64") -> dict:
        buf = (C.c_long * 8)()
        check(lib.tbsla_hip_mirror_info(self._h, dtype_code(value_dtype), buf))
        keys = ["rows", "entries", "x_need", "y_lo", "y_hi", "tiles", "mirror_bytes", "algorithmic_bytes"]
        return dict(zip(keys, list(buf)))

    def Ax_device(self, x, y, vect_incr: int = 0, value_dtype=None, stream=None):
        """y (Ax semantics) from x; x and y are torch tensors on the GPU."""
        vd = dtype_code(value_dtype) if value_dtype is not None else _tensor_dtype(x)
        check(lib.tbsla_hip_Ax(self._h, vd, _tensor_ptr(x), x.numel(), _tensor_dtype(x), _tensor_ptr(y),
                               y.numel(), _tensor_dtype(y), vect_incr, _stream_handle(stream)))
        return y

    def page_rank_device(self, out, epsilon=1e-10, beta=0.85, max_iteration=100, precision="f64",
                         rule="reference", check_every=8, anchor_tau=0.0, nodes=None, stream=None,
                         values=None, teleport=None, gather="auto"):
        """PageRank into the float64 tensor ``out``; returns the run report.

        ``values`` (float64 device tensor, one per stored CSR entry) replaces
        the stored edge values; ``teleport`` (float64 device tensor, n_col)
        is the teleport distribution v of b = beta*A b + (1-beta)(1^T b) v.
        ``gather``: "auto", "tiles" (fixed-order entry tiles, the
        reference's row sums) or "blocked" (propagation-blocked products;
        auto picks them on n >= 2^16)."""
        p = _lib.PageRankParams()
        p.epsilon, p.beta, p.max_iteration = epsilon, beta, max_iteration
        p.precision = dtype_code(precision)
        p.rule = {"reference": 0, "sumnorm": 1}[rule] if isinstance(rule, str) else int(rule)
        p.check_every = check_every
        p.anchor_tau = anchor_tau
        p.gather = {"auto": 0, "tiles": 1, "blocked": 2}[gather] if isinstance(gather, str) else int(gather)
        keep = None
        if nodes is not None:
            keep = np.ascontiguousarray(nodes, dtype=np.int32)
            p.nodes = iptr(keep)
            p.n_nodes = keep.size
        _check_f64(out, "out")
        if out.numel() < self.get_n_col():
            raise ValueError(f"out has {out.numel()} entries, PageRank writes {self.get_n_col()}")
        vals_p, vals_n = _opt_f64(values, "values")
        tele_p, tele_n = _opt_f64(teleport, "teleport")
        rep = _lib.PageRankReport()
        check(lib.tbsla_hip_page_rank_ex(self._h, C.byref(p), vals_p, vals_n, tele_p, tele_n, _tensor_ptr(out),
                                         out.numel(), _stream_handle(stream), C.byref(rep)))
        del keep
        return {"iterations": rep.iterations, "anchors": rep.anchors, "last_error": rep.last_error,
                "bound": rep.bound, "seconds": rep.seconds}

    def page_rank_adjoint(self, pr, w, grad_values=None, grad_teleport=None, beta=0.85, epsilon=1e-7,
                          max_iteration=1000, precision="f32", check_every=8, values=None, teleport=None,
                          stream=None):
        """Gradients of L(p) through the PageRank fixed point ``pr`` given
        ``w`` = dL/dp: dL/dA per stored CSR entry into ``grad_values`` and
        dL/dv into ``grad_teleport`` (float64 device tensors, either may be
        None).  ``values`` / ``teleport`` as in page_rank_device.  Returns
        the run report."""
        _check_f64(pr, "pr")
        _check_f64(w, "w")
        if pr.numel() != w.numel():
            raise ValueError("pr and w must have the same length")
        p = _lib.AdjointParams()
        p.beta, p.epsilon, p.max_iteration = beta, epsilon, max_iteration
        p.precision = dtype_code(precision)
        p.check_every = check_every
        vals_p, vals_n = _opt_f64(values, "values")
        tele_p, tele_n = _opt_f64(teleport, "teleport")
        gv_p, gv_n = _opt_f64(grad_values, "grad_values")
        gt_p, gt_n = _opt_f64(grad_teleport, "grad_teleport")
        if grad_values is not None and gv_n < self.n_entries():
            raise ValueError("grad_values shorter than the stored entries")
        if grad_teleport is not None and gt_n < self.get_n_col():
            raise ValueError("grad_teleport shorter than n_col")
        rep = _lib.AdjointReport()
        check(lib.tbsla_hip_page_rank_adjoint(self._h, C.byref(p), vals_p, vals_n, tele_p, tele_n, _tensor_ptr(pr),
                                              _tensor_ptr(w), pr.numel(), gv_p, gt_p, _stream_handle(stream),
                                              C.byref(rep)))
        return {"iterations": rep.iterations, "last_change": rep.last_change, "s": rep.s,
                "seconds": rep.seconds}

    def n_entries(self) -> int:
        """Stored CSR entries (rowptr[-1] - rowptr[0]): the length of per-entry arrays."""
        if self.FORMAT != _lib.CSR:
            raise TypeError("per-entry arrays need a CSR matrix")
        rp = self._array(1, np.int32)
        return int(rp[-1] - rp[0]) if rp.size > 1 else 0


def _check_f64(t, name):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous float64 CUDA tensor")


def _opt_f64(t, name):
    if t is None:
        return None, 0
    _check_f64(t, name)
    return _tensor_ptr(t), t.numel()


class MatrixCOO(Matrix):
    FORMAT, NAME, _IDX0, _IDX1 = _lib.COO, "COO", "row", "col"

    def __init__(self, n_row=None, n_col=None, values=None, row=None, col=None, _handle=None):
        if _handle is None and values is not None:
            _handle = _from_arrays(self.FORMAT, n_row, n_col, values, row, col)
        super().__init__(_handle)

    def readMM(self, path: str):
        """MatrixMarket coordinate real general/symmetric (MatrixCOO.cpp:230-292):
        1-based indices, symmetric off-diagonal entries mirrored right after
        their original.  Raises ValueError for another header."""
        check(lib.tbsla_matrix_read_mm(self._h, path.encode()))
        return self


class MatrixSCOO(Matrix):
    FORMAT, NAME, _IDX0, _IDX1 = _lib.SCOO, "SCOO", "row", "col"

    def __init__(self, src=None, n_col=None, values=None, row=None, col=None, _handle=None):
        """MatrixSCOO(), MatrixSCOO(coo) or MatrixSCOO(n_row, n_col, values, row, col)."""
        if _handle is None and isinstance(src, MatrixCOO):
            _handle = _convert(src, self.FORMAT)
        elif _handle is None and values is not None:
            _handle = _from_arrays(self.FORMAT, src, n_col, values, row, col)
        super().__init__(_handle)


class MatrixCSR(Matrix):
    FORMAT, NAME, _IDX0, _IDX1 = _lib.CSR, "CSR", "rowptr", "colidx"

    def __init__(self, src=None, n_col=None, values=None, rowptr=None, colidx=None, _handle=None, device=False):
        """MatrixCSR(), MatrixCSR(coo) or MatrixCSR(n_row, n_col, values, rowptr, colidx);
        MatrixCSR(coo, device=True) sorts the entries on the GPU (same arrays)."""
        if _handle is None and isinstance(src, MatrixCOO):
            _handle = _convert(src, self.FORMAT, device)
        elif _handle is None and values is not None:
            _handle = _from_arrays(self.FORMAT, src, n_col, values, rowptr, colidx)
        super().__init__(_handle)


class MatrixELL(Matrix):
    FORMAT, NAME, _IDX0, _IDX1 = _lib.ELL, "ELL", None, "columns"

    def __init__(self, src=None, _handle=None, device=False):
        """MatrixELL() or MatrixELL(coo); device=True sorts the entries on the GPU (same arrays)."""
        if _handle is None and isinstance(src, MatrixCOO):
            _handle = _convert(src, self.FORMAT, device)
        super().__init__(_handle)

    def get_max_col(self):
        return self.info()["max_col"]


class MatrixDENSE(Matrix):
    FORMAT, NAME = _lib.DENSE, "DENSE"

    def __init__(self, src=None, _handle=None):
        if _handle is None and isinstance(src, MatrixCOO):
            _handle = _convert(src, self.FORMAT)
        super().__init__(_handle)


FORMATS = {"COO": MatrixCOO, "SCOO": MatrixSCOO, "CSR": MatrixCSR, "ELL": MatrixELL, "DENSE": MatrixDENSE}


def _convert(coo: MatrixCOO, fmt: int, device: bool = False):
    h = (lib.tbsla_matrix_convert_gpu if device else lib.tbsla_matrix_convert)(coo._h, fmt)
    if not h:
        check(-1)
    return h


def _from_arrays(fmt, n_row, n_col, values, a, b):
    v = np.ascontiguousarray(values, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    h = lib.tbsla_matrix_from_arrays(fmt, n_row, n_col, dptr(v), v.size, iptr(a), a.size, iptr(b), b.size)
    if not h:
        check(-1)
    return h
