"""ctypes front-end to the oracle libraries.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``; never by the product package.

Two back-ends share one interface (:class:`Mat`):

* ``Oracle`` -- ``oracle/liborc.so``, the plain-C restatement of the reference
  path (``tbsla_oracle.c``).  Always available once ``make -C oracle`` ran.
* ``Reference`` -- ``oracle/_ref/libtbsla_ref.so``, the TBSLA reference itself
  compiled from its own sources (``make -C oracle ref``, only where the
  reference tree exists) plus ``ref_shim.cpp``.  Matrix contents are decoded
  from the reference's own binary ``write()`` format.
"""
from __future__ import annotations

import ctypes as C
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_SO = os.path.join(HERE, "liborc.so")
REF_SO = os.path.join(HERE, "_ref", "libtbsla_ref.so")

COO, SCOO, CSR, ELL, DENSE = range(5)
FORMATS = {"COO": COO, "SCOO": SCOO, "CSR": CSR, "ELL": ELL, "DENSE": DENSE}
FORMAT_NAMES = {v: k for k, v in FORMATS.items()}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def decode_binary(fmt: int, blob: bytes, offset: int = 0):
    """Decode one matrix in the reference binary format.

    Layout (MatrixCSR.cpp:179-204 and siblings): n_row n_col ln_row ln_col
    f_row f_col (int32) nnz (int64) pr pc NR NC (int32) [ELL: max_col int32]
    then size_t-prefixed arrays: values (f64); CSR rowptr, colidx / COO,SCOO
    row, col / ELL columns (int32); DENSE values only.
    Returns (info dict, arrays dict, end offset).
    """
    o = offset
    h = struct.unpack_from("<6iq4i", blob, o)
    o += struct.calcsize("<6iq4i")
    keys = ["n_row", "n_col", "ln_row", "ln_col", "f_row", "f_col", "nnz", "pr", "pc", "NR", "NC"]
    info = dict(zip(keys, h))
    info["format"] = FORMAT_NAMES[fmt]
    if fmt == ELL:
        (info["max_col"],) = struct.unpack_from("<i", blob, o)
        o += 4

    def arr(dtype):
        nonlocal o
        (n,) = struct.unpack_from("<Q", blob, o)
        o += 8
        a = np.frombuffer(blob, dtype=dtype, count=n, offset=o).copy()
        o += n * np.dtype(dtype).itemsize
        return a

    out = {"values": arr(np.float64)}
    if fmt == CSR:
        out["rowptr"] = arr(np.int32)
        out["colidx"] = arr(np.int32)
    elif fmt in (COO, SCOO):
        out["row"] = arr(np.int32)
        out["col"] = arr(np.int32)
    elif fmt == ELL:
        out["columns"] = arr(np.int32)
    return info, out, o


class _Backend:
    lib: C.CDLL

    def available(self) -> bool:
        return self.lib is not None


class Mat:
    """A matrix held by one back-end; mirrors tbsla::cpp::Matrix calls."""

    def __init__(self, backend, handle, fmt):
        self.b, self.h, self.fmt = backend, handle, fmt

    def __del__(self):
        try:
            if self.h:
                self.b._free(self.h)
        except Exception:
            pass

    # generators -------------------------------------------------------
    def fill_cdiag(self, n_row, n_col, c, pr=0, pc=0, NR=1, NC=1):
        self.b._fill_cdiag(self.h, n_row, n_col, c, pr, pc, NR, NC)
        return self

    def fill_cqmat(self, n_row, n_col, c, q, seed=1, pr=0, pc=0, NR=1, NC=1, value_seed=0):
        self.b._fill_cqmat(self.h, n_row, n_col, c, q, seed, pr, pc, NR, NC, value_seed)
        return self

    def fill_cqmat_stochastic(self, n_row, n_col, c, q, seed=1, pr=0, pc=0, NR=1, NC=1):
        self.b._fill_cqmat_stochastic(self.h, n_row, n_col, c, q, seed, pr, pc, NR, NC)
        return self

    def normalize_columns(self):
        self.b._normalize(self.h)
        return self

    # contents -----------------------------------------------------------
    def arrays(self):
        return self.b._arrays(self.h, self.fmt)

    @property
    def info(self):
        return self.arrays()[0]

    # products -----------------------------------------------------------
    def spmv(self, x, vect_incr=0):
        out = np.zeros(self.b._info(self.h)["n_row"], dtype=np.float64)
        self.b._spmv(self.h, _f64(x), vect_incr, out)
        return out

    def Ax(self, r, x, vect_incr=0):
        r = _f64(r).copy()
        self.b._Ax(self.h, r, _f64(x), vect_incr)
        return r

    def a_axpx_(self, x):
        n = self.b._info(self.h)["n_row"]
        out = np.zeros(n, dtype=np.float64)
        self.b._a_axpx(self.h, _f64(x), out)
        return out

    def page_rank(self, epsilon, beta, max_iteration):
        n = self.b._info(self.h)["n_col"]
        out = np.zeros(n, dtype=np.float64)
        it = self.b._page_rank(self.h, epsilon, beta, max_iteration, out)
        return out, it

    def personalized_page_rank(self, epsilon, beta, max_iteration, nodes):
        n = self.b._info(self.h)["n_col"]
        out = np.zeros(n, dtype=np.float64)
        it = self.b._ppr(self.h, epsilon, beta, max_iteration, _i32(nodes), out)
        return out, it


class _Oracle(_Backend):
    name = "oracle"

    def __init__(self):
        self.lib = None
        if os.path.exists(ORC_SO):
            self._load()

    def _load(self):
        L = C.CDLL(ORC_SO)
        L.orc_new.restype = C.c_void_p
        L.orc_new.argtypes = [C.c_int]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_fill_cdiag.argtypes = [C.c_void_p] + [C.c_int] * 7
        L.orc_fill_cqmat.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint,
                                     C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]
        L.orc_fill_cqmat_stochastic.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.c_uint, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_normalize_columns.argtypes = [C.c_void_p]
        L.orc_info.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
        L.orc_from_arrays.restype = C.c_void_p
        L.orc_from_arrays.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_size_t, _ip, C.c_size_t,
                                      _ip, C.c_size_t]
        L.orc_convert.restype = C.c_void_p
        L.orc_convert.argtypes = [C.c_void_p, C.c_int]
        L.orc_Ax.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.orc_spmv.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
        L.orc_a_axpx.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
        L.orc_page_rank.restype = C.c_int
        L.orc_page_rank.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp,
                                    C.POINTER(C.c_double)]
        L.orc_personalized_page_rank.restype = C.c_int
        L.orc_personalized_page_rank.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _ip,
                                                 C.c_int, _dp, C.POINTER(C.c_double)]
        L.orc_u01.restype = C.c_double
        L.orc_u01.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_uniform_vector.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, _dp]
        L.orc_spmv_res_cdiag.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_lnv.restype = C.c_size_t
        L.orc_lnv.argtypes = [C.c_size_t, C.c_int, C.c_int]
        L.orc_pflv.restype = C.c_size_t
        L.orc_pflv.argtypes = [C.c_size_t, C.c_int, C.c_int]
        L.orc_write.restype = C.c_long
        L.orc_write.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_read.restype = C.c_long
        L.orc_read.argtypes = [C.c_void_p, C.c_char_p, C.c_long]
        L.orc_cqmat_value.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                      C.c_uint, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_double)]
        L.orc_cqmat_sum_columns.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                            C.c_uint, _dp]
        self.lib = L

    # matrix construction ----------------------------------------------------
    def new(self, fmt) -> Mat:
        fmt = FORMATS[fmt] if isinstance(fmt, str) else fmt
        return Mat(self, self.lib.orc_new(fmt), fmt)

    def from_arrays(self, fmt, n_row, n_col, values, a, b) -> Mat:
        fmt = FORMATS[fmt] if isinstance(fmt, str) else fmt
        v, a, b = _f64(values), _i32(a), _i32(b)
        h = self.lib.orc_from_arrays(fmt, n_row, n_col, v, v.size, a, a.size, b, b.size)
        return Mat(self, h, fmt)

    def convert(self, coo: Mat, fmt) -> Mat:
        fmt = FORMATS[fmt] if isinstance(fmt, str) else fmt
        return Mat(self, self.lib.orc_convert(coo.h, fmt), fmt)

    def read(self, fmt, path, seek=0) -> Mat:
        m = self.new(fmt)
        pos = self.lib.orc_read(m.h, path.encode(), seek)
        if pos < 0:
            raise IOError(path)
        return m

    def write(self, m: Mat, path) -> int:
        return self.lib.orc_write(m.h, path.encode())

    # plumbing used by Mat ----------------------------------------------------
    def _free(self, h):
        self.lib.orc_free(h)

    def _fill_cdiag(self, h, *a):
        self.lib.orc_fill_cdiag(h, *a)

    def _fill_cqmat(self, h, *a):
        self.lib.orc_fill_cqmat(h, *a)

    def _fill_cqmat_stochastic(self, h, *a):
        self.lib.orc_fill_cqmat_stochastic(h, *a)

    def _normalize(self, h):
        self.lib.orc_normalize_columns(h)

    def _info(self, h):
        buf = (C.c_long * 16)()
        self.lib.orc_info(h, buf)
        keys = ["fmt", "n_row", "n_col", "f_row", "f_col", "ln_row", "ln_col", "pr", "pc", "NR",
                "NC", "nnz", "max_col", "n_val", "n_idx0", "n_idx1"]
        return dict(zip(keys, list(buf)))

    def _arrays(self, h, fmt):
        # Serialise through orc_write into memory via a temp file is slow; read
        # the struct directly instead.
        info = self._info(h)

        class M(C.Structure):
            _fields_ = [("fmt", C.c_int), ("n_row", C.c_int), ("n_col", C.c_int),
                        ("f_row", C.c_int), ("f_col", C.c_int), ("ln_row", C.c_int),
                        ("ln_col", C.c_int), ("pr", C.c_int), ("pc", C.c_int), ("NR", C.c_int),
                        ("NC", C.c_int), ("nnz", C.c_long), ("max_col", C.c_int),
                        ("val", C.POINTER(C.c_double)), ("n_val", C.c_size_t),
                        ("cap_val", C.c_size_t), ("idx0", C.POINTER(C.c_int)),
                        ("n_idx0", C.c_size_t), ("cap_idx0", C.c_size_t),
                        ("idx1", C.POINTER(C.c_int)), ("n_idx1", C.c_size_t),
                        ("cap_idx1", C.c_size_t)]

        s = C.cast(C.c_void_p(h), C.POINTER(M)).contents

        def grab(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        out = {"values": grab(s.val, s.n_val, np.float64)}
        if fmt == CSR:
            out["rowptr"] = grab(s.idx0, s.n_idx0, np.int32)
            out["colidx"] = grab(s.idx1, s.n_idx1, np.int32)
        elif fmt in (COO, SCOO):
            out["row"] = grab(s.idx0, s.n_idx0, np.int32)
            out["col"] = grab(s.idx1, s.n_idx1, np.int32)
        elif fmt == ELL:
            out["columns"] = grab(s.idx1, s.n_idx1, np.int32)
        keys = ["n_row", "n_col", "ln_row", "ln_col", "f_row", "f_col", "nnz", "pr", "pc", "NR",
                "NC"]
        meta = {k: info[k] for k in keys}
        meta["format"] = FORMAT_NAMES[fmt]
        if fmt == ELL:
            meta["max_col"] = info["max_col"]
        return meta, out

    def _spmv(self, h, x, vi, out):
        self.lib.orc_spmv(h, x, vi, out)

    def _Ax(self, h, r, x, vi):
        self.lib.orc_Ax(h, r, x, vi)

    def _a_axpx(self, h, x, out):
        self.lib.orc_a_axpx(h, x, 0, out)

    def _page_rank(self, h, eps, beta, it, out):
        err = C.c_double()
        return self.lib.orc_page_rank(h, eps, beta, it, out, C.byref(err))

    def _ppr(self, h, eps, beta, it, nodes, out):
        err = C.c_double()
        return self.lib.orc_personalized_page_rank(h, eps, beta, it, nodes, nodes.size, out,
                                                   C.byref(err))

    # generators ---------------------------------------------------------------
    def u01(self, seed, idx):
        return self.lib.orc_u01(seed, idx)

    def uniform_vector(self, seed, n, first=0):
        out = np.empty(n, dtype=np.float64)
        self.lib.orc_uniform_vector(seed, first, n, out)
        return out

    def spmv_res_cdiag(self, nr, nc, c, v):
        out = np.zeros(nr, dtype=np.float64)
        self.lib.orc_spmv_res_cdiag(nr, nc, c, _f64(v), out)
        return out

    def cqmat_sum_columns(self, nr, nc, c, q, seed):
        out = np.zeros(nc, dtype=np.float64)
        self.lib.orc_cqmat_sum_columns(nr, nc, c, q, seed, out)
        return out

    def lnv(self, size, p, P):
        return self.lib.orc_lnv(size, p, P)

    def pflv(self, size, p, P):
        return self.lib.orc_pflv(size, p, P)


class _Reference(_Backend):
    name = "reference"

    def __init__(self):
        self.lib = None
        if os.path.exists(REF_SO):
            self._load()

    def _load(self):
        L = C.CDLL(REF_SO)
        L.ref_new.restype = C.c_void_p
        L.ref_new.argtypes = [C.c_int]
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_fill_cdiag.argtypes = [C.c_void_p] + [C.c_int] * 7
        L.ref_fill_cqmat.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint,
                                     C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_fill_cqmat_stochastic.argtypes = L.ref_fill_cqmat.argtypes
        L.ref_normalize_columns.argtypes = [C.c_void_p]
        L.ref_convert.restype = C.c_void_p
        L.ref_convert.argtypes = [C.c_void_p, C.c_int]
        L.ref_from_arrays.restype = C.c_void_p
        L.ref_from_arrays.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_size_t, _ip, C.c_size_t,
                                      _ip, C.c_size_t]
        L.ref_serialize.restype = C.c_void_p
        L.ref_serialize.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
        L.ref_blob_free.argtypes = [C.c_void_p]
        L.ref_deserialize.restype = C.c_void_p
        L.ref_deserialize.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
        L.ref_get_nnz.restype = C.c_long
        L.ref_get_nnz.argtypes = [C.c_void_p]
        L.ref_get_n_row.restype = C.c_int
        L.ref_get_n_row.argtypes = [C.c_void_p]
        L.ref_get_n_col.restype = C.c_int
        L.ref_get_n_col.argtypes = [C.c_void_p]
        L.ref_spmv.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_int, _dp]
        L.ref_Ax.argtypes = [C.c_void_p, _dp, C.c_size_t, _dp, C.c_size_t, C.c_int]
        L.ref_Ax_repeat.argtypes = [C.c_void_p, _dp, C.c_size_t, _dp, C.c_size_t, C.c_int]
        L.ref_a_axpx.argtypes = [C.c_void_p, _dp, C.c_size_t, _dp]
        L.ref_Ax_times.argtypes = [C.c_void_p, _dp, C.c_size_t, _dp, C.c_size_t, C.c_int, _dp]
        L.ref_spmv_times.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_int, _dp]
        L.ref_page_rank.restype = C.c_int
        L.ref_page_rank.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp]
        L.ref_read_mm.restype = C.c_int
        L.ref_read_mm.argtypes = [C.c_void_p, C.c_char_p]
        L.ref_personalized_page_rank.restype = C.c_int
        L.ref_personalized_page_rank.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _ip,
                                                 C.c_int, _dp]
        self.lib = L

    def new(self, fmt) -> Mat:
        fmt = FORMATS[fmt] if isinstance(fmt, str) else fmt
        return Mat(self, self.lib.ref_new(fmt), fmt)

    def from_arrays(self, fmt, n_row, n_col, values, a, b) -> Mat:
        fmt = FORMATS[fmt] if isinstance(fmt, str) else fmt
        v, a, b = _f64(values), _i32(a), _i32(b)
        return Mat(self, self.lib.ref_from_arrays(fmt, n_row, n_col, v, v.size, a, a.size, b,
                                                  b.size), fmt)

    def convert(self, coo: Mat, fmt) -> Mat:
        fmt = FORMATS[fmt] if isinstance(fmt, str) else fmt
        return Mat(self, self.lib.ref_convert(coo.h, fmt), fmt)

    def serialize(self, m: Mat) -> bytes:
        """The bytes the reference's write() produces for m."""
        return self._serialize(m.h)

    def _serialize(self, h) -> bytes:
        p = C.c_void_p()
        n = C.c_size_t()
        blob = self.lib.ref_serialize(h, C.byref(p), C.byref(n))
        try:
            return C.string_at(p, n.value)
        finally:
            self.lib.ref_blob_free(blob)

    def deserialize(self, fmt, data: bytes) -> Mat:
        """A matrix read by the reference's read() from `data`."""
        fmt = FORMATS[fmt] if isinstance(fmt, str) else fmt
        return Mat(self, self.lib.ref_deserialize(fmt, data, len(data)), fmt)

    def _free(self, h):
        self.lib.ref_free(h)

    def _fill_cdiag(self, h, *a):
        self.lib.ref_fill_cdiag(h, *a)

    def _fill_cqmat(self, h, *a):
        if a[-1]:
            raise ValueError("the reference has no value overlay")
        self.lib.ref_fill_cqmat(h, *a[:-1])

    def _fill_cqmat_stochastic(self, h, *a):
        self.lib.ref_fill_cqmat_stochastic(h, *a)

    def _normalize(self, h):
        self.lib.ref_normalize_columns(h)

    def _info(self, h):
        return {"n_row": self.lib.ref_get_n_row(h), "n_col": self.lib.ref_get_n_col(h),
                "nnz": self.lib.ref_get_nnz(h)}

    def _arrays(self, h, fmt):
        info, arrs, _ = decode_binary(fmt, self._serialize(h))
        return info, arrs

    def _spmv(self, h, x, vi, out):
        self.lib.ref_spmv(h, x, x.size, vi, out)

    def _Ax(self, h, r, x, vi):
        self.lib.ref_Ax(h, r, r.size, x, x.size, vi)

    def _a_axpx(self, h, x, out):
        self.lib.ref_a_axpx(h, x, x.size, out)

    def _page_rank(self, h, eps, beta, it, out):
        return self.lib.ref_page_rank(h, eps, beta, it, out)

    def _ppr(self, h, eps, beta, it, nodes, out):
        return self.lib.ref_personalized_page_rank(h, eps, beta, it, nodes, nodes.size, out)

    def read_mm(self, path: str) -> Mat:
        """MatrixCOO::readMM (MatrixCOO.cpp:230-292); ValueError where the
        reference throws MatrixFormatReadException."""
        m = self.new("COO")
        if not self.lib.ref_read_mm(m.h, path.encode()):
            raise ValueError("MatrixFormatReadException")
        return m

    def Ax_times(self, m: Mat, nr, x, reps):
        """(seconds of each of `reps` Ax calls on preallocated vectors, last result)."""
        r = np.zeros(nr)
        t = np.zeros(reps)
        x = _f64(x)
        self.lib.ref_Ax_times(m.h, r, r.size, x, x.size, reps, t)
        return t, r

    def spmv_times(self, m: Mat, x, reps):
        """Seconds of each of `reps` spmv() calls (each allocates its result)."""
        t = np.zeros(reps)
        x = _f64(x)
        self.lib.ref_spmv_times(m.h, x, x.size, reps, t)
        return t

    def Ax_repeat(self, m: Mat, r, x, reps):
        r = _f64(r).copy()
        self.lib.ref_Ax_repeat(m.h, r, r.size, _f64(x), x.size, reps)
        return r


Oracle = _Oracle()
Reference = _Reference()


def build(with_reference: bool | None = None) -> None:
    """Compile the oracle (``make -C oracle``); the reference only if present."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    ref_tree = os.path.isdir("/root/reference/src/cpp")
    if with_reference or (with_reference is None and ref_tree):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    global Oracle, Reference
    Oracle = _Oracle()
    Reference = _Reference()
