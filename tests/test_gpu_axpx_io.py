"""GPU parity of a_axpx_ (Matrix.cpp:6-11) and of the C ABI file path
(interface/c/matrix/MatrixCSR.cpp:46-93): matrices read from files written
by the compiled reference, or written and read back through
C_CPP_Matrix<F>_write / _read, multiplied on the GPU by C_CPP_Matrix<F>_spmv.
Every product is compared with the reference's own output or the oracle,
bit for bit where the kernel keeps the reference's summation order."""
import os

import numpy as np
import pytest

from conftest import FMTS, GOLDEN

pytestmark = pytest.mark.gpu

BIN = os.path.join(GOLDEN, "ref_binary")
EXACT = {"COO", "SCOO", "CSR", "ELL"}


def res_cdiag(nr, nc, c, v):
    """utils/vector.cpp:16-29 (compute_spmv_res_cdiag)."""
    from test_axpx_io import res_cdiag as f
    return f(nr, nc, c, v)


def test_a_axpx_cdiag_closed_form(tb, gpu):
    """tests/test_a_axpx__cdiag.cpp: x_i = 2i + 1, the nine matrices of each
    shape (COO; SCOO, CSR, ELL, DENSE both converted and filled), equal to
    the closed form A(Ax + x) of utils/vector.cpp:52-76.  Integer data, so
    every format is exact."""
    from test_axpx_io import AXPX_SHAPES, axpx_closed_form
    for (n, c) in AXPX_SHAPES:
        v = 2.0 * np.arange(n) + 1.0
        expect = axpx_closed_form(n, c, v)
        coo = tb.MatrixCOO().fill_cdiag(n, n, c)
        mats = [("COO", coo)]
        for fmt in ["SCOO", "CSR", "ELL", "DENSE"]:
            mats += [(fmt + "(coo)", tb.FORMATS[fmt](coo)), (fmt, tb.FORMATS[fmt]().fill_cdiag(n, n, c))]
        for name, m in mats:
            assert np.array_equal(m.a_axpx_(v), expect), (name, n, c)


def test_a_axpx_matches_oracle(tb, oracle, gpu):
    """Non-integer data -- the value overlay and 1/count values -- in every
    format and block layout, v of the global length: equal to the oracle bit
    for bit (DENSE within 1e-13 of |A|(|A||v| + |v|))."""
    n = 600
    v = oracle.uniform_vector(5, n)
    for fmt in FMTS:
        for blk in [(0, 0, 1, 1), (1, 0, 2, 2), (1, 2, 2, 3)]:
            for gen, kw in (("fill_cqmat", {"value_seed": 42}), ("fill_cqmat_stochastic", {})):
                p = getattr(tb.FORMATS[fmt](), gen)(n, n, 9, 0.3, 11, *blk, **kw)
                o = getattr(oracle.new(fmt), gen)(n, n, 9, 0.3, 11, *blk, **kw)
                y, yo = p.a_axpx_(v), o.a_axpx_(v)
                if fmt in EXACT:
                    assert np.array_equal(y, yo), (fmt, blk, gen, np.abs(y - yo).max())
                else:
                    # values and v are non-negative, so yo is |A|(|A||v| + |v|)
                    assert np.all(np.abs(y - yo) <= 1e-13 * yo + 1e-300), (fmt, blk, gen)


def test_a_axpx_needs_a_square_operand(tb, gpu):
    from tbsla_amd._lib import TbslaError
    m = tb.MatrixCSR().fill_cqmat(100, 60, 5, 0.2, 1)
    with pytest.raises(TbslaError):
        m.a_axpx_(np.ones(60))


def _vec_from_file(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    assert len(raw) == 8 + 8 * n
    return np.frombuffer(raw[8:], dtype=np.float64).copy()


def _vec_to_file(path, x):
    with open(path, "wb") as f:
        f.write(np.uint64(x.size).tobytes())
        f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())


def test_c_abi_spmv_on_reference_files(tb, golden, digests, gpu, tmp_path):
    """C_CPP_Matrix<F>_read of a file written by the reference, then
    C_CPP_Matrix<F>_spmv on x = 0, 1, 2, ... (C_CPP_Vector_fill): equal to the
    reference's own SpMV of that matrix (fixture); DENSE within 1e-14 of
    sum|a||x|."""
    from tbsla_amd._lib import lib
    for fn in sorted(os.listdir(BIN)):
        if fn.startswith("appended"):
            continue
        F = fn.rsplit(".", 1)[1]
        m = getattr(lib, f"C_CPP_Matrix{F}_create")()
        assert getattr(lib, f"C_CPP_Matrix{F}_read")(m, os.path.join(BIN, fn).encode(), 0), fn
        v = lib.C_CPP_Vector_create()
        lib.C_CPP_Vector_fill(v, digests["binary"][fn]["nx"], 0)
        r = getattr(lib, f"C_CPP_Matrix{F}_spmv")(m, v)
        out = str(tmp_path / f"y_{fn}")
        assert lib.C_CPP_Vector_write(r, out.encode())
        y = _vec_from_file(out)
        ref = golden[f"binary/{fn}/spmv_iota"]
        if F in EXACT:
            assert np.array_equal(y, ref), fn
        else:
            assert np.allclose(y, ref, rtol=1e-14, atol=0), fn
        lib.C_CPP_Vector_destroy(r)
        lib.C_CPP_Vector_destroy(v)
        getattr(lib, f"C_CPP_Matrix{F}_destroy")(m)


def test_c_abi_write_read_spmv(tb, oracle, gpu, tmp_path):
    """The YML flow through the C ABI: fill, write (appended twice), read the
    second copy into a new handle, read x from a vector file
    (C_CPP_Vector_read), spmv, write the result: equal to the oracle's
    product of the same matrix."""
    from tbsla_amd._lib import lib
    for F in FMTS:
        m = getattr(lib, f"C_CPP_Matrix{F}_create")()
        getattr(lib, f"C_CPP_Matrix{F}_fill_cqmat")(m, 500, 500, 8, 0.3, 4, 1, 0, 2, 2)
        path = str(tmp_path / f"m.{F}")
        assert getattr(lib, f"C_CPP_Matrix{F}_write")(m, path.encode())
        size1 = os.path.getsize(path)
        assert getattr(lib, f"C_CPP_Matrix{F}_write")(m, path.encode())
        m2 = getattr(lib, f"C_CPP_Matrix{F}_create")()
        assert getattr(lib, f"C_CPP_Matrix{F}_read")(m2, path.encode(), size1)
        nx = 500 if F == "COO" else 250
        x = oracle.uniform_vector(9, nx)
        xf = str(tmp_path / f"x.{F}")
        _vec_to_file(xf, x)
        v = lib.C_CPP_Vector_create()
        assert lib.C_CPP_Vector_read(v, xf.encode(), 0)
        r = getattr(lib, f"C_CPP_Matrix{F}_spmv")(m2, v)
        out = str(tmp_path / f"y.{F}")
        assert lib.C_CPP_Vector_write(r, out.encode())
        y = _vec_from_file(out)
        o = oracle.new(F).fill_cqmat(500, 500, 8, 0.3, 4, 1, 0, 2, 2)
        yo = o.spmv(x)
        if F in EXACT:
            assert np.array_equal(y, yo), F
        else:
            assert np.all(np.abs(y - yo) <= 1e-14 * o.spmv(np.abs(x)) + 1e-300), F
        for h in (r, v):
            lib.C_CPP_Vector_destroy(h)
        getattr(lib, f"C_CPP_Matrix{F}_destroy")(m)
        getattr(lib, f"C_CPP_Matrix{F}_destroy")(m2)


def test_c_abi_spmv_error_is_reported(tb, gpu, tmp_path):
    """C_CPP_Matrix<F>_spmv with an x shorter than the matrix reads: an empty
    vector (the reference would read past x) and the reason in
    tbsla_last_error(); a later successful call clears it."""
    from tbsla_amd._lib import lib
    m = lib.C_CPP_MatrixCSR_create()
    lib.C_CPP_MatrixCSR_fill_cqmat(m, 300, 300, 6, 0.2, 1, 0, 0, 1, 1)
    v = lib.C_CPP_Vector_create()
    lib.C_CPP_Vector_fill(v, 100, 0)
    r = lib.C_CPP_MatrixCSR_spmv(m, v)
    out = str(tmp_path / "short")
    assert lib.C_CPP_Vector_write(r, out.encode())
    assert _vec_from_file(out).size == 0
    assert b"x has 100 entries" in lib.tbsla_last_error()
    lib.C_CPP_Vector_fill(v, 300, 0)
    r2 = lib.C_CPP_MatrixCSR_spmv(m, v)
    assert lib.tbsla_last_error() == b""
    for h in (r, r2, v):
        lib.C_CPP_Vector_destroy(h)
    lib.C_CPP_MatrixCSR_destroy(m)
