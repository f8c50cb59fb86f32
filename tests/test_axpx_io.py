"""Pins for a_axpx_ and the binary file format, on the CPU.

a_axpx_ (Matrix.cpp:6-11): the oracle's orc_a_axpx against the reference's
own closed form (tests/test_a_axpx__cdiag.cpp:19-84 with
utils/vector.cpp:52-76, x_i = 2i + 1), against planted mistakes, and against
the compiled reference on non-integer values.

Binary files (MatrixCSR.cpp:179-232, MatrixCOO.cpp:174-228,
MatrixSCOO.cpp:174-228, MatrixELL.cpp:142-190, MatrixDENSE.cpp:98-135):
files written by the compiled reference (tests/golden/ref_binary, made by
tests/golden/make_golden.py) are read by the product and by the oracle and
written back byte for byte; files written here are read by the reference.
The C ABI read/write (interface/c/matrix/MatrixCSR.cpp:46-70) is exercised
the same way.  Products need no GPU here.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import FMTS, GOLDEN, INFO_KEYS, single_thread_omp

BIN = os.path.join(GOLDEN, "ref_binary")


def res_cdiag(nr, nc, c, v):
    """utils/vector.cpp:16-29 (compute_spmv_res_cdiag), restated."""
    r = np.zeros(nr)
    i = 0
    while i < min(c, nr):
        r[i] = v[i + c] if i < nc - c else 0
        i += 1
    while i < min(nr, nc - c):
        r[i] = v[i] if c == 0 else v[i + c] + v[i - c]
        i += 1
    while i < nr:
        r[i] = v[i - c] if i < nc + c else 0
        i += 1
    return r


def axpx_closed_form(n, c, v):
    """utils/vector.cpp:52-76: A(Av + v) for the cdiag matrix."""
    return res_cdiag(n, n, c, res_cdiag(n, n, c, v) + v)


# the square shapes of tests/test_a_axpx__cdiag.cpp:73-84
AXPX_SHAPES = [(10, i) for i in range(13)] + [(30, 2 * i) for i in range(13)]


def cdiag_variants(lib_new, lib_convert, n, c):
    """The nine matrices test_a_axpx__cdiag.cpp:42-70 builds: COO, then every
    other format both converted from that COO and filled directly."""
    coo = lib_new("COO").fill_cdiag(n, n, c)
    out = [("COO", coo)]
    for fmt in ["SCOO", "CSR", "ELL", "DENSE"]:
        out.append((fmt + "(coo)", lib_convert(coo, fmt)))
        out.append((fmt, lib_new(fmt).fill_cdiag(n, n, c)))
    return out


def test_oracle_a_axpx_closed_form(oracle):
    for (n, c) in AXPX_SHAPES:
        v = 2.0 * np.arange(n) + 1.0
        expect = axpx_closed_form(n, c, v)
        for name, m in cdiag_variants(oracle.new, oracle.convert, n, c):
            assert np.array_equal(m.a_axpx_(v), expect), (name, n, c)


def test_closed_form_against_dense_band():
    """The closed form itself against a dense +-c band (mathematics only)."""
    for (n, c) in AXPX_SHAPES:
        A = np.zeros((n, n))
        for i in range(n):
            for j in ({i - c, i + c} if c else {i}):
                if 0 <= j < n:
                    A[i, j] = 1
        v = 2.0 * np.arange(n) + 1.0
        assert np.array_equal(A @ (A @ v + v), axpx_closed_form(n, c, v)), (n, c)


def test_closed_form_catches_planted_mistakes(oracle):
    """Mistakes an a_axpx_ restatement could make -- the '+ v' dropped, one
    product instead of two, v added after the second product -- all fail the
    closed-form check on the reference shape grid."""
    caught = {"no_plus_v": False, "one_product": False, "plus_v_after": False}
    for (n, c) in AXPX_SHAPES:
        v = 2.0 * np.arange(n) + 1.0
        expect = axpx_closed_form(n, c, v)
        m = oracle.new("CSR").fill_cdiag(n, n, c)
        wrong = {"no_plus_v": m.spmv(m.spmv(v)), "one_product": m.spmv(v) + v,
                 "plus_v_after": m.spmv(m.spmv(v)) + v}
        for k, w in wrong.items():
            caught[k] |= not np.array_equal(w, expect)
    assert all(caught.values()), caught


def test_oracle_a_axpx_against_reference(oracle, reference):
    """Non-integer values: 1/count values of fill_cqmat_stochastic (block
    layouts, every format, global-length v as the reference requires) and
    the value overlay through the array constructors -- bit-exact, with the
    reference's COO/SCOO reductions in their single-thread order."""
    with single_thread_omp():
        _a_axpx_against_reference(oracle, reference)


def _a_axpx_against_reference(oracle, reference):
    n = 120
    v = oracle.uniform_vector(5, n)
    for fmt in FMTS:
        for blk in [(0, 0, 1, 1), (1, 0, 2, 2), (1, 2, 2, 3)]:
            o = oracle.new(fmt).fill_cqmat_stochastic(n, n, 7, 0.4, 3, *blk)
            r = reference.new(fmt).fill_cqmat_stochastic(n, n, 7, 0.4, 3, *blk)
            assert np.array_equal(o.a_axpx_(v), r.a_axpx_(v)), (fmt, blk)
    # the value overlay through the COO array constructor
    _, a = oracle.new("COO").fill_cqmat(n, n, 9, 0.3, 11, value_seed=42).arrays()
    coo = oracle.from_arrays("COO", n, n, a["values"], a["row"], a["col"])
    rcoo = reference.from_arrays("COO", n, n, a["values"], a["row"], a["col"])
    assert np.array_equal(coo.a_axpx_(v), rcoo.a_axpx_(v))
    # the conversions of a generated COO (the reference's array constructor
    # leaves nnz unset, MatrixCOO.cpp:30-44, and its ELL and DENSE products
    # return early when the nnz they copy is 0); 1/count values make
    # repeated (row, col) entries equal, so the reference's unstable sort
    # cannot change the sums
    coo = oracle.new("COO").fill_cqmat_stochastic(n, n, 9, 0.3, 11)
    rcoo = reference.new("COO").fill_cqmat_stochastic(n, n, 9, 0.3, 11)
    for fmt in ["SCOO", "CSR", "ELL", "DENSE"]:
        assert np.array_equal(oracle.convert(coo, fmt).a_axpx_(v), reference.convert(rcoo, fmt).a_axpx_(v)), fmt


# ----------------------------------------------------------- binary files

def ref_files():
    return sorted(f for f in os.listdir(BIN) if not f.startswith("appended"))


def test_reference_files_match_digests(digests):
    for fn in ref_files() + ["appended2.CSR"]:
        raw = open(os.path.join(BIN, fn), "rb").read()
        assert hashlib.sha256(raw).hexdigest() == digests["binary"][fn]["sha256"], fn


def test_reference_files_read_and_rewritten(tb, oracle, tmp_path):
    """Product and oracle read each reference-written file and write it back
    byte for byte; the arrays are those the file holds."""
    from oracle.oracle import FORMATS, decode_binary
    for fn in ref_files():
        fmt = fn.rsplit(".", 1)[1]
        raw = open(os.path.join(BIN, fn), "rb").read()
        info, arrs, end = decode_binary(FORMATS[fmt], raw)
        assert end == len(raw)
        p = tb.FORMATS[fmt]().read(os.path.join(BIN, fn))
        out = str(tmp_path / ("p_" + fn))
        p.write(out)
        assert open(out, "rb").read() == raw, fn
        pi = p.info()
        for k in INFO_KEYS + ["nnz"]:
            assert pi[k] == info[k], (fn, k)
        for k, a in arrs.items():
            assert np.array_equal(p.arrays()[k], a), (fn, k)
        o = oracle.read(fmt, os.path.join(BIN, fn))
        out = str(tmp_path / ("o_" + fn))
        oracle.write(o, out)
        assert open(out, "rb").read() == raw, fn


def test_appended_file_seek(tb, oracle, digests):
    """Two matrices appended in one file (C_CPP_Matrix<F>_write appends): the
    second is read at its byte offset."""
    path = os.path.join(BIN, "appended2.CSR")
    at = digests["binary"]["appended2.CSR"]["second_at"]
    raw = open(path, "rb").read()
    first = tb.MatrixCSR().read(path, 0)
    second = tb.MatrixCSR().read(path, at)
    assert first.info()["n_row"] == 40 and second.info()["n_row"] == 50
    assert second.count_stochastic() is False  # read values carry no generator tag
    from oracle.oracle import CSR, decode_binary
    _, a2, end = decode_binary(CSR, raw, at)
    assert end == len(raw)
    for k, a in a2.items():
        assert np.array_equal(second.arrays()[k], a)
    o = oracle.read("CSR", path, at)
    assert np.array_equal(o.arrays()[1]["colidx"], a2["colidx"])


def test_product_files_read_by_reference(tb, oracle, reference, tmp_path):
    """Files written here -- value overlay, block layouts, stochastic values --
    are read by the reference's read(), which writes the same bytes back, and
    its products over them equal the oracle's."""
    cases = [("fill_cqmat", (60, 60, 7, 0.3, 5, 1, 1, 2, 2), {"value_seed": 9}),
             ("fill_cqmat_stochastic", (70, 70, 6, 0.5, 2, 0, 0, 1, 1), {}),
             ("fill_cdiag", (40, 40, 3, 0, 1, 2, 2), {})]
    for gen, args, kw in cases:
        for fmt in FMTS:
            p = getattr(tb.FORMATS[fmt](), gen)(*args, **kw)
            path = str(tmp_path / f"{gen}.{fmt}")
            p.write(path)
            raw = open(path, "rb").read()
            r = reference.deserialize(fmt, raw)
            assert reference.serialize(r) == raw, (gen, fmt)
            o = getattr(oracle.new(fmt), gen)(*args, **kw)
            nx = o.info["n_col"] if fmt == "COO" else o.info["ln_col"]
            x = oracle.uniform_vector(3, nx)
            with single_thread_omp():
                yr, yo = r.spmv(x), o.spmv(x)
            assert np.array_equal(yr, yo), (gen, fmt)
            oracle.write(o, str(tmp_path / f"o_{gen}.{fmt}"))
            oraw = open(str(tmp_path / f"o_{gen}.{fmt}"), "rb").read()
            if gen == "fill_cdiag" and fmt in ("CSR", "SCOO"):
                # the reference leaves nnz unset there (header bytes 24-31);
                # the product sets it, the oracle leaves 0
                oraw, raw = oraw[:24] + oraw[32:], raw[:24] + raw[32:]
            assert oraw == raw, (gen, fmt)


def _c_cpp(lib, F, op):
    return getattr(lib, f"C_CPP_Matrix{F}_{op}")


def test_c_abi_write_read_round_trip(tb, tmp_path):
    """C_CPP_Matrix<F>_write appends (interface/c/matrix/MatrixCSR.cpp:60-70);
    _read at the second matrix's offset restores it; both agree byte for byte
    with the reference-written files."""
    from tbsla_amd._lib import lib
    for fn in ref_files():
        F = fn.rsplit(".", 1)[1]
        raw = open(os.path.join(BIN, fn), "rb").read()
        m = _c_cpp(lib, F, "create")()
        assert _c_cpp(lib, F, "read")(m, os.path.join(BIN, fn).encode(), 0)
        path = str(tmp_path / fn)
        assert _c_cpp(lib, F, "write")(m, path.encode())
        assert _c_cpp(lib, F, "write")(m, path.encode())
        assert open(path, "rb").read() == raw + raw, fn
        m2 = _c_cpp(lib, F, "create")()
        assert _c_cpp(lib, F, "read")(m2, path.encode(), len(raw))
        path2 = str(tmp_path / ("again_" + fn))
        assert _c_cpp(lib, F, "write")(m2, path2.encode())
        assert open(path2, "rb").read() == raw, fn
        _c_cpp(lib, F, "destroy")(m)
        _c_cpp(lib, F, "destroy")(m2)
    assert not lib.C_CPP_MatrixCSR_read(None, b"x", 0)
    m = lib.C_CPP_MatrixCSR_create()
    assert not lib.C_CPP_MatrixCSR_read(m, str(tmp_path / "missing").encode(), 0)
    lib.C_CPP_MatrixCSR_destroy(m)


def test_truncated_and_corrupt_files_rejected(tb, tmp_path):
    """A truncated file, or one whose array size exceeds the file, is an
    error that leaves the matrix as it was (the reference reads past the
    end); no allocation of the bogus size is attempted."""
    from tbsla_amd._lib import TbslaError, lib
    raw = open(os.path.join(BIN, "stoch50c5q0.5s3.CSR"), "rb").read()
    good = tb.MatrixCSR().fill_cqmat(30, 30, 4, 0.2, 1)
    before = str(tmp_path / "before")
    good.write(before)
    for cut in (10, 48, 60, len(raw) - 4):
        bad = str(tmp_path / f"cut{cut}")
        open(bad, "wb").write(raw[:cut])
        with pytest.raises(TbslaError):
            good.read(bad)
        m = lib.C_CPP_MatrixCSR_create()
        assert not lib.C_CPP_MatrixCSR_read(m, bad.encode(), 0)
        assert b"truncated" in lib.tbsla_last_error() or b"holds" in lib.tbsla_last_error()
        lib.C_CPP_MatrixCSR_destroy(m)
    corrupt = bytearray(raw)
    corrupt[48:56] = (1 << 60).to_bytes(8, "little")  # size of the values array
    bad = str(tmp_path / "corrupt")
    open(bad, "wb").write(bytes(corrupt))
    with pytest.raises(TbslaError, match="holds"):
        good.read(bad)
    after = str(tmp_path / "after")
    good.write(after)
    assert open(after, "rb").read() == open(before, "rb").read()
    # C_CPP_Vector_read (Vector.cpp:73-88) with the same corrupt size
    v = lib.C_CPP_Vector_create()
    vf = str(tmp_path / "vec")
    open(vf, "wb").write((1 << 60).to_bytes(8, "little") + b"\0" * 16)
    assert not lib.C_CPP_Vector_read(v, vf.encode(), 0)
    lib.C_CPP_Vector_destroy(v)
