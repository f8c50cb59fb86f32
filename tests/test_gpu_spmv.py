"""GPU SpMV parity (through the C ABI of libtbsla_hip.so) against the oracle.

The reference tests are the model: the 5x5 known answer (tests/csr.cpp),
the cdiag closed form over the reference shape grid (tests/test_spmv_cdiag.cpp),
exact cross-format equality on cqmat matrices (tests/test_spmv_cqmat.cpp).
FP64 products are compared bit for bit where the kernel keeps the reference
summation order (CSR, ELL, COO, SCOO) and within 1e-14 relative, elementwise
against sum|a||x|, for DENSE.
"""
import numpy as np
import pytest

from conftest import CDIAG_SHAPES, CQMAT_SHAPES, FMTS, grid_q

pytestmark = pytest.mark.gpu

EXACT = {"COO", "SCOO", "CSR", "ELL"}


def close(fmt, y, ref, absy):
    if fmt in EXACT:
        assert np.array_equal(y, ref), (fmt, np.abs(y - ref).max())
    else:
        assert np.all(np.abs(y - ref) <= 1e-14 * absy + 1e-300), (fmt, np.abs(y - ref).max())


def test_known_answer_5x5(tb, golden, gpu):
    csr = tb.MatrixCSR(5, 5, golden["ka_values"], golden["ka_rowptr"], golden["ka_colidx"])
    coo = tb.MatrixCOO(5, 5, golden["ka_values"], golden["ka_row"], golden["ka_colidx"])
    assert np.array_equal(csr.spmv(golden["ka_x"]), golden["ka_expected"])
    assert np.array_equal(coo.spmv(golden["ka_x"]), golden["ka_expected"])
    for cls in (tb.MatrixSCOO, tb.MatrixCSR, tb.MatrixELL, tb.MatrixDENSE):
        assert np.array_equal(cls(coo).spmv(golden["ka_x"]), golden["ka_expected"]), cls.NAME


def test_cdiag_closed_form_all_formats(tb, oracle, gpu):
    """tests/test_spmv_cdiag.cpp:41-105: direct fill and via COO, exact."""
    for (nr, nc, c) in CDIAG_SHAPES:
        x = np.arange(nc, dtype=np.float64)
        expect = oracle.spmv_res_cdiag(nr, nc, c, x)
        coo = tb.MatrixCOO().fill_cdiag(nr, nc, c)
        for fmt in FMTS:
            m = tb.FORMATS[fmt]().fill_cdiag(nr, nc, c)
            assert np.array_equal(m.spmv(x), expect), (fmt, nr, nc, c)
            if fmt != "COO":
                assert np.array_equal(tb.FORMATS[fmt](coo).spmv(x), expect), (fmt, nr, nc, c)


def test_cqmat_cross_format_exact(tb, oracle, gpu):
    """tests/test_spmv_cqmat.cpp:19-151: every format, direct and converted,
    equal to the COO product bit for bit (values 1, x = iota), and equal to
    the oracle."""
    for (nr, nc, c) in CQMAT_SHAPES[::2]:
        for s in range(4):
            for q in grid_q()[::2]:
                x = np.arange(nc, dtype=np.float64)
                coo = tb.MatrixCOO().fill_cqmat(nr, nc, c, q, s)
                rc = coo.spmv(x)
                assert np.array_equal(rc, oracle.new("COO").fill_cqmat(nr, nc, c, q, s).spmv(x))
                for fmt in ["SCOO", "CSR", "ELL", "DENSE"]:
                    assert np.array_equal(tb.FORMATS[fmt](coo).spmv(x), rc), (fmt, nr, nc, c, s, q)
                    assert np.array_equal(tb.FORMATS[fmt]().fill_cqmat(nr, nc, c, q, s).spmv(x), rc), (fmt, nr, nc, c)


def test_blocks_vect_incr_and_accumulate(tb, oracle, gpu):
    """Block matrices (x local to the block), Ax accumulation for COO/SCOO/DENSE,
    vect_incr for COO/SCOO, non-integer values: same bits as the oracle's
    sequential loops (DENSE within 1e-14)."""
    rng = np.random.default_rng(5)
    for (n, c, q, s) in [(3000, 16, 0.1, 42), (2500, 9, 0.5, 3)]:
        for blk in [(0, 0, 1, 1), (1, 1, 2, 3), (2, 0, 3, 2)]:
            for fmt in FMTS:
                p = tb.FORMATS[fmt]().fill_cqmat(n, n, c, q, s, *blk, value_seed=42)
                o = oracle.new(fmt).fill_cqmat(n, n, c, q, s, *blk, value_seed=42)
                info = p.info()
                nx = n if fmt == "COO" else info["ln_col"]
                x = oracle.uniform_vector(1, nx, first=0 if fmt == "COO" else info["f_col"])
                r0 = rng.random(info["n_row"] + 7)
                for vi in ([0, 3, 7] if fmt in ("COO", "SCOO") else [0]):
                    y = p.Ax(r0, x, vi)
                    yo = o.Ax(r0, x, vi)
                    absm = np.abs(o.Ax(np.zeros_like(r0), np.abs(x), vi)) + np.abs(r0)
                    close(fmt, y, yo, absm)
                close(fmt, p.spmv(x), o.spmv(x), np.abs(o.spmv(np.abs(x))))


def test_unsorted_coo_is_exact(tb, oracle, gpu):
    """A COO whose rows are not sorted is viewed row-sorted on the device
    (stable), so each row still sums in entry order."""
    rng = np.random.default_rng(1)
    n, nnz = 500, 6000
    row = rng.integers(0, n, nnz).astype(np.int32)
    col = rng.integers(0, n, nnz).astype(np.int32)
    val = rng.random(nnz)
    x = rng.random(n)
    p = tb.MatrixCOO(n, n, val, row, col)
    o = oracle.from_arrays("COO", n, n, val, row, col)
    assert np.array_equal(p.spmv(x), o.spmv(x))
    r0 = rng.random(n)
    assert np.array_equal(p.Ax(r0, x), o.Ax(r0, x))


LONG_ROW = 1536  # half a 3072-entry tile: longer rows take the per-row path


def test_long_rows(tb, oracle, gpu):
    """Rows longer than half a tile (1536 entries) take the
    one-workgroup-per-row path (fixed-order tree, within 1e-13); shorter rows
    that run past their 3072-entry tile stay bit-exact."""
    n = 6000
    lens = np.zeros(n, dtype=np.int64)
    lens[:18] = [5000, 1, 0, 12000, 1024, 1025, 700, 3, 1536, 1537, 2047, 2048, 1535, 3071, 3072, 3073, 1536, 1]
    lens[18:40] = 37
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nnz = rowptr[-1]
    rng = np.random.default_rng(2)
    col = rng.integers(0, n, nnz).astype(np.int32)
    val = rng.random(nnz)
    x = rng.random(n)
    p = tb.MatrixCSR(n, n, val, rowptr, col)
    o = oracle.from_arrays("CSR", n, n, val, rowptr, col)
    y, yo = p.spmv(x), o.spmv(x)
    assert np.all(np.abs(y - yo) <= 1e-13 * np.abs(yo) + 1e-300)
    short = lens <= LONG_ROW
    assert np.array_equal(y[short], yo[short])
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    coo = tb.MatrixCOO(n, n, val, rows.astype(np.int32), col)
    yc = coo.spmv(x)
    assert np.all(np.abs(yc - yo) <= 1e-13 * np.abs(yo) + 1e-300)


def test_device_vectors_and_precisions(tb, oracle, gpu):
    """Ax on torch device tensors; FP32 / BF16 value storage with FP32
    accumulation against the FP64 oracle."""
    import torch
    n = 1 << 16
    p = tb.MatrixCSR().fill_cqmat(n, n, 16, 0.1, 42, value_seed=42)
    o = oracle.new("CSR").fill_cqmat(n, n, 16, 0.1, 42, value_seed=42)
    x = oracle.uniform_vector(1, n)
    yo = o.spmv(x)
    xd = torch.from_numpy(x).to(gpu)
    yd = torch.zeros(n, dtype=torch.float64, device=gpu)
    p.Ax_device(xd, yd)
    torch.cuda.synchronize()
    assert np.array_equal(yd.cpu().numpy(), yo)
    for vd, tol in (("f32", 2e-6), ("bf16", 2e-2)):
        y32 = torch.zeros(n, dtype=torch.float32, device=gpu)
        p.Ax_device(xd.float(), y32, value_dtype=vd)
        torch.cuda.synchronize()
        err = np.abs(y32.cpu().numpy().astype(np.float64) - yo) / np.abs(yo).max()
        assert err.max() < tol, (vd, err.max())
    e = tb.MatrixELL().fill_cqmat(n, n, 16, 0.1, 42, value_seed=42)
    y32 = torch.zeros(n, dtype=torch.float32, device=gpu)
    e.Ax_device(xd.float(), y32, value_dtype="f32")
    torch.cuda.synchronize()
    assert (np.abs(y32.cpu().numpy() - yo) / np.abs(yo).max()).max() < 2e-6


def test_ell_slot_counts_and_partial_slices(tb, gpu):
    """ELL rows of every length up to K for K around the kernel's group of
    eight slots, on row counts around its 128-row slices; random values and
    x.  Expected: each row's slots summed left to right from 0 with the
    product rounded first (MatrixELL.cpp:105-113), evaluated slot by slot in
    numpy over all rows, padding slots included."""
    rng = np.random.default_rng(11)
    for n in (1, 127, 128, 129, 1000):
        for K in (1, 7, 8, 9, 16, 17, 23, 24, 25, 40):
            lens = rng.integers(0, K + 1, n)
            lens[rng.integers(0, n)] = K
            row = np.repeat(np.arange(n, dtype=np.int32), lens)
            col = rng.integers(0, n, row.size).astype(np.int32)
            val = rng.standard_normal(row.size)
            ell = tb.MatrixELL(tb.MatrixCOO(n, n, val, row, col))
            arr = ell.arrays()
            assert ell.info()["max_col"] == K
            v = arr["values"].reshape(n, K)
            c = arr["columns"].reshape(n, K)
            x = rng.standard_normal(n)
            expect = np.zeros(n)
            for j in range(K):
                expect = expect + v[:, j] * x[np.maximum(c[:, j], 0)]
            assert np.array_equal(ell.spmv(x), expect), (n, K)


def test_ell_low_precision_slots(tb, gpu):
    """The ELL kernel with FP32 and BF16 stored values (FP32 x and y), for K
    around the 8-slot groups and row counts around the 128-row slices:
    within the storage precision of the FP64 product, relative to
    sum |a| |x| per row."""
    import torch
    rng = np.random.default_rng(12)
    for n in (127, 129, 1000):
        for K in (3, 8, 9, 17, 25):
            lens = rng.integers(0, K + 1, n)
            lens[rng.integers(0, n)] = K
            row = np.repeat(np.arange(n, dtype=np.int32), lens)
            col = rng.integers(0, n, row.size).astype(np.int32)
            val = rng.standard_normal(row.size)
            ell = tb.MatrixELL(tb.MatrixCOO(n, n, val, row, col))
            x = rng.standard_normal(n).astype(np.float32)
            ref = ell.spmv(x.astype(np.float64))
            absr = tb.MatrixELL(tb.MatrixCOO(n, n, np.abs(val), row, col)).spmv(np.abs(x).astype(np.float64))
            xd = torch.from_numpy(x).to(gpu)
            for vd, tol in (("f32", 1e-6), ("bf16", 1e-2)):
                y = torch.zeros(n, dtype=torch.float32, device=gpu)
                ell.Ax_device(xd, y, value_dtype=vd)
                torch.cuda.synchronize()
                err = np.abs(y.cpu().numpy().astype(np.float64) - ref)
                assert np.all(err <= tol * absr + 1e-30), (n, K, vd, float((err / (absr + 1e-300)).max()))


def test_empty_matrices_and_empty_row_ranges(tb, oracle, gpu):
    """Matrices with no entries (n x n and 0 x 0) leave spmv at zero and Ax
    at r in every format (MatrixCSR.cpp:130-131, MatrixELL.cpp:102-103,
    MatrixDENSE.cpp:65-66). A matrix whose entries sit in ten rows of 3000
    puts every row into one CSR entry tile, so the tile walks 3000 rows,
    almost all empty; COO, SCOO, CSR and ELL match the oracle bit for bit,
    DENSE (rows summed by a butterfly) within 1e-14 of sum |a||x|."""
    rng = np.random.default_rng(5)
    n = 3000
    x, r = rng.random(n), rng.random(n)
    none_f, none_i = np.zeros(0), np.zeros(0, dtype=np.int32)
    cases = [(n, none_f, none_i, none_i)]
    row = np.sort(rng.integers(2000, 2010, 64)).astype(np.int32)
    cases.append((n, rng.random(64), row, rng.integers(0, n, 64).astype(np.int32)))
    for size, val, row, col in cases:
        p = tb.MatrixCOO(size, size, val, row, col)
        o = oracle.from_arrays("COO", size, size, val, row, col)
        absy = np.zeros(size)
        np.add.at(absy, row, np.abs(val) * np.abs(x[col]))
        for fmt in FMTS:
            pm = p if fmt == "COO" else tb.FORMATS[fmt](p)
            om = o if fmt == "COO" else oracle.convert(o, fmt)
            y = pm.spmv(x)
            close(fmt, y, om.spmv(x), absy)
            close(fmt, pm.Ax(r, x), om.Ax(r, x), absy + np.abs(r))
            if val.size == 0:
                assert not y.any(), fmt
    zero = tb.MatrixCOO(0, 0, none_f, none_i, none_i)
    for fmt in FMTS:
        pm = zero if fmt == "COO" else tb.FORMATS[fmt](zero)
        assert pm.spmv(none_f).shape == (0,), fmt


def test_short_vectors_are_rejected(tb, gpu):
    """A vector shorter than the matrix's indices is an error, not a fault."""
    from tbsla_amd._lib import TbslaError
    m = tb.MatrixCSR().fill_cqmat(100, 100, 5, 0.5, 1)
    with pytest.raises(TbslaError):
        m.spmv(np.ones(50))
    c = tb.MatrixCOO().fill_cqmat(100, 100, 5, 0.5, 1)
    with pytest.raises(TbslaError):
        c.Ax(np.zeros(10), np.ones(100))


def test_reference_c_abi_spmv(tb, oracle, gpu, tmp_path):
    """C_CPP_Matrix<F>_* / C_CPP_Vector_* used the way the YML spmv component
    does (yml/impl/spmv.query:17-39): fill, spmv into a new vector, write it
    out (Vector.cpp:90-103 format: size_t n, n doubles)."""
    from tbsla_amd._lib import lib
    for F in ["COO", "SCOO", "CSR", "ELL", "DENSE"]:
        m = getattr(lib, f"C_CPP_Matrix{F}_create")()
        getattr(lib, f"C_CPP_Matrix{F}_fill_cqmat")(m, 400, 400, 8, 0.3, 4, 0, 1, 2, 2)
        v = lib.C_CPP_Vector_create()
        lib.C_CPP_Vector_fill(v, 400 if F == "COO" else 200, 0)
        r = getattr(lib, f"C_CPP_Matrix{F}_spmv")(m, v)
        path = str(tmp_path / f"r_{F}.bin")
        assert lib.C_CPP_Vector_write(r, path.encode())
        raw = open(path, "rb").read()
        n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
        y = np.frombuffer(raw[8:8 + 8 * n], dtype=np.float64)
        o = oracle.new(F).fill_cqmat(400, 400, 8, 0.3, 4, 0, 1, 2, 2)
        assert np.array_equal(y, o.spmv(np.arange(400 if F == "COO" else 200, dtype=np.float64))), F
        lib.C_CPP_Vector_destroy(r)
        lib.C_CPP_Vector_destroy(v)
        getattr(lib, f"C_CPP_Matrix{F}_destroy")(m)


@pytest.mark.slow
def test_full_size_block_formats(tb, oracle, gpu):
    """One rank's block (0,0) of fill_cqmat(2^24, 2^24, 16, 0.1, 42) on a 4x4
    grid, values 1 - u01(42, i), x_j = u01(1, j): GPU == oracle bit for bit
    (CSR, ELL, SCOO; COO on the rank-0 slice with a full-length x)."""
    n = 1 << 24
    for fmt in ["CSR", "ELL", "SCOO", "COO"]:
        print(f"full-size {fmt}", flush=True)
        p = tb.FORMATS[fmt]().fill_cqmat(n, n, 16, 0.1, 42, 0, 0, 4, 4, value_seed=42)
        nx = n if fmt == "COO" else p.get_ln_col()
        x = oracle.uniform_vector(1, nx)
        y = p.spmv(x)
        o = oracle.new(fmt).fill_cqmat(n, n, 16, 0.1, 42, 0, 0, 4, 4, value_seed=42)
        yo = o.spmv(x)
        assert np.array_equal(y, yo), (fmt, np.abs(y - yo).max())
        del p, o
    # DENSE: an 8192 x 8192 block of fill_cqmat(32768, ...)
    p = tb.MatrixDENSE().fill_cqmat(32768, 32768, 16, 0.1, 42, 0, 0, 4, 4, value_seed=42)
    o = oracle.new("DENSE").fill_cqmat(32768, 32768, 16, 0.1, 42, 0, 0, 4, 4, value_seed=42)
    x = oracle.uniform_vector(1, 8192)
    close("DENSE", p.spmv(x), o.spmv(x), o.spmv(np.abs(x)))


def test_cdiag_10000_config(tb, golden, gpu):
    """fill_cdiag(10000, 10000, 10) with x = u01(1, j): equal to the
    reference's own product (fixture)."""
    y = tb.MatrixCSR().fill_cdiag(10000, 10000, 10).spmv(golden["cdiag10000_x"])
    assert np.array_equal(y, golden["cdiag10000_y"])


def test_cdiag_10000_config_literal_band(tb, oracle, gpu):
    """The literal reading of the 10,000-row configuration: ten diagonals at
    offsets 0..9 (columns i..i+9 clipped to the matrix), values 1.0, built
    through the MatrixCSR array constructor (99,955 entries).  Equal to the
    oracle, and to the closed form sum(x[i:i+10]) summed left to right."""
    n = 10000
    cols = [np.arange(i, min(i + 10, n)) for i in range(n)]
    rowptr = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int32)
    colidx = np.concatenate(cols).astype(np.int32)
    assert rowptr[-1] == 99955
    val = np.ones(rowptr[-1])
    x = oracle.uniform_vector(1, n)
    y = tb.MatrixCSR(n, n, val, rowptr, colidx).spmv(x)
    assert np.array_equal(y, oracle.from_arrays("CSR", n, n, val, rowptr, colidx).spmv(x))
    closed = np.zeros(n)
    for k in range(10):
        closed[: n - k] += x[k:]
    assert np.array_equal(y, closed)
