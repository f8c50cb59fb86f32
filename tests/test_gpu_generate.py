"""GPU generation of cqmat COO, SCOO, CSR and ELL matrices
(tbsla_matrix_fill_cqmat_gpu) and GPU COO -> CSR / ELL conversion
(tbsla_matrix_convert_gpu).

The device restates the reference's entry stream (values_generation.cpp:26-45)
with its own restatement of glibc's rand_r.  Every array must equal the host
generator's, which is pinned to the reference itself, and at full size the
arrays must hash to the digests of the reference-generated matrices."""
import hashlib
import time

import numpy as np
import pytest

from conftest import BLOCKS, CQMAT_SHAPES, INFO_KEYS, grid_q

pytestmark = pytest.mark.gpu


def same(a, b):
    ia, ib = a.info(), b.info()
    for k in INFO_KEYS + ["nnz"]:
        assert ia[k] == ib[k], (k, ia[k], ib[k])
    xa, xb = a.arrays(), b.arrays()
    assert set(xa) == set(xb)
    for k in xa:
        assert xa[k].dtype == xb[k].dtype and np.array_equal(xa[k], xb[k]), k


GPU_FMTS = ["CSR", "ELL", "SCOO", "COO"]


@pytest.mark.parametrize("fmt", GPU_FMTS)
def test_gpu_cqmat_grid_equals_host(tb, gpu, fmt):
    """The reference test grid (shapes, q in 0..1, seeds 0..3, 2D blocks), with
    and without the value overlay, plus shapes where the partial rows and the
    c > n_col clipping matter."""
    shapes = list(CQMAT_SHAPES) + [(40, 17, 6), (17, 40, 6), (30, 30, 31), (64, 64, 1), (9, 9, 9)]
    cls = tb.FORMATS[fmt]
    for (nr, nc, c) in shapes:
        for s in range(4):
            for q in grid_q()[::3] + [1.0]:
                for blk in BLOCKS:
                    for vs in (0, 7):
                        h = cls().fill_cqmat(nr, nc, c, q, s, *blk, value_seed=vs)
                        d = cls().fill_cqmat(nr, nc, c, q, s, *blk, value_seed=vs, device=True)
                        same(h, d)
                    if c > 0:
                        h = cls().fill_cqmat_stochastic(nr, nc, c, q, s, *blk)
                        d = cls().fill_cqmat_stochastic(nr, nc, c, q, s, *blk, device=True)
                        same(h, d)
                        assert d.count_stochastic() == h.count_stochastic(), (fmt, nr, nc, c, s, q, blk)


@pytest.mark.parametrize("fmt", GPU_FMTS)
def test_gpu_cqmat_larger_blocks_equal_host(tb, gpu, fmt):
    cls = tb.FORMATS[fmt]
    for (n, c, q, s, blk) in [(100000, 16, 0.1, 42, (1, 2, 4, 4)), (70000, 12, 0.3, 1001, (0, 0, 1, 1)),
                              (50000, 8, 0.5, 7, (2, 1, 3, 3))]:
        same(cls().fill_cqmat(n, n, c, q, s, *blk, value_seed=42),
             cls().fill_cqmat(n, n, c, q, s, *blk, value_seed=42, device=True))
        same(cls().fill_cqmat_stochastic(n, n, c, q, s, *blk),
             cls().fill_cqmat_stochastic(n, n, c, q, s, *blk, device=True))


def test_gpu_generation_rejects_dense(tb, gpu):
    from tbsla_amd import _lib
    with pytest.raises(_lib.TbslaError):
        tb.MatrixDENSE().fill_cqmat(10, 10, 3, 0.1, 1, device=True)
    with pytest.raises(_lib.TbslaError):
        tb.MatrixCSR().fill_cqmat(10, 10, 3, 0.1, 1, 2, 0, 2, 2, device=True)  # pr >= NR


@pytest.mark.slow
def test_gpu_generation_matches_reference_digests(tb, digests, gpu):
    """The benchmark block and the PageRank graphs, generated on the GPU, hash
    to the digests of the matrices the compiled reference generated."""
    cases = {
        "cqmat2^24c16q0.1s42/CSR/block0.0.4.4": lambda: tb.MatrixCSR().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4, device=True),
        "stoch2^24c8q0.5s7/CSR": lambda: tb.MatrixCSR().fill_cqmat_stochastic(2 ** 24, 2 ** 24, 8, 0.5, 7, device=True),
        "stoch2^22c16q0.3s1256/CSR": lambda: tb.MatrixCSR().fill_cqmat_stochastic(2 ** 22, 2 ** 22, 16, 0.3, 1256, device=True),
        "stoch2^20c16q0.2s3/CSR": lambda: tb.MatrixCSR().fill_cqmat_stochastic(2 ** 20, 2 ** 20, 16, 0.2, 3, device=True),
        "cqmat2^24c16q0.1s42/ELL/block0.0.4.4": lambda: tb.MatrixELL().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4, device=True),
        "cqmat2^24c16q0.1s42/SCOO/block0.0.4.4": lambda: tb.MatrixSCOO().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4, device=True),
        "cqmat2^24c16q0.1s42/COO/rank0of16": lambda: tb.MatrixCOO().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4, device=True),
    }
    for key, make in cases.items():
        t = time.time()
        m = make()
        dt = time.time() - t
        ref = digests["big"][key]
        info = m.info()
        for k in INFO_KEYS + ["nnz"]:
            assert info[k] == ref[k], (key, k)
        for k, a in m.arrays().items():
            assert a.size == ref[f"len_{k}"], (key, k)
            assert hashlib.sha256(a.tobytes()).hexdigest() == ref[f"sha256_{k}"], (key, k)
        print(f"{key}: generated on the GPU in {dt:.2f} s")
        del m


# ---------------------------------------------------------------- conversions

def _digest(m, fmt):
    import json
    info, arrs = m.info(), m.arrays()
    h = hashlib.sha256()
    h.update(json.dumps([info[k] for k in INFO_KEYS] + [info["max_col"] if fmt == "ELL" else -1]).encode())
    for k in sorted(arrs):
        h.update(k.encode())
        h.update(np.ascontiguousarray(arrs[k]).tobytes())
    return h.hexdigest()


def test_gpu_conversion_matches_reference_digests(tb, digests, gpu):
    """COO -> CSR / ELL with the (row, column) sort on the GPU hashes to the
    digests of the compiled reference's own conversions (MatrixCSR.cpp:29-74,
    MatrixELL.cpp:12-54) over its test grid."""
    d = digests["conversions"]
    for (nr, nc, c) in [(10, 10, 3), (30, 30, 8), (20, 30, 4), (30, 20, 6), (50, 50, 5)]:
        for s in range(4):
            for q in [0.0, 0.1, 0.5, 1.0]:
                coo = tb.MatrixCOO().fill_cqmat(nr, nc, c, q, s)
                for fmt, cls in (("CSR", tb.MatrixCSR), ("ELL", tb.MatrixELL)):
                    assert _digest(cls(coo, device=True), fmt) == d[f"convert/{nr}x{nc}c{c}s{s}q{q}/{fmt}"], (nr, nc, c, s, q, fmt)


def test_gpu_conversion_equals_host(tb, gpu):
    """Unsorted COO with duplicate (row, column) pairs and distinct values:
    ties keep their input order on both paths; empty rows; block COO slices."""
    from tbsla_amd import _lib
    rng = np.random.default_rng(4)
    n = 5000
    nnz = 60000
    row = rng.integers(0, n, nnz).astype(np.int32)
    col = rng.integers(0, n // 3, nnz).astype(np.int32)  # many duplicates
    row[:500] = 7  # one long row
    val = rng.random(nnz)
    coo = tb.MatrixCOO(n, n, val, row, col)
    for cls in (tb.MatrixCSR, tb.MatrixELL):
        same(cls(coo), cls(coo, device=True))
    for blk in [(0, 0, 2, 2), (1, 3, 2, 4)]:
        coo = tb.MatrixCOO().fill_cqmat(4000, 4000, 9, 0.3, 5, *blk, value_seed=3)
        for cls in (tb.MatrixCSR, tb.MatrixELL):
            same(cls(coo), cls(coo, device=True))
    # negative column indices (the sort key then holds all 32 column bits,
    # sign flipped) and a row count that needs several 8-bit digit passes
    n2 = 70000
    row = rng.integers(0, n2, 200000).astype(np.int32)
    col = rng.integers(-300, 300, 200000).astype(np.int32)
    coo = tb.MatrixCOO(n2, n2, rng.random(200000), row, col)
    for cls in (tb.MatrixCSR, tb.MatrixELL):
        same(cls(coo), cls(coo, device=True))
    empty = tb.MatrixCOO(10, 10, np.zeros(0), np.zeros(0, np.int32), np.zeros(0, np.int32))
    for cls in (tb.MatrixCSR, tb.MatrixELL):
        same(cls(empty), cls(empty, device=True))
    bad = tb.MatrixCOO(5, 5, np.ones(2), np.array([0, 5], np.int32), np.array([0, 1], np.int32))
    with pytest.raises(_lib.TbslaError):  # row 5 outside [0, 5)
        tb.MatrixCSR(bad, device=True)


@pytest.mark.slow
def test_gpu_conversion_full_size(tb, gpu):
    """The COO slice of the benchmark matrix (16.8 M entries) to CSR and ELL."""
    coo = tb.MatrixCOO().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4, value_seed=42)
    for cls in (tb.MatrixCSR, tb.MatrixELL):
        t = time.time()
        h = cls(coo)
        th = time.time() - t
        t = time.time()
        g = cls(coo, device=True)
        tg = time.time() - t
        same(h, g)
        print(f"{cls.__name__} from a 16.8 M-entry COO: host {th:.2f} s, GPU {tg:.2f} s")
        del h, g
