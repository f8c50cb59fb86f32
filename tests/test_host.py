"""Host side of libtbsla_hip.so (no GPU): generators, conversions, binary I/O
and the C-ABI surface.  Generators and conversions must be bit-identical to
the oracle (itself pinned to the reference) and to the reference digests."""
import hashlib
import os
import re

import numpy as np
import pytest

from conftest import BLOCKS, CDIAG_SHAPES, CQMAT_SHAPES, FMTS, INFO_KEYS, ROOT, grid_q


def same_matrix(prod, orc, fmt, check_nnz=True):
    a = prod.arrays()
    oi, oa = orc.arrays()
    info = prod.info()
    for k in INFO_KEYS:
        assert info[k] == oi[k], (fmt, k, info[k], oi[k])
    if fmt == "ELL":
        assert info["max_col"] == oi["max_col"]
    if check_nnz:
        assert info["nnz"] == oi["nnz"], (fmt, info["nnz"], oi["nnz"])
    assert set(a) == set(oa)
    for k in oa:
        assert a[k].dtype == oa[k].dtype and np.array_equal(a[k], oa[k], equal_nan=True), (fmt, k)


def test_cqmat_grid_bit_exact(tb, oracle):
    for (nr, nc, c) in CQMAT_SHAPES:
        for fmt in FMTS:
            for blk in BLOCKS:
                for s in range(4):
                    for q in grid_q()[::3]:
                        p = tb.FORMATS[fmt]().fill_cqmat(nr, nc, c, q, s, *blk)
                        o = oracle.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk)
                        same_matrix(p, o, fmt)


def test_cdiag_grid_bit_exact(tb, oracle):
    for (nr, nc, c) in CDIAG_SHAPES:
        for fmt in FMTS:
            for blk in BLOCKS:
                p = tb.FORMATS[fmt]().fill_cdiag(nr, nc, c, *blk)
                o = oracle.new(fmt).fill_cdiag(nr, nc, c, *blk)
                # the reference never sets nnz in CSR/SCOO fill_cdiag
                same_matrix(p, o, fmt, check_nnz=fmt not in ("CSR", "SCOO"))


def test_value_overlay_and_stochastic(tb, oracle):
    for (nr, nc, c, q, s) in [(300, 300, 16, 0.1, 42), (257, 300, 12, 0.3, 5), (300, 250, 8, 0.5, 7)]:
        for fmt in FMTS:
            for blk in BLOCKS:
                p = tb.FORMATS[fmt]().fill_cqmat(nr, nc, c, q, s, *blk, value_seed=42)
                o = oracle.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk, value_seed=42)
                same_matrix(p, o, fmt)
                p = tb.FORMATS[fmt]().fill_cqmat_stochastic(nr, nc, c, q, s, *blk)
                o = oracle.new(fmt).fill_cqmat_stochastic(nr, nc, c, q, s, *blk)
                same_matrix(p, o, fmt)
                assert p.count_stochastic()
                p = tb.FORMATS[fmt]().fill_cqmat(nr, nc, c, q, s, *blk, value_seed=7)
                o = oracle.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk, value_seed=7)
                if o.arrays()[1]["values"].size:
                    p.normalize_columns()
                    o.normalize_columns()
                    same_matrix(p, o, fmt)
                    assert not p.count_stochastic()


def test_conversions(tb, oracle):
    for (nr, nc, c, q, s) in [(30, 30, 8, 0.5, 3), (50, 40, 5, 1.0, 1), (64, 64, 12, 0.2, 9)]:
        pc = tb.MatrixCOO().fill_cqmat(nr, nc, c, q, s, value_seed=11)
        oc = oracle.new("COO").fill_cqmat(nr, nc, c, q, s, value_seed=11)
        for fmt, cls in [("SCOO", tb.MatrixSCOO), ("CSR", tb.MatrixCSR), ("ELL", tb.MatrixELL),
                         ("DENSE", tb.MatrixDENSE)]:
            same_matrix(cls(pc), oracle.convert(oc, fmt), fmt)


def test_array_constructors(tb, golden):
    m = tb.MatrixCSR(5, 5, golden["ka_values"], golden["ka_rowptr"], golden["ka_colidx"])
    a = m.arrays()
    assert np.array_equal(a["rowptr"], golden["ka_rowptr"]) and m.get_n_row() == 5
    m = tb.MatrixCOO(5, 5, golden["ka_values"], golden["ka_row"], golden["ka_colidx"])
    assert np.array_equal(m.arrays()["row"], golden["ka_row"])


def test_binary_io_matches_oracle_format(tb, oracle, tmp_path):
    """write() appends the reference binary layout; read() restores it; the
    oracle's reader (written from MatrixCSR.cpp:206-232) reads our files."""
    for fmt in FMTS:
        p = tb.FORMATS[fmt]().fill_cqmat(40, 40, 6, 0.3, 2, 1, 0, 2, 2, value_seed=3)
        f = str(tmp_path / f"m.{fmt}")
        p.write(f)
        p.write(f)  # appended twice, like C_CPP_Matrix<F>_write
        size1 = os.path.getsize(f) // 2
        q = tb.FORMATS[fmt]().read(f, size1)
        same_matrix(q, oracle.read(fmt, f, 0), fmt)
        same_matrix(p, oracle.read(fmt, f, size1), fmt)


MM_GENERAL = """%%MatrixMarket matrix coordinate real general
% comment lines are skipped
%
3 4 5
1 1 2.5
3 4 -1
2 2 0.125
1 4 7
3 1 1e-3
"""
MM_SYMMETRIC = """%%MatrixMarket matrix coordinate real symmetric
4 4 4
1 1 1.5
3 1 2
4 2 -3
4 4 0.5
"""


def test_read_matrix_market(tb, tmp_path):
    """readMM (MatrixCOO.cpp:230-292): 1-based indices in file order, and for
    'symmetric' each off-diagonal entry followed by its mirror."""
    from tbsla_amd import _lib
    f = tmp_path / "g.mtx"
    f.write_text(MM_GENERAL)
    m = tb.MatrixCOO().readMM(str(f))
    a, info = m.arrays(), m.info()
    assert (info["n_row"], info["n_col"], info["nnz"]) == (3, 4, 5)
    assert a["row"].tolist() == [0, 2, 1, 0, 2] and a["col"].tolist() == [0, 3, 1, 3, 0]
    assert a["values"].tolist() == [2.5, -1.0, 0.125, 7.0, 1e-3]
    f = tmp_path / "s.mtx"
    f.write_text(MM_SYMMETRIC)
    m = tb.MatrixCOO().readMM(str(f))
    a, info = m.arrays(), m.info()
    assert (info["n_row"], info["n_col"], info["nnz"]) == (4, 4, 6)
    assert a["row"].tolist() == [0, 2, 0, 3, 1, 3] and a["col"].tolist() == [0, 0, 2, 1, 3, 3]
    assert a["values"].tolist() == [1.5, 2.0, 2.0, -3.0, -3.0, 0.5]
    for header in ("%%MatrixMarket matrix array real general", "%%MatrixMarket matrix coordinate complex general",
                   "%%MatrixMarket matrix coordinate real hermitian"):
        f.write_text(header + "\n1 1 1\n1 1 1\n")
        with pytest.raises(_lib.TbslaError) as e:
            tb.MatrixCOO().readMM(str(f))
        assert e.value.code == -1
    with pytest.raises(_lib.TbslaError) as e:
        tb.MatrixCOO().readMM(str(tmp_path / "missing.mtx"))
    assert e.value.code == -3
    f.write_text(MM_GENERAL)
    assert _lib.lib.tbsla_matrix_read_mm(tb.MatrixCSR()._h, str(f).encode()) == -1


def test_read_matrix_market_matches_reference(tb, reference, tmp_path):
    """The same files through the compiled reference's own readMM."""
    for name, text in (("g", MM_GENERAL), ("s", MM_SYMMETRIC)):
        f = tmp_path / f"{name}.mtx"
        f.write_text(text)
        same_matrix(tb.MatrixCOO().readMM(str(f)), reference.read_mm(str(f)), "COO")
    f = tmp_path / "bad.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real skew-symmetric\n1 1 1\n1 1 1\n")
    with pytest.raises(ValueError):
        reference.read_mm(str(f))


@pytest.mark.slow
def test_full_size_generators_match_reference_digests(tb, digests):
    cases = {
        "cqmat2^24c16q0.1s42/CSR/block0.0.4.4": lambda: tb.MatrixCSR().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4),
        "cqmat2^24c16q0.1s42/ELL/block0.0.4.4": lambda: tb.MatrixELL().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4),
        "cqmat2^24c16q0.1s42/SCOO/block0.0.4.4": lambda: tb.MatrixSCOO().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4),
        "cqmat2^24c16q0.1s42/COO/rank0of16": lambda: tb.MatrixCOO().fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4),
        "cqmat32768c16q0.1s42/DENSE/block0.0.4.4": lambda: tb.MatrixDENSE().fill_cqmat(32768, 32768, 16, 0.1, 42, 0, 0, 4, 4),
        "stoch2^24c8q0.5s7/CSR": lambda: tb.MatrixCSR().fill_cqmat_stochastic(2 ** 24, 2 ** 24, 8, 0.5, 7),
        "stoch2^22c16q0.3s1256/CSR": lambda: tb.MatrixCSR().fill_cqmat_stochastic(2 ** 22, 2 ** 22, 16, 0.3, 1256),
        "stoch2^20c16q0.2s3/CSR": lambda: tb.MatrixCSR().fill_cqmat_stochastic(2 ** 20, 2 ** 20, 16, 0.2, 3),
        "stoch262144c12q0.3s1003/CSR": lambda: tb.MatrixCSR().fill_cqmat_stochastic(262144, 262144, 12, 0.3, 1003),
        "cdiag10000/CSR": lambda: tb.MatrixCSR().fill_cdiag(10000, 10000, 10),
    }
    for key, make in cases.items():
        m = make()
        ref = digests["big"][key]
        info = m.info()
        for k in INFO_KEYS:
            assert info[k] == ref[k], (key, k)
        if not key.startswith("cdiag"):
            assert info["nnz"] == ref["nnz"], key
        for k, a in m.arrays().items():
            assert a.size == ref[f"len_{k}"], (key, k)
            assert hashlib.sha256(a.tobytes()).hexdigest() == ref[f"sha256_{k}"], (key, k)
        del m


# ----------------------------------------------------------------- the C ABI

def declared_symbols():
    names = set()
    for h in ("tbsla_hip.h", "tbsla_cpp_c.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(tbsla_(?:matrix|hip|batch)_\w+|tbsla_abi_version|tbsla_last_error)\s*\(", txt))
        for F in re.findall(r"TBSLA_C_MATRIX_DECL\((\w+)\)\n", txt):
            for op in ("create", "destroy", "fill_cdiag", "fill_cqmat", "read", "write", "print", "spmv"):
                names.add(f"C_CPP_Matrix{F}_{op}")
        names |= set(re.findall(r"\b(C_CPP_Vector_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol(tb):
    from tbsla_amd import _lib
    names = declared_symbols()
    assert len(names) >= 75
    for n in sorted(names):
        assert hasattr(_lib.lib, n), n
    assert set(_lib.SIGNATURES) == names
    assert _lib.lib.tbsla_abi_version() == 2


def test_c_abi_errors_without_device_calls(tb):
    from tbsla_amd import _lib
    assert _lib.lib.tbsla_matrix_create(9) is None
    assert b"unknown format" in _lib.lib.tbsla_last_error()
    m = tb.MatrixCSR()
    with pytest.raises(_lib.TbslaError):
        m.fill_cqmat(10, 10, 3, 0.1, 1, 2, 0, 2, 2)  # pr >= NR
    # the GPU generation and conversion entry points validate before any device call
    with pytest.raises(_lib.TbslaError):
        m.fill_cqmat(10, 10, 3, 0.1, 1, 2, 0, 2, 2, device=True)  # pr >= NR
    with pytest.raises(_lib.TbslaError):
        tb.MatrixDENSE().fill_cqmat(10, 10, 3, 0.1, 1, device=True)  # not a GPU generator
    assert _lib.lib.tbsla_matrix_convert_gpu(m._h, 2) is None  # needs a COO
    assert b"needs a COO" in _lib.lib.tbsla_last_error()
    coo = tb.MatrixCOO().fill_cqmat(10, 10, 3, 0.1, 1)
    assert _lib.lib.tbsla_matrix_convert_gpu(coo._h, 4) is None  # DENSE: not a GPU conversion
    assert b"CSR or ELL" in _lib.lib.tbsla_last_error()
    # PageRank parameters are checked before any device call (the output
    # pointer is never touched): precision (TBSLA_PR_PATTERN_F32 = 4 is the
    # last), stopping rule, product structure
    import ctypes
    from tbsla_amd.matrix import dtype_code
    assert dtype_code("pattern_f32") == _lib.PR_PATTERN_F32 == 4
    einval = -1  # TBSLA_EINVAL
    g = tb.MatrixCSR().fill_cqmat_stochastic(50, 50, 4, 0.3, 1)
    for field, bad in (("precision", 5), ("precision", -1), ("rule", 2), ("gather", 3)):
        prm = _lib.PageRankParams(epsilon=1e-10, beta=0.85, max_iteration=10, precision=0, rule=1, check_every=8,
                                  anchor_tau=0.0, nodes=None, n_nodes=0, gather=0)
        setattr(prm, field, bad)
        rc = _lib.lib.tbsla_hip_page_rank_ex(g._h, ctypes.byref(prm), None, 0, None, 0, ctypes.c_void_p(8), 50,
                                             None, None)
        assert rc == einval, (field, bad, rc)
    # an output shorter than n_col is refused before any launch (the pointer
    # is a dummy that no kernel may write through)
    prm = _lib.PageRankParams(epsilon=1e-10, beta=0.85, max_iteration=10, precision=0, rule=1, check_every=8,
                              anchor_tau=0.0, nodes=None, n_nodes=0, gather=0)
    for fn, args in ((_lib.lib.tbsla_hip_page_rank_ex, (None, 0, None, 0, ctypes.c_void_p(8), 49, None, None)),
                     (_lib.lib.tbsla_hip_page_rank, (ctypes.cast(ctypes.c_void_p(8), _lib._dp), 49, None, None))):
        assert fn(g._h, ctypes.byref(prm), *args) == einval
        assert b"PageRank writes 50" in _lib.lib.tbsla_last_error()
    v = _lib.lib.C_CPP_Vector_create()
    _lib.lib.C_CPP_Vector_fill(v, 5, 3)
    assert not _lib.lib.C_CPP_Vector_add(v, None)
    _lib.lib.C_CPP_Vector_destroy(v)
    _lib.lib.C_CPP_MatrixCSR_destroy(None)


def test_batch_workload_generator(tb, oracle):
    """The batch workload's counter generator equals the value-overlay
    generator (u01), and the graph sizes / teleport nodes are in range."""
    from tbsla_amd import batch
    x = oracle.uniform_vector(1000, 8)
    assert [batch.u01(1000, i) for i in range(8)] == list(x)
    assert np.array_equal(batch.u01_array(7, 0, 1 << 16), oracle.uniform_vector(7, 1 << 16))
    u = oracle.uniform_vector(3, 2000)
    w = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])
    assert np.array_equal(batch.normal_array(3, 1000), w)
    w = batch.personalised_workload()
    assert len(w) == 257
    for (n, c, q, seed, sets) in w[:-1]:
        assert 1024 <= n <= 1 << 18 and c == 12 and q == 0.3 and len(sets) == 1 and 0 <= sets[0][0] < n
    n, c, q, seed, sets = w[-1]
    assert (n, c, q, seed, len(sets)) == (1 << 22, 16, 0.3, 1256, 64)
    assert all(0 <= s[0] < n for s in sets)


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU's block; --gpus N > 1 exits with a message."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_consumers_link_and_start():
    """The C++ and C consumers built by __graft_entry__.build() load
    libtbsla_hip.so and reject bad arguments without touching a GPU."""
    import subprocess
    cons = os.path.join(ROOT, "tests", "consumers")
    for exe, args in (("tbsla_perf_gpu", ["--format", "XYZ"]), ("c_consumer", [])):
        path = os.path.join(cons, exe)
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", cons], check=True)
        r = subprocess.run([path] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, (exe, r.returncode, r.stderr)
        assert "unrecognized" in r.stderr or "usage" in r.stderr
