"""Programs built against the public interfaces, run as separate processes:
tests/consumers/tbsla_perf_gpu (the C++ API through the tbsla::cpp::Matrix
base, as the reference's src/cpp/apps/tbsla_perf.cpp:92-143 uses it) and
tests/consumers/c_consumer (the C ABI, as the reference's YML components
use it).  Their outputs are compared with the oracle and the reference
fixtures."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import FMTS, GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CONS = os.path.join(ROOT, "tests", "consumers")
EXACT = {"COO", "SCOO", "CSR", "ELL"}


def vec_write(path, x):
    with open(path, "wb") as f:
        f.write(np.uint64(x.size).tobytes())
        f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())


def vec_read(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    return np.frombuffer(raw[8:8 + 8 * n], dtype=np.float64).copy()


def perf(tmp_path, **kw):
    exe = os.path.join(CONS, "tbsla_perf_gpu")
    assert os.path.exists(exe), "build the consumers (__graft_entry__.build())"
    out = str(tmp_path / "out.vec")
    args = [exe]
    for k, v in kw.items():
        args += [f"--{k}", str(v)]
    args += ["--out", out]
    print("run", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1]), vec_read(out)


def test_cpp_consumer_products(oracle, gpu, tmp_path):
    """spmv, Ax and a_axpx_ through the virtual base on every format, block
    layouts included, equal to the oracle."""
    n = 400
    for fmt in FMTS:
        o = oracle.new(fmt).fill_cqmat(n, n, 8, 0.3, 4, 1, 0, 2, 2)
        nx = n if fmt == "COO" else o.info["ln_col"]
        x = oracle.uniform_vector(2, nx)
        xf = str(tmp_path / "x.vec")
        vec_write(xf, x)
        common = dict(format=fmt, matrix="cqmat", NR=n, NC=n, C=8, Q=0.3, S=4, pr=1, pc=0, GR=2, GC=2, x=xf)
        info, y = perf(tmp_path, op="spmv", **common)
        assert info["n_row"] == n and info["nnz"] == o.info["nnz"] and info["test"] == "spmv"
        ref = o.spmv(x)
        assert np.array_equal(y, ref) if fmt in EXACT else np.allclose(y, ref, rtol=1e-14, atol=0), fmt
        _, y = perf(tmp_path, op="Ax", **common)
        ref = o.Ax(np.zeros(n), x)
        assert np.array_equal(y, ref) if fmt in EXACT else np.allclose(y, ref, rtol=1e-14, atol=0), fmt
        xg = oracle.uniform_vector(3, n)
        vec_write(xf, xg)
        _, y = perf(tmp_path, op="a_axpx", **common)
        ref = o.a_axpx_(xg)
        assert np.array_equal(y, ref) if fmt in EXACT else np.allclose(y, ref, rtol=1e-13, atol=0), fmt


def test_cpp_consumer_page_rank(oracle, gpu, tmp_path):
    """page_rank and personalized_page_rank through the virtual base: the
    oracle's iteration count (+-1) and vector (L1 <= 1e-12)."""
    n = 3000
    o = oracle.new("CSR").fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    for fmt in ["CSR", "ELL", "COO"]:
        common = dict(format=fmt, matrix="stochastic", NR=n, NC=n, C=8, Q=0.5, S=7, epsilon=1e-10, beta=0.85,
                      max_iteration=500)
        info, pr = perf(tmp_path, op="page_rank", **common)
        ref, it = o.page_rank(1e-10, 0.85, 500)
        assert abs(info["iterations"] - it) <= 1 and np.abs(pr - ref).sum() <= 1e-12, fmt
        info, pr = perf(tmp_path, op="personalized_page_rank", nodes="1,1000,2999", **common)
        ref, it = o.personalized_page_rank(1e-10, 0.85, 500, np.array([1, 1000, 2999], dtype=np.int32))
        assert abs(info["iterations"] - it) <= 1 and np.abs(pr - ref).sum() <= 1e-12, fmt


def test_cpp_consumer_reads_reference_files(golden, digests, gpu, tmp_path):
    """read() of a file written by the reference, then spmv: the reference's
    own product of that matrix."""
    d = os.path.join(GOLDEN, "ref_binary")
    for fn in ["cqmat40c6q0.3s2_block1.0.2.2.CSR", "stoch50c5q0.5s3.ELL", "cdiag30c4_block1.2.2.3.SCOO"]:
        xf = str(tmp_path / "x.vec")
        vec_write(xf, np.arange(digests["binary"][fn]["nx"], dtype=np.float64))
        _, y = perf(tmp_path, format=fn.rsplit(".", 1)[1], matrix=os.path.join(d, fn), op="spmv", x=xf)
        assert np.array_equal(y, golden[f"binary/{fn}/spmv_iota"]), fn


def test_c_consumer(oracle, gpu, tmp_path):
    """The C program: fill_cqmat, write, read into a new handle, read x,
    spmv, write y -- equal to the oracle."""
    exe = os.path.join(CONS, "c_consumer")
    assert os.path.exists(exe), "build the consumers (__graft_entry__.build())"
    n = 600
    for fmt in FMTS:
        o = oracle.new(fmt).fill_cqmat(n, n, 8, 0.2, 9, 0, 1, 2, 2)
        nx = n if fmt == "COO" else o.info["ln_col"]
        x = oracle.uniform_vector(4, nx)
        xf, yf, mf = (str(tmp_path / f"{k}.{fmt}") for k in ("x", "y", "m"))
        vec_write(xf, x)
        print("c_consumer", fmt, flush=True)
        r = subprocess.run([exe, fmt, str(n), str(n), "8", "0.2", "9", "0", "1", "2", "2", mf, xf, yf],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        y, ref = vec_read(yf), o.spmv(x)
        assert np.array_equal(y, ref) if fmt in EXACT else np.allclose(y, ref, rtol=1e-14, atol=0), fmt
