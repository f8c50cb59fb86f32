"""Generate the golden fixtures from the compiled TBSLA reference.

Run here (where /root/reference exists) after ``make -C oracle ref``:

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py [--big]

Writes, next to this script:

* ``golden_small.npz``  -- small full vectors (SpMV and PageRank outputs).
* ``digests.json``      -- SHA-256 of every generated index/value array for the
                           reference's own test grids (tests/test_spmv_cqmat.cpp,
                           tests/test_spmv_cdiag.cpp) and, with ``--big``, for
                           the full-size workloads (one rank's block of the
                           n=2^24 matrix, the PageRank graphs, ...).
* ``ref_binary/``       -- small matrix files written by the reference's own
                           write() (MatrixCSR.cpp:179-204 and siblings), one
                           per format and case, for file interchange tests.
* with ``--cfg3``: the FP64 PageRank limit of
  fill_cqmat_stochastic(2^24, 2^24, 8, 0.5, 7), beta 0.85 -- the reference
  iteration run into its 2-cycle, both cycle states as SHA-256 digests and
  sampled entries (digests.json "pagerank_cfg3", golden_small.npz).

Every number here comes from the reference itself (oracle/_ref), never from the
restatement under test.  OMP_NUM_THREADS=1 makes the COO/SCOO vector reduction
order the sequential one.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import FORMATS, Reference as R  # noqa: E402

FMT_ORDER = ["COO", "SCOO", "CSR", "ELL", "DENSE"]
INFO_KEYS = ["n_row", "n_col", "ln_row", "ln_col", "f_row", "f_col", "pr", "pc", "NR", "NC"]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def grid_q():
    """q values of tests/test_spmv_cqmat.cpp:155 (accumulated 0.1 steps)."""
    out, q = [], 0.0
    while q <= 1:
        out.append(q)
        q += 0.1
    return out


CQMAT_SHAPES = ([(10, 5, 8), (5, 10, 7), (10, 10, 12), (10, 10, 3), (10, 5, 3), (10, 5, 6)]
                + [(10, 10, i) for i in range(13)] + [(5, 10, i) for i in range(13)]
                + [(10, 5, i) for i in range(13)] + [(30, 30, 2 * i) for i in range(13)]
                + [(20, 30, 2 * i) for i in range(13)] + [(30, 20, 2 * i) for i in range(13)])
CDIAG_SHAPES = ([(10, 10, i) for i in range(13)] + [(5, 10, i) for i in range(13)]
                + [(10, 5, i) for i in range(13)] + [(30, 30, 2 * i) for i in range(13)]
                + [(20, 30, 2 * i) for i in range(13)] + [(30, 20, 2 * i) for i in range(13)])
BLOCKS = [(0, 0, 1, 1), (1, 0, 2, 2), (1, 2, 2, 3)]


def matrix_digest(m, fmt, h):
    """Fold one matrix (arrays + deterministic header fields) into hash h."""
    info, arrs = m.arrays()
    h.update(json.dumps([info[k] for k in INFO_KEYS] + [info.get("max_col", -1)]).encode())
    for k in sorted(arrs):
        h.update(k.encode())
        h.update(np.ascontiguousarray(arrs[k]).tobytes())
    return info


def small_grids():
    out = {"cqmat": {}, "cdiag": {}}
    for (nr, nc, c) in CQMAT_SHAPES:
        for fmt in FMT_ORDER:
            for blk in BLOCKS:
                h = hashlib.sha256()
                for s in range(4):
                    for q in grid_q():
                        m = R.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk)
                        info = matrix_digest(m, fmt, h)
                        x = np.arange(info["ln_col"] if fmt != "COO" else nc, dtype=np.float64)
                        h.update(m.spmv(x).tobytes())
                out["cqmat"][f"{nr}x{nc}c{c}/{fmt}/{'.'.join(map(str, blk))}"] = h.hexdigest()
    for (nr, nc, c) in CDIAG_SHAPES:
        for fmt in FMT_ORDER:
            for blk in BLOCKS:
                h = hashlib.sha256()
                m = R.new(fmt).fill_cdiag(nr, nc, c, *blk)
                info = matrix_digest(m, fmt, h)
                x = np.arange(info["ln_col"] if fmt != "COO" else nc, dtype=np.float64)
                h.update(m.spmv(x).tobytes())
                out["cdiag"][f"{nr}x{nc}c{c}/{fmt}/{'.'.join(map(str, blk))}"] = h.hexdigest()
    return out


def conversions_and_stochastic():
    """Digests of X(const MatrixCOO&), fill_cqmat_stochastic, normalize_columns."""
    out = {}
    for (nr, nc, c) in [(10, 10, 3), (30, 30, 8), (20, 30, 4), (30, 20, 6), (50, 50, 5)]:
        for s in range(4):
            for q in [0.0, 0.1, 0.5, 1.0]:
                coo = R.new("COO").fill_cqmat(nr, nc, c, q, s)
                for fmt in ["SCOO", "CSR", "ELL", "DENSE"]:
                    h = hashlib.sha256()
                    matrix_digest(R.convert(coo, fmt), fmt, h)
                    out[f"convert/{nr}x{nc}c{c}s{s}q{q}/{fmt}"] = h.hexdigest()
                for fmt in FMT_ORDER:
                    for blk in [(0, 0, 1, 1), (1, 1, 2, 2)]:
                        h = hashlib.sha256()
                        matrix_digest(R.new(fmt).fill_cqmat_stochastic(nr, nc, c, q, s, *blk), fmt, h)
                        out[f"stochastic/{nr}x{nc}c{c}s{s}q{q}/{fmt}/{blk}"] = h.hexdigest()
                        m = R.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk)
                        if m.arrays()[1]["values"].size == 0:
                            continue  # the reference's normalize_columns runs off an empty rowptr
                        m.normalize_columns()
                        h = hashlib.sha256()
                        matrix_digest(m, fmt, h)
                        out[f"normalize/{nr}x{nc}c{c}s{s}q{q}/{fmt}/{blk}"] = h.hexdigest()
    return out


def pagerank_small(vectors):
    out = {}
    for (n, c, q, s) in [(30, 8, 0.5, 3), (200, 8, 0.5, 7), (1000, 8, 0.5, 7), (2000, 12, 0.3, 5),
                         (1500, 16, 0.1, 1)]:
        for fmt in FMT_ORDER:
            m = R.new(fmt).fill_cqmat_stochastic(n, n, c, q, s)
            for (eps, beta, it) in [(1e-10, 0.85, 500), (0.85, 0.01, 10)]:
                pr, its = m.page_rank(eps, beta, it)
                key = f"pr/{n}c{c}q{q}s{s}/{fmt}/eps{eps}b{beta}"
                out[key] = its
                vectors[key] = pr
                nodes = np.array([1, n // 3, n - 1], dtype=np.int32)
                pr, its = m.personalized_page_rank(eps, beta, it, nodes)
                key = f"ppr/{n}c{c}q{q}s{s}/{fmt}/eps{eps}b{beta}"
                out[key] = its
                vectors[key] = pr
    return out


def big_configs():
    """Digests of the full-size workloads (reference generator, single block)."""
    out = {}
    t = time.time()

    def rec(key, m):
        info, arrs = m.arrays()
        d = {k: info[k] for k in INFO_KEYS}
        d["nnz"] = int(R._info(m.h)["nnz"])
        if "max_col" in info:
            d["max_col"] = info["max_col"]
        for k, a in arrs.items():
            d[f"sha256_{k}"] = sha(a)
            d[f"len_{k}"] = int(a.size)
        out[key] = d
        print(f"  {key}: {time.time() - t:.1f}s", flush=True)
        return arrs

    # cdiag CSR (n=10,000, C=10) and its SpMV on x_j = u01(1, j)
    rec("cdiag10000/CSR", R.new("CSR").fill_cdiag(10000, 10000, 10))
    # one rank's block of the n=2^24, c=16, q=0.1, seed 42 matrix on a 4x4 grid
    n = 1 << 24
    for fmt in ["CSR", "ELL", "SCOO"]:
        rec(f"cqmat2^24c16q0.1s42/{fmt}/block0.0.4.4", R.new(fmt).fill_cqmat(n, n, 16, 0.1, 42, 0, 0, 4, 4))
    rec("cqmat2^24c16q0.1s42/COO/rank0of16", R.new("COO").fill_cqmat(n, n, 16, 0.1, 42, 0, 0, 4, 4))
    rec("cqmat32768c16q0.1s42/DENSE/block0.0.4.4", R.new("DENSE").fill_cqmat(32768, 32768, 16, 0.1, 42, 0, 0, 4, 4))
    # PageRank graph n=2^24, c=8, q=0.5, seed 7 (full matrix)
    rec("stoch2^24c8q0.5s7/CSR", R.new("CSR").fill_cqmat_stochastic(n, n, 8, 0.5, 7))
    # adjoint graph n=2^20, c=16, q=0.2, seed 3
    rec("stoch2^20c16q0.2s3/CSR", R.new("CSR").fill_cqmat_stochastic(1 << 20, 1 << 20, 16, 0.2, 3))
    # batched: the n=2^22, c=16, q=0.3 graph (seed 1256) and a sample of small graphs
    rec("stoch2^22c16q0.3s1256/CSR", R.new("CSR").fill_cqmat_stochastic(1 << 22, 1 << 22, 16, 0.3, 1256))
    for (ng, s) in [(1024, 1000), (5000, 1001), (65536, 1002), (262144, 1003)]:
        rec(f"stoch{ng}c12q0.3s{s}/CSR", R.new("CSR").fill_cqmat_stochastic(ng, ng, 12, 0.3, s))
    return out


def pagerank_big(vectors):
    out = {}
    n = 1 << 20
    m = R.new("CSR").fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    pr, its = m.page_rank(1e-10, 0.85, 10000)
    out["pr/2^20c8q0.5s7/CSR/eps1e-10b0.85"] = {"iterations": its, "sha256": sha(pr),
                                                 "sum": float(pr.sum())}
    idx = np.random.default_rng(0).choice(n, 4096, replace=False)
    vectors["pr2^20_sample_idx"] = idx.astype(np.int64)
    vectors["pr2^20_sample_val"] = pr[idx]
    return out


BINARY_CASES = {
    # name: (generator, args)  -- reference generators only (values 1.0 or 1/count)
    "cqmat40c6q0.3s2_block1.0.2.2": ("fill_cqmat", (40, 40, 6, 0.3, 2, 1, 0, 2, 2)),
    "stoch50c5q0.5s3": ("fill_cqmat_stochastic", (50, 50, 5, 0.5, 3)),
    "cdiag30c4_block1.2.2.3": ("fill_cdiag", (30, 30, 4, 1, 2, 2, 3)),
}


def binary_files(vectors):
    """Files written by the reference's write(), and its SpMV of each (x =
    0, 1, 2, ... over the x the format reads), for the interchange tests."""
    d = os.path.join(HERE, "ref_binary")
    os.makedirs(d, exist_ok=True)
    out = {}
    for name, (gen, args) in BINARY_CASES.items():
        for fmt in FMT_ORDER:
            m = getattr(R.new(fmt), gen)(*args)
            blob = R.serialize(m)
            fn = f"{name}.{fmt}"
            with open(os.path.join(d, fn), "wb") as f:
                f.write(blob)
            info, _ = m.arrays()
            nx = info["n_col"] if fmt == "COO" else info["ln_col"]
            x = np.arange(nx, dtype=np.float64)
            vectors[f"binary/{fn}/spmv_iota"] = m.spmv(x)
            out[fn] = {"sha256": hashlib.sha256(blob).hexdigest(), "bytes": len(blob), "nx": nx}
    # two matrices appended to one file, as C_CPP_Matrix<F>_write does
    a = R.new("CSR").fill_cqmat(40, 40, 6, 0.3, 2, 1, 0, 2, 2)
    b = R.new("CSR").fill_cqmat_stochastic(50, 50, 5, 0.5, 3)
    ba, bb = R.serialize(a), R.serialize(b)
    with open(os.path.join(d, "appended2.CSR"), "wb") as f:
        f.write(ba + bb)
    out["appended2.CSR"] = {"sha256": hashlib.sha256(ba + bb).hexdigest(), "bytes": len(ba) + len(bb),
                            "second_at": len(ba)}
    return out


CFG3_ITERATIONS = (48, 49)


def pagerank_cfg3(vectors):
    """The reference FP64 PageRank of fill_cqmat_stochastic(2^24, 2^24, 8,
    0.5, 7), beta 0.85, run for 48 and 49 iterations (epsilon 0: no early
    stop).  From about 40 iterations on, the max-normalised iterate cycles
    between two states; their sum-normalised forms differ by ~1e-16 in L1,
    so either is the FP64 fixed point p*."""
    n = 1 << 24
    t = time.time()
    m = R.new("CSR").fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    print(f"  cfg3 graph {time.time() - t:.1f}s", flush=True)
    idx = np.random.default_rng(24).choice(n, 8192, replace=False).astype(np.int64)
    vectors["cfg3_sample_idx"] = idx
    out = {"graph": "fill_cqmat_stochastic(2^24,2^24,8,0.5,7)", "beta": 0.85, "epsilon": 0.0}
    states = []
    for k in CFG3_ITERATIONS:
        pr, its = m.page_rank(0.0, 0.85, k)
        assert its == k
        states.append(pr)
        out[f"it{k}"] = {"sha256": sha(pr), "sum": float(pr.sum())}
        vectors[f"cfg3_it{k}_sample_val"] = pr[idx]
        print(f"  cfg3 {k} iterations {time.time() - t:.1f}s", flush=True)
    out["l1_between_cycle_states"] = float(np.abs(states[0] - states[1]).sum())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true")
    ap.add_argument("--cfg3", action="store_true", help="only add the n=2^24 PageRank fixture")
    a = ap.parse_args()
    if not R.available():
        sys.exit("oracle/_ref/libtbsla_ref.so missing: run `make -C oracle ref` first")
    if os.environ.get("OMP_NUM_THREADS") != "1" and not a.cfg3:
        # (the cfg3 fixture is CSR only, whose row sums do not depend on the
        # thread count)
        sys.exit("run with OMP_NUM_THREADS=1 (sequential COO/SCOO reduction order)")
    path = os.path.join(HERE, "digests.json")
    dig = json.load(open(path)) if os.path.exists(path) else {}
    vec = dict(np.load(os.path.join(HERE, "golden_small.npz"))) if os.path.exists(
        os.path.join(HERE, "golden_small.npz")) else {}
    t = time.time()
    if a.cfg3:
        dig["pagerank_cfg3"] = pagerank_cfg3(vec)
        json.dump(dig, open(path, "w"), indent=1, sort_keys=True)
        np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **vec)
        print(f"done {time.time() - t:.1f}s")
        return
    dig["grids"] = small_grids()
    print(f"grids {time.time() - t:.1f}s", flush=True)
    dig["conversions"] = conversions_and_stochastic()
    dig["pagerank_small"] = pagerank_small(vec)
    dig["binary"] = binary_files(vec)
    print(f"small {time.time() - t:.1f}s", flush=True)
    # 5x5 known answer (tests/csr.cpp:8-25, tests/coo.cpp:8-25)
    vec["ka_values"] = np.arange(1, 11, dtype=np.float64)
    vec["ka_rowptr"] = np.array([0, 3, 4, 5, 7, 10], dtype=np.int32)
    vec["ka_colidx"] = np.array([0, 1, 4, 2, 1, 1, 3, 0, 2, 4], dtype=np.int32)
    vec["ka_row"] = np.array([0, 0, 0, 1, 2, 3, 3, 4, 4, 4], dtype=np.int32)
    vec["ka_x"] = np.arange(1, 6, dtype=np.float64)
    vec["ka_expected"] = np.array([20, 12, 10, 40, 85], dtype=np.float64)
    # cdiag(10000, 10000, 10) SpMV on x_j = u01(1, j) through the reference
    from oracle.oracle import Oracle as O
    x = O.uniform_vector(1, 10000)
    vec["cdiag10000_x"] = x
    vec["cdiag10000_y"] = R.new("CSR").fill_cdiag(10000, 10000, 10).spmv(x)
    if a.big:
        dig["big"] = big_configs()
        dig["pagerank_big"] = pagerank_big(vec)
    dig["_about"] = ("SHA-256 digests of arrays produced by the compiled TBSLA reference "
                     "(oracle/_ref) with OMP_NUM_THREADS=1; see tests/golden/make_golden.py")
    json.dump(dig, open(path, "w"), indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **vec)
    print(f"done {time.time() - t:.1f}s")


if __name__ == "__main__":
    main()
