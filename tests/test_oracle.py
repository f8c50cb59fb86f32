"""Pins the CPU restatement (oracle/tbsla_oracle.c) before it is trusted.

Each function of the restatement is checked against (a) fixtures produced by
the compiled reference (tests/golden, made by tests/golden/make_golden.py),
(b) the compiled reference itself where it is built (oracle/_ref), and
(c) what the mathematics fixes: closed forms, brute force, invariants,
published generator values.
"""
import hashlib
import json

import numpy as np
import pytest

from conftest import BLOCKS, CDIAG_SHAPES, CQMAT_SHAPES, FMTS, INFO_KEYS, grid_q


def fold(h, info, arrs):
    h.update(json.dumps([info[k] for k in INFO_KEYS] + [info.get("max_col", -1)]).encode())
    for k in sorted(arrs):
        h.update(k.encode())
        h.update(np.ascontiguousarray(arrs[k]).tobytes())


# ---------------------------------------------------------------- partition

def test_lnv_pflv_brute_force(oracle):
    """range.cpp:7-26: blocks tile [0, size) exactly, sizes differ by <= 1."""
    for size in [0, 1, 7, 100, 2 ** 24 + 3]:
        for P in [1, 2, 3, 4, 7, 16]:
            sizes = [oracle.lnv(size, p, P) for p in range(P)]
            firsts = [oracle.pflv(size, p, P) for p in range(P)]
            assert sum(sizes) == size
            assert max(sizes) - min(sizes) <= 1
            assert firsts[0] == 0
            for p in range(1, P):
                assert firsts[p] == firsts[p - 1] + sizes[p - 1]


# --------------------------------------------------------------- generators

def test_splitmix64_published_values(oracle):
    """SplitMix64 from seed 1234567: the published test sequence (e.g. the
    Rosetta Code SplitMix64 task), independent of this code."""
    z = 1234567
    outs = []
    for _ in range(5):
        outs.append(oracle.lib.orc_splitmix64(z))
        z = (z + 0x9E3779B97F4A7C15) % 2 ** 64
    assert outs == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                    4593380528125082431, 16408922859458223821]
    u = oracle.uniform_vector(1, 100000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.005


@pytest.mark.parametrize("shape", CQMAT_SHAPES[:24])
def test_cqmat_grid_against_golden(oracle, digests, shape):
    """fill_cqmat in every format over the reference's test grid
    (tests/test_spmv_cqmat.cpp:153-193), 3 block layouts, digests from the
    compiled reference."""
    nr, nc, c = shape
    for fmt in FMTS:
        for blk in BLOCKS:
            h = hashlib.sha256()
            for s in range(4):
                for q in grid_q():
                    m = oracle.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk)
                    info, arrs = m.arrays()
                    fold(h, info, arrs)
                    x = np.arange(info["ln_col"] if fmt != "COO" else nc, dtype=np.float64)
                    h.update(m.spmv(x).tobytes())
            key = f"{nr}x{nc}c{c}/{fmt}/{'.'.join(map(str, blk))}"
            assert h.hexdigest() == digests["grids"]["cqmat"][key], key


def test_cqmat_grid_against_golden_rest(oracle, digests):
    for (nr, nc, c) in CQMAT_SHAPES[24:]:
        for fmt in FMTS:
            for blk in BLOCKS:
                h = hashlib.sha256()
                for s in range(4):
                    for q in grid_q():
                        m = oracle.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk)
                        info, arrs = m.arrays()
                        fold(h, info, arrs)
                        x = np.arange(info["ln_col"] if fmt != "COO" else nc, dtype=np.float64)
                        h.update(m.spmv(x).tobytes())
                key = f"{nr}x{nc}c{c}/{fmt}/{'.'.join(map(str, blk))}"
                assert h.hexdigest() == digests["grids"]["cqmat"][key], key


def test_cdiag_grid_against_golden(oracle, digests):
    for (nr, nc, c) in CDIAG_SHAPES:
        for fmt in FMTS:
            for blk in BLOCKS:
                h = hashlib.sha256()
                m = oracle.new(fmt).fill_cdiag(nr, nc, c, *blk)
                info, arrs = m.arrays()
                fold(h, info, arrs)
                x = np.arange(info["ln_col"] if fmt != "COO" else nc, dtype=np.float64)
                h.update(m.spmv(x).tobytes())
                key = f"{nr}x{nc}c{c}/{fmt}/{'.'.join(map(str, blk))}"
                assert h.hexdigest() == digests["grids"]["cdiag"][key], key


def test_cdiag_closed_form(oracle):
    """tests/test_spmv_cdiag.cpp:18-105 with utils/vector.cpp:16-29, exact."""
    for (nr, nc, c) in CDIAG_SHAPES:
        x = np.arange(nc, dtype=np.float64)
        expect = oracle.spmv_res_cdiag(nr, nc, c, x)
        for fmt in FMTS:
            m = oracle.new(fmt).fill_cdiag(nr, nc, c)
            assert np.array_equal(m.spmv(x), expect), (fmt, nr, nc, c)
            coo = oracle.new("COO").fill_cdiag(nr, nc, c)
            if fmt != "COO":
                assert np.array_equal(oracle.convert(coo, fmt).spmv(x), expect), (fmt, nr, nc, c)


def test_cdiag_closed_form_independent():
    """The closed form restated in numpy against a dense +-c band (mathematics,
    not the reference code)."""
    for (nr, nc, c) in CDIAG_SHAPES:
        A = np.zeros((nr, nc))
        for i in range(nr):
            for j in ({i - c, i + c} if c else {i}):
                if 0 <= j < nc:
                    A[i, j] = 1
        x = np.arange(nc, dtype=np.float64)
        from oracle.oracle import Oracle
        assert np.array_equal(A @ x, Oracle.spmv_res_cdiag(nr, nc, c, x)), (nr, nc, c)


def test_known_answer_5x5(oracle, golden):
    """tests/csr.cpp:8-31 and tests/coo.cpp:8-31: r == {20,12,10,40,85}."""
    csr = oracle.from_arrays("CSR", 5, 5, golden["ka_values"], golden["ka_rowptr"], golden["ka_colidx"])
    coo = oracle.from_arrays("COO", 5, 5, golden["ka_values"], golden["ka_row"], golden["ka_colidx"])
    assert np.array_equal(csr.spmv(golden["ka_x"]), golden["ka_expected"])
    assert np.array_equal(coo.spmv(golden["ka_x"]), golden["ka_expected"])
    dense = np.zeros((5, 5))
    for v, r, c in zip(golden["ka_values"], golden["ka_row"], golden["ka_colidx"]):
        dense[r, c] += v
    assert np.array_equal(dense @ golden["ka_x"], golden["ka_expected"])


def test_conversions_stochastic_normalize_golden(oracle, digests):
    d = digests["conversions"]
    for (nr, nc, c) in [(10, 10, 3), (30, 30, 8), (20, 30, 4), (30, 20, 6), (50, 50, 5)]:
        for s in range(4):
            for q in [0.0, 0.1, 0.5, 1.0]:
                coo = oracle.new("COO").fill_cqmat(nr, nc, c, q, s)
                for fmt in ["SCOO", "CSR", "ELL", "DENSE"]:
                    h = hashlib.sha256()
                    fold(h, *oracle.convert(coo, fmt).arrays())
                    assert h.hexdigest() == d[f"convert/{nr}x{nc}c{c}s{s}q{q}/{fmt}"]
                for fmt in FMTS:
                    for blk in [(0, 0, 1, 1), (1, 1, 2, 2)]:
                        h = hashlib.sha256()
                        fold(h, *oracle.new(fmt).fill_cqmat_stochastic(nr, nc, c, q, s, *blk).arrays())
                        assert h.hexdigest() == d[f"stochastic/{nr}x{nc}c{c}s{s}q{q}/{fmt}/{blk}"]
                        key = f"normalize/{nr}x{nc}c{c}s{s}q{q}/{fmt}/{blk}"
                        if key in d:
                            m = oracle.new(fmt).fill_cqmat(nr, nc, c, q, s, *blk).normalize_columns()
                            h = hashlib.sha256()
                            fold(h, *m.arrays())
                            assert h.hexdigest() == d[key], key


def test_conversion_invariants(oracle):
    """COO -> CSR: rows sorted by (row, col), the (row, col, value) multiset
    kept, rowptr the running row count."""
    coo = oracle.new("COO").fill_cqmat(200, 200, 9, 0.4, 3, value_seed=5)
    _, a = coo.arrays()
    _, c = oracle.convert(coo, "CSR").arrays()
    rows = np.repeat(np.arange(200), np.diff(c["rowptr"]))
    assert np.array_equal(np.bincount(a["row"], minlength=200), np.diff(c["rowptr"]))
    key = rows.astype(np.int64) * 1000 + c["colidx"]
    assert np.all(np.diff(key) >= 0)
    m1 = sorted(zip(a["row"], a["col"], a["values"]))
    m2 = sorted(zip(rows, c["colidx"], c["values"]))
    assert m1 == m2


# ------------------------------------------------------------------ PageRank

def test_pagerank_small_golden(oracle, digests, golden):
    """page_rank / personalized_page_rank (Matrix.cpp:13-114) against the
    compiled reference's vectors and iteration counts (bit-exact)."""
    d = digests["pagerank_small"]
    for (n, c, q, s) in [(30, 8, 0.5, 3), (200, 8, 0.5, 7), (1000, 8, 0.5, 7), (2000, 12, 0.3, 5),
                         (1500, 16, 0.1, 1)]:
        for fmt in FMTS:
            m = oracle.new(fmt).fill_cqmat_stochastic(n, n, c, q, s)
            for (eps, beta, it) in [(1e-10, 0.85, 500), (0.85, 0.01, 10)]:
                pr, its = m.page_rank(eps, beta, it)
                key = f"pr/{n}c{c}q{q}s{s}/{fmt}/eps{eps}b{beta}"
                assert its == d[key] and np.array_equal(pr, golden[key]), key
                nodes = np.array([1, n // 3, n - 1], dtype=np.int32)
                pr, its = m.personalized_page_rank(eps, beta, it, nodes)
                key = f"ppr/{n}c{c}q{q}s{s}/{fmt}/eps{eps}b{beta}"
                assert its == d[key] and np.array_equal(pr, golden[key]), key


def test_pagerank_perron_vector(oracle):
    """At tiny n the fixed point is the Perron vector of
    beta*A + (1-beta)/n * 11^T (numpy eig), sum-normalised."""
    for (n, c, q, s) in [(40, 5, 0.5, 3), (64, 8, 0.3, 1)]:
        m = oracle.new("CSR").fill_cqmat_stochastic(n, n, c, q, s)
        _, a = m.arrays()
        A = np.zeros((n, n))
        rows = np.repeat(np.arange(n), np.diff(a["rowptr"]))
        np.add.at(A, (rows, a["colidx"]), a["values"])
        M = 0.85 * A + 0.15 / n * np.ones((n, n))
        w, V = np.linalg.eig(M)
        v = np.abs(np.real(V[:, np.argmax(np.real(w))]))
        v /= v.sum()
        pr, its = m.page_rank(1e-13, 0.85, 2000)
        assert np.abs(pr - v).sum() < 1e-11, (n, np.abs(pr - v).sum())
        # personalised: teleport only to the listed nodes
        nodes = np.array([0, 5, 7], dtype=np.int32)
        e = np.zeros(n)
        e[nodes] = 1.0 / 3
        Mp = 0.85 * A + 0.15 * np.outer(e, np.ones(n))
        w, V = np.linalg.eig(Mp)
        v = np.abs(np.real(V[:, np.argmax(np.real(w))]))
        v /= v.sum()
        pr, _ = m.personalized_page_rank(1e-13, 0.85, 2000, nodes)
        assert np.abs(pr - v).sum() < 1e-11


def test_pagerank_2_20_golden(oracle, digests, golden):
    """n=2^20, c=8, q=0.5, seed 7: the reference stops after 44 iterations."""
    g = digests["pagerank_big"]["pr/2^20c8q0.5s7/CSR/eps1e-10b0.85"]
    m = oracle.new("CSR").fill_cqmat_stochastic(1 << 20, 1 << 20, 8, 0.5, 7)
    pr, its = m.page_rank(1e-10, 0.85, 10000)
    assert its == g["iterations"] == 44
    assert hashlib.sha256(pr.tobytes()).hexdigest() == g["sha256"]
    assert np.array_equal(pr[golden["pr2^20_sample_idx"]], golden["pr2^20_sample_val"])


# ------------------------------------------------------- full-size digests

@pytest.mark.slow
def test_full_size_digests(oracle, digests):
    """Index/value arrays of the full-size workloads equal the reference's
    (one rank's 4x4 block of n=2^24, the PageRank graphs)."""
    cases = {
        "cdiag10000/CSR": lambda: oracle.new("CSR").fill_cdiag(10000, 10000, 10),
        "cqmat2^24c16q0.1s42/CSR/block0.0.4.4": lambda: oracle.new("CSR").fill_cqmat(2 ** 24, 2 ** 24, 16, 0.1, 42, 0, 0, 4, 4),
        "cqmat32768c16q0.1s42/DENSE/block0.0.4.4": lambda: oracle.new("DENSE").fill_cqmat(32768, 32768, 16, 0.1, 42, 0, 0, 4, 4),
        "stoch2^20c16q0.2s3/CSR": lambda: oracle.new("CSR").fill_cqmat_stochastic(2 ** 20, 2 ** 20, 16, 0.2, 3),
        "stoch65536c12q0.3s1002/CSR": lambda: oracle.new("CSR").fill_cqmat_stochastic(65536, 65536, 12, 0.3, 1002),
    }
    for key, make in cases.items():
        m = make()
        info, arrs = m.arrays()
        ref = digests["big"][key]
        for k, a in arrs.items():
            assert hashlib.sha256(a.tobytes()).hexdigest() == ref[f"sha256_{k}"], (key, k)
        if not key.startswith("cdiag"):  # MatrixCSR::fill_cdiag leaves nnz unset
            assert info["nnz"] == ref["nnz"], key


def test_overlay_spmv_against_reference(oracle, reference):
    """Non-integer values: the restatement's SpMV equals the reference's
    (same summation order) on matrices built through the array constructor."""
    m = oracle.new("CSR").fill_cqmat(3000, 3000, 16, 0.1, 42, 0, 0, 2, 2, value_seed=42)
    _, a = m.arrays()
    r = reference.from_arrays("CSR", 1500, 1500, a["values"], a["rowptr"], a["colidx"])
    x = oracle.uniform_vector(1, 1500)
    assert np.array_equal(m.spmv(x)[:1500], r.spmv(x))


def test_cdiag10000_golden(oracle, golden):
    y = oracle.new("CSR").fill_cdiag(10000, 10000, 10).spmv(golden["cdiag10000_x"])
    assert np.array_equal(y, golden["cdiag10000_y"])


@pytest.mark.slow
def test_pagerank_2_24_limit_cycle_golden(oracle, digests, golden):
    """fill_cqmat_stochastic(2^24, 2^24, 8, 0.5, 7), beta 0.85, no early stop:
    after 48 and 49 iterations the restatement's vectors are the reference's
    two limit-cycle states bit for bit (digest and sampled entries)."""
    d = digests["pagerank_cfg3"]
    n = 1 << 24
    m = oracle.new("CSR").fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    idx = golden["cfg3_sample_idx"]
    for k in (48, 49):
        pr, its = m.page_rank(0.0, 0.85, k)
        assert its == k
        assert hashlib.sha256(pr.tobytes()).hexdigest() == d[f"it{k}"]["sha256"], k
        assert np.array_equal(pr[idx], golden[f"cfg3_it{k}_sample_val"]), k
