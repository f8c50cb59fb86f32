"""Batched personalised PageRank (tbsla_batch_*) against the reference
semantics of Matrix::personalized_page_rank (Matrix.cpp:63-114), one oracle
call per teleport set, as the reference would run them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_batch(tb, oracle):
    """Five single-set graphs of uneven size (one with a 3-node set), one
    graph with 9 one-hot sets (SpMM group), one with 70 sets (two groups)."""
    spec = [(1024, 12, 0.3, 1000, [[5]]), (1500, 12, 0.3, 1001, [[1499]]), (3000, 8, 0.5, 7, [[0, 17, 2999]]),
            (5000, 12, 0.3, 1002, [[2500]]), (2100, 16, 0.1, 3, [[77]]),
            (4096, 16, 0.3, 1256, [[(37 * r) % 4096] for r in range(9)]),
            (2048, 12, 0.3, 1003, [[(101 * r + 5) % 2048] for r in range(70)])]
    mats, orcs = [], []
    for (n, c, q, s, sets) in spec:
        mats.append(tb.MatrixCSR().fill_cqmat_stochastic(n, n, c, q, s))
        orcs.append(oracle.new("CSR").fill_cqmat_stochastic(n, n, c, q, s))
    return spec, mats, orcs


def run_batch(tb, mats, spec, precision, **kw):
    b = tb.PageRankBatch(precision)
    for m, (n, c, q, s, sets) in zip(mats, spec):
        b.add_graph(m, sets)
    out, rep, its = b.run(**kw)
    return b.layout(), out.cpu().numpy(), rep, its


def test_batch_f64_matches_reference_per_set(tb, oracle, gpu):
    spec, mats, orcs = small_batch(tb, oracle)
    off, out, rep, its = run_batch(tb, mats, spec, "f64", epsilon=1e-10, max_iteration=500)
    k = 0
    for o, (n, c, q, s, sets) in zip(orcs, spec):
        for nodes in sets:
            pr, it = o.personalized_page_rank(1e-10, 0.85, 500, np.array(nodes, dtype=np.int32))
            got = out[off[k]:off[k + 1]]
            assert abs(int(its[k]) - it) <= 1, (n, nodes, its[k], it)
            assert np.abs(got - pr).sum() <= 1e-13, (n, nodes, np.abs(got - pr).sum())
            k += 1
    assert k == len(its) and rep["iterations"] == its.max()


def test_batch_long_rows_and_empty_rows(tb, oracle, gpu):
    """Rows of 0 to 70 entries (the SpMM reads a row 16 entries at a time;
    the single-set tiles send rows past half a tile to the per-row path):
    an SpMM group of 9 sets and a single-set graph, FP64, per set against the
    reference iteration."""
    rng = np.random.default_rng(11)
    graphs = []
    for n in (3000, 2500):
        lens = rng.integers(0, 24, n)
        lens[::37] = rng.integers(17, 71, lens[::37].size)  # rows past one 16-entry chunk
        lens[5::101] = 0
        if n == 2500:
            lens[7] = 1100  # past half a 2048-entry tile
        rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        col = rng.integers(0, n, rowptr[-1]).astype(np.int32)
        cnt = np.bincount(col, minlength=n).astype(np.float64)
        val = 1.0 / cnt[col]
        graphs.append((n, rowptr, col, val))
    sets = [[[(53 * r + 11) % 3000] for r in range(9)], [[17]]]
    b = tb.PageRankBatch("f64")
    for (n, rowptr, col, val), st in zip(graphs, sets):
        b.add_graph(tb.MatrixCSR(n, n, val, rowptr, col), st)
    out, rep, its = b.run(epsilon=1e-10, max_iteration=500, graph=True, autotune=True)
    out = out.cpu().numpy()
    off = b.layout()
    k = 0
    for (n, rowptr, col, val), st in zip(graphs, sets):
        o = oracle.from_arrays("CSR", n, n, val, rowptr, col)
        for nodes in st:
            pr, it = o.personalized_page_rank(1e-10, 0.85, 500, np.array(nodes, dtype=np.int32))
            got = out[off[k]:off[k + 1]]
            assert abs(int(its[k]) - it) <= 1, (n, nodes, its[k], it)
            assert np.abs(got - pr).sum() <= 1e-13, (n, nodes, np.abs(got - pr).sum())
            k += 1


def test_batch_graph_replay_equals_plain_launches(tb, oracle, gpu):
    """hipGraph replay, plain launches and every autotuned launch shape give
    bit-identical results (the products and reductions are order-fixed)."""
    spec, mats, _ = small_batch(tb, oracle)
    _, a, _, ia = run_batch(tb, mats, spec, "f64", epsilon=1e-10, max_iteration=300, graph=True, autotune=True)
    _, b, _, ib = run_batch(tb, mats, spec, "f64", epsilon=1e-10, max_iteration=300, graph=False, autotune=False)
    assert np.array_equal(ia, ib)
    assert np.array_equal(a, b)


def test_batch_autotuner_trials_and_cache(tb, oracle, gpu):
    """The autotuner's trials are iterations of the run (the result equals a
    run with a fixed shape bit for bit); a second batch of the same
    signature takes the cached shape with no trial iterations."""
    spec, mats, _ = small_batch(tb, oracle)
    b1 = tb.PageRankBatch("f64")
    b2 = tb.PageRankBatch("f64")
    for m, sp in zip(mats, spec):
        b1.add_graph(m, sp[4])
        b2.add_graph(m, sp[4])
    o1, r1, i1 = b1.run(epsilon=1e-10, max_iteration=300, autotune=True)
    o2, r2, i2 = b2.run(epsilon=1e-10, max_iteration=300, autotune=True)
    if not r1["tune_cached"]:  # (an earlier test may have tuned this signature already)
        assert r1["tune_iterations"] >= 6 and r1["tune_iterations"] % 2 == 0
    assert r2["tune_cached"] and r2["tune_iterations"] == 0 and r2["config"] == r1["config"]
    _, c, _, ic = run_batch(tb, mats, spec, "f64", epsilon=1e-10, max_iteration=300, graph=True, autotune=False)
    assert np.array_equal(i1, ic) and np.array_equal(o1.cpu().numpy(), c)
    assert np.array_equal(o1.cpu().numpy(), o2.cpu().numpy())


def test_batch_every_shape_same_bits(tb, gpu):
    """The autotuner's trials cover every tile size (1024, 2048, 3072
    entries) and SpMM rows per item (32, 64, 128) as iterations of the run:
    rows longer than 512 entries are long rows at every tile size, so the
    result equals a fixed-shape run's bit for bit.  Rows of up to 1600
    entries (past half of every tile size) in the single-set graph, and an
    SpMM group whose rows end inside and past each work item size."""
    rng = np.random.default_rng(2027)
    graphs = []
    for n in (4000, 3300):
        lens = rng.integers(0, 20, n)
        if n == 4000:
            lens[[3, 900, 1800, 3999]] = [513, 1025, 1537, 1600]
        rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        col = rng.integers(0, n, rowptr[-1]).astype(np.int32)
        cnt = np.bincount(col, minlength=n).astype(np.float64)
        graphs.append((n, rowptr, col, 1.0 / cnt[col]))
    sets = [[[123]], [[(29 * r + 3) % 3300] for r in range(5)]]

    def run(autotune):
        b = tb.PageRankBatch("f64")
        for (n, rowptr, col, val), st in zip(graphs, sets):
            b.add_graph(tb.MatrixCSR(n, n, val, rowptr, col), st)
        out, rep, its = b.run(epsilon=1e-12, max_iteration=300, graph=True, autotune=autotune)
        return out.cpu().numpy(), rep, its

    a, ra, ia = run(True)
    b, rb, ib = run(False)
    if not ra["tune_cached"]:
        assert ra["tune_iterations"] >= 10, ra
    assert ra["config"]["tile_nnz"] in (1024, 2048, 3072) and ra["config"]["spmm_rows"] in (32, 64, 128)
    assert rb["config"]["tile_nnz"] == 2048 and rb["config"]["spmm_rows"] == 64
    assert np.array_equal(ia, ib) and np.array_equal(a, b)


def test_batch_f32_within_tolerance(tb, oracle, gpu):
    """FP32 values, iterates and products (FP64 residual and sums), stopped by
    the sum-normalised bound at 1e-6: within 1e-5 L1 of the FP64 result."""
    spec, mats, orcs = small_batch(tb, oracle)
    off, out, rep, its = run_batch(tb, mats, spec, "f32", epsilon=1e-6, rule="sumnorm", max_iteration=500)
    off64, out64, _, _ = run_batch(tb, mats, spec, "f64", epsilon=1e-13, rule="sumnorm", max_iteration=500)
    assert np.array_equal(off, off64)
    for k in range(len(off) - 1):
        p32, p64 = out[off[k]:off[k + 1]], out64[off[k]:off[k + 1]]
        assert abs(p32.sum() - 1) < 1e-12
        assert np.abs(p32 - p64).sum() <= 1e-5, (k, np.abs(p32 - p64).sum())


def test_batch_max_iteration_and_zero(tb, oracle, gpu):
    spec, mats, orcs = small_batch(tb, oracle)
    off, out, rep, its = run_batch(tb, mats[:2], spec[:2], "f64", epsilon=0.0, max_iteration=3)
    assert list(its) == [3, 3]
    for k, o in enumerate(orcs[:2]):
        pr, it = o.personalized_page_rank(0.0, 0.85, 3, np.array(spec[k][4][0], dtype=np.int32))
        assert it == 3
        assert np.abs(out[off[k]:off[k + 1]] - pr).sum() <= 1e-14


@pytest.mark.slow
def test_full_size_personalised_batch(tb, gpu):
    """The full batch (256 graphs, n log-uniform in [2^10, 2^18], plus the
    2^22 graph with 64 one-hot sets): every set sums to 1, FP32 lands within
    1e-5 L1 of FP64 for every set, and sampled sets equal the single-graph
    FP64 engine run with the same stopping rule."""
    import torch
    spec = tb.personalised_workload()
    mats = [tb.MatrixCSR().fill_cqmat_stochastic(n, n, c, q, s) for (n, c, q, s, _) in spec]
    outs = {}
    for prec, eps in (("f32", 1e-6), ("f64", 1e-10)):
        b = tb.PageRankBatch(prec)
        for m, sp in zip(mats, spec):
            b.add_graph(m, sp[4])
        out, rep, its = b.run(epsilon=eps, rule="sumnorm", max_iteration=1000)
        off = b.layout()
        outs[prec] = out.cpu().numpy()
        del b, out
    assert off[-1] == sum(n * len(sets) for (n, _, _, _, sets) in spec)
    sums = np.add.reduceat(outs["f64"], off[:-1])
    assert np.all(np.abs(sums - 1) < 1e-12)
    for k in range(len(off) - 1):
        assert np.abs(outs["f32"][off[k]:off[k + 1]] - outs["f64"][off[k]:off[k + 1]]).sum() <= 1e-5, k
    # sampled sets against the single-graph engine (FP64, same rule)
    k_of = np.cumsum([0] + [len(sp[4]) for sp in spec])
    for g in (0, 77, 255, 256):
        m, sets = mats[g], spec[g][4]
        for j in (0, len(sets) - 1):
            o = torch.zeros(m.get_n_col(), dtype=torch.float64, device=gpu)
            m.page_rank_device(o, epsilon=1e-10, max_iteration=1000, rule="sumnorm", nodes=sets[j], gather="tiles")
            k = k_of[g] + j
            assert np.abs(o.cpu().numpy() - outs["f64"][off[k]:off[k + 1]]).sum() <= 1e-12, (g, j)
