"""GPU PageRank parity (Matrix.cpp:13-114 semantics) against the reference
fixtures and the oracle, and the precision variants against the FP64 fixed
point."""
import numpy as np
import pytest

from conftest import FMTS

pytestmark = pytest.mark.gpu


def test_page_rank_small_matches_reference(tb, digests, golden, gpu):
    """Reference vectors (compiled reference, fixtures): same iteration count,
    L1 distance <= 1e-13 (the GPU sums teleport/residual in a different
    order than the reference's sequential loops)."""
    d = digests["pagerank_small"]
    for (n, c, q, s) in [(30, 8, 0.5, 3), (200, 8, 0.5, 7), (1000, 8, 0.5, 7), (2000, 12, 0.3, 5),
                         (1500, 16, 0.1, 1)]:
        for fmt in FMTS:
            m = tb.FORMATS[fmt]().fill_cqmat_stochastic(n, n, c, q, s)
            for (eps, beta, it) in [(1e-10, 0.85, 500), (0.85, 0.01, 10)]:
                pr, its = m.page_rank(eps, beta, it)
                key = f"pr/{n}c{c}q{q}s{s}/{fmt}/eps{eps}b{beta}"
                assert abs(its - d[key]) <= 1, (key, its, d[key])
                assert np.abs(pr - golden[key]).sum() <= 1e-13, (key, np.abs(pr - golden[key]).sum())
                nodes = np.array([1, n // 3, n - 1], dtype=np.int32)
                pr, its = m.personalized_page_rank(eps, beta, it, nodes)
                key = f"ppr/{n}c{c}q{q}s{s}/{fmt}/eps{eps}b{beta}"
                assert abs(its - d[key]) <= 1, (key, its, d[key])
                assert np.abs(pr - golden[key]).sum() <= 1e-13, key


def test_page_rank_small_blocked_matches_reference(tb, digests, golden, gpu):
    """The reference-rule fixtures again with propagation-blocked products
    (the default from 2^16 nodes on), for CSR: the iteration count within 1
    and the vector within 1e-13 L1."""
    import torch
    d = digests["pagerank_small"]
    for (n, c, q, s) in [(200, 8, 0.5, 7), (1000, 8, 0.5, 7), (2000, 12, 0.3, 5), (1500, 16, 0.1, 1)]:
        m = tb.MatrixCSR().fill_cqmat_stochastic(n, n, c, q, s)
        out = torch.zeros(n, dtype=torch.float64, device=gpu)
        for (eps, beta, it) in [(1e-10, 0.85, 500), (0.85, 0.01, 10)]:
            nodes = np.array([1, n // 3, n - 1], dtype=np.int32)
            for kind, kw in (("pr", {}), ("ppr", {"nodes": nodes})):
                rep = m.page_rank_device(out, epsilon=eps, beta=beta, max_iteration=it, rule="reference",
                                         gather="blocked", **kw)
                key = f"{kind}/{n}c{c}q{q}s{s}/CSR/eps{eps}b{beta}"
                assert abs(rep["iterations"] - d[key]) <= 1, (key, rep["iterations"], d[key])
                err = np.abs(out.cpu().numpy() - golden[key]).sum()
                assert err <= 1e-13, (key, err)


@pytest.fixture(scope="module")
def graph_2_20(tb, oracle):
    n = 1 << 20
    p = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    o = oracle.new("CSR").fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    pr_ref, its = o.page_rank(1e-10, 0.85, 10000)
    # the FP64 fixed point: the oracle iteration run until its residual stops falling
    pstar, _ = o.page_rank(0.0, 0.85, 120)
    return p, pr_ref, its, pstar


def reference_errors(oracle, n, ks):
    """The reference's stopping quantity (Matrix.cpp:38-45) after k
    iterations, from the oracle, for each k in ks."""
    import ctypes
    o = oracle.new("CSR").fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    out = np.zeros(n)
    errs = {}
    for k in ks:
        e = ctypes.c_double()
        oracle.lib.orc_page_rank(o.h, 0.0, 0.85, k, out, ctypes.byref(e))
        errs[k] = e.value
    return errs


def test_page_rank_2_20_reference_rule(graph_2_20, oracle, digests, gpu):
    """n=2^20, c=8, q=0.5, seed 7: the reference stops after 44 iterations.
    From about iteration 41 on its stopping quantity is rounding noise
    (3.0e-10, 1.3e-10, 1.1e-10, 7.8e-11, 1.2e-10, ... while the iterates
    differ by ~1e-16 L1), so where it first falls below 1e-10 depends on the
    summation order.  The fixed-order tiles (the reference's row sums) stop
    within one iteration of the reference; page_rank() (propagation-blocked
    products at this size, rows summed in a run-dependent order) stops where
    the reference's own quantity is already at that floor (below 2e-10).
    Both vectors are within 1e-13 L1 of the reference's."""
    import torch
    p, pr_ref, its_ref, pstar = graph_2_20
    assert its_ref == digests["pagerank_big"]["pr/2^20c8q0.5s7/CSR/eps1e-10b0.85"]["iterations"]
    out = torch.zeros(p.get_n_col(), dtype=torch.float64, device=gpu)
    rep = p.page_rank_device(out, epsilon=1e-10, max_iteration=10000, rule="reference", gather="tiles")
    assert abs(rep["iterations"] - its_ref) <= 1
    assert np.abs(out.cpu().numpy() - pr_ref).sum() <= 1e-13
    pr, its = p.page_rank(1e-10, 0.85, 10000)
    print(f"page_rank(): {its} iterations (reference {its_ref})", flush=True)
    if abs(its - its_ref) > 1:
        errs = reference_errors(oracle, p.get_n_col(), [its])
        assert errs[its] < 2e-10, (its, errs)
    assert np.abs(pr - pr_ref).sum() <= 1e-13, np.abs(pr - pr_ref).sum()
    assert abs(pr.sum() - 1) < 1e-14


@pytest.mark.parametrize("gather", ["tiles", "blocked"])
@pytest.mark.parametrize("precision", ["f64", "pattern", "f32", "bf16", "pattern_f32"])
def test_page_rank_precisions_reach_1e10(graph_2_20, gpu, precision, gather):
    """Every precision variant, with fixed-order tile products or with
    propagation-blocked products, stops on the sum-normalised bound and lands
    within 1e-10 L1 of the FP64 fixed point (BF16/FP32 values and the
    value-free FP32 deltas need the FP64 anchor product to get there)."""
    import torch
    p, _, _, pstar = graph_2_20
    out = torch.zeros(p.get_n_col(), dtype=torch.float64, device=gpu)
    rep = p.page_rank_device(out, epsilon=1e-10, beta=0.85, max_iteration=500, precision=precision,
                             rule="sumnorm", gather=gather)
    torch.cuda.synchronize()
    err = np.abs(out.cpu().numpy() - pstar).sum()
    assert err <= 1e-10, (precision, err, rep)
    assert rep["iterations"] < 500


def test_pattern_path_is_bit_identical(graph_2_20, gpu):
    """Value-free products equal the FP64-value products bit for bit, so the
    whole run is identical."""
    import torch
    p = graph_2_20[0]
    a = torch.zeros(p.get_n_col(), dtype=torch.float64, device=gpu)
    b = torch.zeros_like(a)
    ra = p.page_rank_device(a, epsilon=1e-10, max_iteration=60, precision="f64", rule="reference", gather="tiles")
    rb = p.page_rank_device(b, epsilon=1e-10, max_iteration=60, precision="pattern", rule="reference",
                            gather="tiles")
    torch.cuda.synchronize()
    assert ra["iterations"] == rb["iterations"]
    assert torch.equal(a, b)


def test_personalized_device_matches_oracle(tb, oracle, gpu):
    import torch
    n = 50000
    p = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 12, 0.3, 1001)
    o = oracle.new("CSR").fill_cqmat_stochastic(n, n, 12, 0.3, 1001)
    nodes = np.array([17], dtype=np.int32)
    ref, its = o.personalized_page_rank(1e-12, 0.85, 1000, nodes)
    out = torch.zeros(n, dtype=torch.float64, device=gpu)
    rep = p.page_rank_device(out, epsilon=1e-12, max_iteration=1000, nodes=nodes)
    torch.cuda.synchronize()
    assert abs(rep["iterations"] - its) <= 1
    assert np.abs(out.cpu().numpy() - ref).sum() <= 1e-12
    # propagation-blocked products, reference stopping rule: same run up to rounding
    rep = p.page_rank_device(out, epsilon=1e-12, max_iteration=1000, nodes=nodes, gather="blocked")
    assert abs(rep["iterations"] - its) <= 1
    assert np.abs(out.cpu().numpy() - ref).sum() <= 1e-12


def test_blocked_products_with_teleport_vector_and_bins(tb, oracle, gpu):
    """Propagation-blocked products on a graph spanning several bins (8192
    rows each, last one partial) and column superblocks (8192 columns) with
    an explicit teleport distribution: equal within rounding to the
    fixed-order tile products."""
    import torch
    n = 5 * 8192 + 1234
    p = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 12, 0.4, 77)
    v = oracle.uniform_vector(5, n) + 0.05
    v /= v.sum()
    vd = torch.from_numpy(v).to(gpu)
    a = torch.zeros(n, dtype=torch.float64, device=gpu)
    b = torch.zeros_like(a)
    for prec in ("f64", "pattern"):
        ra = p.page_rank_device(a, epsilon=1e-13, max_iteration=400, rule="sumnorm", teleport=vd,
                                precision=prec, gather="tiles")
        rb = p.page_rank_device(b, epsilon=1e-13, max_iteration=400, rule="sumnorm", teleport=vd,
                                precision=prec, gather="blocked")
        # (blocked runs keep the iterate near sum 1 instead of max 1, which
        # makes their stopping bound tighter: they may stop earlier)
        assert rb["iterations"] <= ra["iterations"] + 1
        assert float((a - b).abs().sum()) <= 1e-12, prec
        # the reference rule (FP64 blocked runs keep b unnormalised and carry
        # 1/max to its readers): the same iterations within one, the same vector
        ra = p.page_rank_device(a, epsilon=1e-9, max_iteration=400, rule="reference", teleport=vd,
                                precision=prec, gather="tiles")
        rb = p.page_rank_device(b, epsilon=1e-9, max_iteration=400, rule="reference", teleport=vd,
                                precision=prec, gather="blocked")
        assert abs(rb["iterations"] - ra["iterations"]) <= 1, (prec, ra, rb)
        assert float((a - b).abs().sum()) <= 1e-12, prec


def test_mixed_precisions_with_teleport_vector_and_node_set(tb, oracle, gpu):
    """The mixed schemes (FP32 and BF16 values, value-free FP32) with an
    explicit teleport distribution and with a personalised node set, on both
    product structures: within 1e-10 L1 of the FP64 run of the same problem
    (the anchor products are FP64 in every teleport mode).  The graph spans
    three 16384-column superblocks of the wide scatter order, the last one
    partial, which the value-free anchor products and their moves (the
    first from a zero anchor, the others from a moved one) run over."""
    import torch
    n = 5 * 8192 + 1234
    p = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 12, 0.4, 77)
    v = oracle.uniform_vector(5, n) + 0.05
    v /= v.sum()
    vd = torch.from_numpy(v).to(gpu)
    nodes = np.array([17, 1000, 30000, n - 1], dtype=np.int32)
    ref = torch.zeros(n, dtype=torch.float64, device=gpu)
    out = torch.zeros_like(ref)
    for kw in ({"teleport": vd}, {"nodes": nodes}):
        p.page_rank_device(ref, epsilon=1e-13, max_iteration=500, rule="sumnorm", precision="f64", gather="tiles",
                           **kw)
        for prec in ("f32", "bf16", "pattern_f32"):
            for gather in ("tiles", "blocked"):
                rep = p.page_rank_device(out, epsilon=1e-12, max_iteration=500, rule="sumnorm", precision=prec,
                                         gather=gather, **kw)
                err = float((out - ref).abs().sum())
                assert err <= 1e-10, (list(kw), prec, gather, err, rep)
                assert abs(float(out.sum()) - 1.0) < 1e-12


def test_blocked_products_small_and_odd_sizes(tb, gpu):
    """Propagation-blocked products where one bin holds everything (no far
    entries), where a bin has fewer than four local entries, and on a
    hand-made graph with empty rows and columns and a long row: equal within
    rounding to the fixed-order tile products, in every value precision."""
    import torch
    rng = np.random.default_rng(8)
    mats = [tb.MatrixCSR().fill_cqmat_stochastic(3000, 3000, 5, 0.3, 4),
            tb.MatrixCSR().fill_cqmat_stochastic(8195, 8195, 3, 0.9, 6)]
    n = 9000
    lens = rng.integers(0, 9, n)
    lens[::50] = 0
    lens[4321] = 3000
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = rng.integers(0, n, rowptr[-1]).astype(np.int32)
    col[col % 97 == 0] = 1  # some columns never appear
    cnt = np.bincount(col, minlength=n).astype(np.float64)
    mats.append(tb.MatrixCSR(n, n, 1.0 / cnt[col], rowptr, col))
    for p in mats:
        m = p.get_n_col()
        a = torch.zeros(m, dtype=torch.float64, device=gpu)
        b = torch.zeros_like(a)
        precs = ("f64", "f32", "bf16") + (("pattern_f32",) if p.count_stochastic() else ())
        for prec in precs:
            ra = p.page_rank_device(a, epsilon=1e-12, max_iteration=400, rule="sumnorm", precision=prec,
                                    gather="tiles")
            rb = p.page_rank_device(b, epsilon=1e-12, max_iteration=400, rule="sumnorm", precision=prec,
                                    gather="blocked")
            # the sum-normalised bound of the blocked runs is the tighter one
            # (iterates scaled to sum ~1, not max 1): they may stop earlier
            assert rb["iterations"] <= ra["iterations"] + 2, (m, prec, ra, rb)
            assert float((a - b).abs().sum()) <= 1e-11, (m, prec, float((a - b).abs().sum()))


def test_blocked_layout_gpu_build_matches_host_build(tb, graph_2_20, gpu):
    """The propagation-blocked layout that page_rank builds on the GPU
    (device radix sorts and scans) equals, array by array, the layout built
    by a sequential pass on the host cores: on generated graphs, on 2^20
    nodes, and on hand-made graphs with empty rows, a long row, a column
    that every bin points to (runs across many bins), local rows more than
    seven rows apart (groups ended early), a single node and no entries."""
    from tbsla_amd._lib import lib
    rng = np.random.default_rng(11)
    mats = [tb.MatrixCSR().fill_cqmat_stochastic(3000, 3000, 5, 0.3, 4),
            tb.MatrixCSR().fill_cqmat_stochastic(8195, 8195, 3, 0.9, 6),
            tb.MatrixCSR().fill_cqmat_stochastic(40000, 40000, 16, 0.1, 2),
            graph_2_20[0]]
    n = 50000
    lens = rng.integers(0, 9, n)
    lens[::50] = 0
    lens[4321] = 30000
    lens[20000:28192:9] = 1  # one local entry every 9 rows
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = rng.integers(0, n, rowptr[-1]).astype(np.int32)
    col[col % 97 == 0] = 5  # a heavy column, seen from every bin
    for r in range(20000, 28192, 9):
        col[rowptr[r]:rowptr[r + 1]] = r
    mats.append(tb.MatrixCSR(n, n, rng.random(rowptr[-1]), rowptr, col))
    mats.append(tb.MatrixCSR(1, 1, np.array([1.0]), np.array([0, 1]), np.array([0])))
    mats.append(tb.MatrixCSR(70000, 70000, np.ones(3), np.array([0] * 8 + [3] * 69993), np.array([69999, 3, 40000])))
    mats.append(tb.MatrixCSR(100, 100, np.zeros(0), np.zeros(101), np.zeros(0)))
    for m in mats:
        assert lib.tbsla_hip_pb_layout_check(m._h) == 0, (m.get_n_col(), lib.tbsla_last_error())


def test_value_free_precisions_need_count_stochastic_values(tb, gpu):
    """PATTERN and PATTERN_F32 read no values, so they are refused for a
    matrix whose values are not 1/count of their column; PATTERN_F32 also
    needs the sum-normalised rule, like the other mixed schemes."""
    import torch
    from tbsla_amd._lib import TbslaError
    m = tb.MatrixCSR().fill_cqmat(1000, 1000, 6, 0.3, 3)
    out = torch.zeros(1000, dtype=torch.float64, device=gpu)
    for prec in ("pattern", "pattern_f32"):
        with pytest.raises(TbslaError):
            m.page_rank_device(out, epsilon=1e-10, max_iteration=50, precision=prec, rule="sumnorm")
    s = tb.MatrixCSR().fill_cqmat_stochastic(1000, 1000, 6, 0.3, 3)
    with pytest.raises(TbslaError):
        s.page_rank_device(out, epsilon=1e-10, max_iteration=50, precision="pattern_f32", rule="reference")


def test_launch_batches_do_not_change_the_run(tb, gpu):
    """How many iterations the host launches between two reads of the state
    (check_every, and the batch sized from the residual decay) changes only
    the launches that return at once: the iteration count and the vector are
    the same (fixed-order tiles, reference rule, FP64)."""
    import torch
    n = 20000
    p = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 8, 0.5, 7)
    ref = torch.zeros(n, dtype=torch.float64, device=gpu)
    r0 = p.page_rank_device(ref, epsilon=1e-12, max_iteration=500, rule="reference", gather="tiles", check_every=1)
    out = torch.zeros_like(ref)
    for ce in (2, 3, 8, 64):
        r = p.page_rank_device(out, epsilon=1e-12, max_iteration=500, rule="reference", gather="tiles",
                               check_every=ce)
        assert r["iterations"] == r0["iterations"], (ce, r, r0)
        assert torch.equal(out, ref), ce


def test_page_rank_zero_iterations(tb, gpu):
    m = tb.MatrixCSR().fill_cqmat_stochastic(100, 100, 4, 0.5, 2)
    pr, its = m.page_rank(1e-10, 0.85, 0)
    assert its == 0 and np.allclose(pr, 1.0 / 100)


def oracle_pstar_2_24(g, oracle, digests, golden):
    """The FP64 fixed point p* of fill_cqmat_stochastic(2^24, 2^24, 8, 0.5, 7),
    beta 0.85, on the host: the oracle's 48-iteration reference iteration over
    the product's arrays, whose digests are the reference's (checked here),
    and p* itself is checked against the reference's own vector digest."""
    import hashlib
    big = digests["big"]["stoch2^24c8q0.5s7/CSR"]
    a = g.arrays()
    for k in ("values", "rowptr", "colidx"):
        assert hashlib.sha256(a[k].tobytes()).hexdigest() == big[f"sha256_{k}"], k
    n = g.get_n_col()
    o = oracle.from_arrays("CSR", n, n, a["values"], a["rowptr"], a["colidx"])
    del a
    pstar, its = o.page_rank(0.0, 0.85, 48)
    assert its == 48
    assert hashlib.sha256(pstar.tobytes()).hexdigest() == digests["pagerank_cfg3"]["it48"]["sha256"]
    assert np.array_equal(pstar[golden["cfg3_sample_idx"]], golden["cfg3_it48_sample_val"])
    return pstar


@pytest.mark.slow
def test_full_size_pagerank_2_24(tb, oracle, digests, golden, gpu):
    """fill_cqmat_stochastic(2^24, 2^24, 8, 0.5, 7) at full size against the
    FP64 fixed point p* computed on the host by the oracle (and equal to the
    reference's own limit vector): the GPU's FP64 limit (48 reference-rule
    iterations, tile products) is within 1e-12 L1 of p*, and every precision
    with propagation-blocked products, and FP64 with tiles, stopped by the
    sum-normalised bound at 1e-10, lands within 1e-10 L1 of p* and sums to 1."""
    import torch
    n = 1 << 24
    print("generating the 2^24 graph on the GPU", flush=True)
    g = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 8, 0.5, 7, device=True)
    print("oracle p* on the host (48 iterations)", flush=True)
    pstar = torch.from_numpy(oracle_pstar_2_24(g, oracle, digests, golden)).to(gpu)
    out = torch.zeros_like(pstar)
    g.page_rank_device(out, epsilon=0.0, max_iteration=48, precision="f64", rule="reference", gather="tiles")
    lim = float((out - pstar).abs().sum())
    # the row sums are the reference's bit for bit; the teleport sum is not:
    # the reference adds its 2^24 terms left to right (typical relative
    # error ~sqrt(n) * 1.1e-16 ~ 5e-13), the GPU with a compensated sum
    assert lim <= 1e-12, lim
    for prec, gather in [("f64", "tiles"), ("f64", "blocked"), ("pattern", "blocked"), ("f32", "blocked"),
                         ("bf16", "blocked"), ("pattern_f32", "blocked")]:
        print(f"2^24 {prec} {gather}", flush=True)
        rep = g.page_rank_device(out, epsilon=1e-10, max_iteration=300, precision=prec, rule="sumnorm",
                                 gather=gather)
        err = float((out - pstar).abs().sum())
        assert err <= 1e-10, (prec, gather, err, rep)
        assert abs(float(out.sum()) - 1.0) < 1e-12
