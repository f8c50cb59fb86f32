"""Adjoint PageRank: the FP64 CPU gradient (oracle/adjoint.py) pinned against
a dense closed form and finite differences on the CPU, then the HIP adjoint
(tbsla_hip_page_rank_adjoint) and its torch.autograd wrapper against that
gradient and against central finite differences of the HIP forward.

Workload of the full-size case: fill_cqmat_stochastic(2^20, 2^20, 16, 0.2, 3),
beta 0.85, uniform teleport, L = sum_i w_i p_i with w_i ~ N(0,1) from a
Box-Muller transform of u01(3, .), 64 edges drawn with u01(5, .)."""
import numpy as np
import pytest

from oracle import adjoint as adj


def tiny_graph(n=24, seed=0):
    rng = np.random.default_rng(seed)
    Ad = (rng.random((n, n)) < 0.25) * rng.random((n, n))
    Ad += np.diag(rng.random(n) * 0.1)  # no empty column
    return Ad / Ad.sum(axis=0, keepdims=True)


def test_oracle_adjoint_matches_dense_closed_form_and_fd():
    """Iterative sparse gradient == exact dense solve == central differences
    of the Perron-vector forward, for the edge values and the teleport."""
    import scipy.sparse as sp
    n, beta = 24, 0.85
    Ad = tiny_graph(n)
    rng = np.random.default_rng(1)
    v = rng.random(n) + 0.2
    v /= v.sum()
    w = rng.standard_normal(n)
    GA, gt_dense, p_dense = adj.dense_adjoint(Ad, w, beta, v)
    A = sp.csr_matrix(Ad)
    p = adj.fixed_point(A, beta, v)
    assert np.abs(p - p_dense).sum() < 1e-13
    gv, gt, _ = adj.adjoint(A, p, w, beta, v)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    scale = np.abs(GA).max()
    assert np.abs(gv - GA[rows, A.indices]).max() <= 1e-10 * scale
    assert np.abs(gt - gt_dense).max() <= 1e-10 * np.abs(gt_dense).max()
    # central differences of L = w . p(A, v)
    L = lambda Am, vv: w @ adj.dense_fixed_point(Am, beta, vv)
    h = 1e-6
    for k in range(0, A.nnz, 7):
        i, j = rows[k], A.indices[k]
        Ap, Am = Ad.copy(), Ad.copy()
        Ap[i, j] += h
        Am[i, j] -= h
        fd = (L(Ap, v) - L(Am, v)) / (2 * h)
        assert abs(fd - gv[k]) <= 1e-6 * scale, (k, fd, gv[k])
    for k in range(0, n, 5):
        vp, vm = v.copy(), v.copy()
        vp[k] += h
        vm[k] -= h
        fd = (L(Ad, vp) - L(Ad, vm)) / (2 * h)
        assert abs(fd - gt[k]) <= 1e-6 * np.abs(gt).max(), (k, fd, gt[k])


def test_oracle_adjoint_mistakes_fail():
    """Plausible mistakes -- g without the (I - 1 p^T) projection, or the
    adjoint solved with J instead of J^T -- disagree with the closed form
    by far more than the tolerances above: the check can see them."""
    n, beta = 24, 0.85
    Ad = tiny_graph(n, 3)
    w = np.random.default_rng(4).standard_normal(n)
    GA, _, p = adj.dense_adjoint(Ad, w, beta)
    M = beta * Ad + (1 - beta) / n
    s = (M @ p).sum()
    J = (np.eye(n) - np.outer(p, np.ones(n))) @ M / s
    lam = np.linalg.solve(np.eye(n) - J.T, w)
    g_noproj = lam / s
    assert np.abs(beta * np.outer(g_noproj, p) - GA).max() > 1e-3 * np.abs(GA).max()
    lam_t = np.linalg.solve(np.eye(n) - J, w)
    g_t = (lam_t - p @ lam_t) / s
    assert np.abs(beta * np.outer(g_t, p) - GA).max() > 1e-3 * np.abs(GA).max()


def box_muller_w(oracle, seed, n):
    u = oracle.uniform_vector(seed, 2 * n)
    u1, u2 = 1.0 - u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def csr_scipy(m):
    a = m.arrays()
    rp = a["rowptr"] - a["rowptr"][0]
    ne = rp[-1]
    return adj.csr_matrix(m.get_n_col(), rp, a["colidx"][:ne], a["values"][:ne])


@pytest.mark.gpu
@pytest.mark.parametrize("teleport", [False, True])
def test_adjoint_small_against_cpu_gradient(tb, oracle, gpu, teleport):
    import torch
    n, beta = 3000, 0.85
    m = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 12, 0.3, 1001)
    A = csr_scipy(m)
    v = None
    vd = None
    if teleport:
        v = oracle.uniform_vector(9, n) + 0.1
        v /= v.sum()
        vd = torch.from_numpy(v).to(gpu)
    p = torch.empty(n, dtype=torch.float64, device=gpu)
    m.page_rank_device(p, epsilon=1e-15, max_iteration=2000, rule="sumnorm", teleport=vd)
    p_cpu = adj.fixed_point(A, beta, v)
    assert np.abs(p.cpu().numpy() - p_cpu).sum() <= 1e-12
    w = box_muller_w(oracle, 3, n)
    gv_ref, gt_ref, _ = adj.adjoint(A, p.cpu().numpy(), w, beta, v)
    wd = torch.from_numpy(w).to(gpu)
    for prec, tol in (("f64", 1e-9), ("f32", 1e-4)):
        gv = torch.zeros(m.n_entries(), dtype=torch.float64, device=gpu)
        gt = torch.zeros(n, dtype=torch.float64, device=gpu)
        rep = m.page_rank_adjoint(p, wd, gv, gt, beta=beta, precision=prec,
                                  epsilon=1e-14 if prec == "f64" else 1e-7, teleport=vd)
        assert rep["iterations"] > 10
        err_v = np.abs(gv.cpu().numpy() - gv_ref).max() / np.abs(gv_ref).max()
        err_t = np.abs(gt.cpu().numpy() - gt_ref).max() / np.abs(gt_ref).max()
        assert err_v <= tol and err_t <= tol, (prec, err_v, err_t)


@pytest.fixture(scope="module")
def graph_cfg5(tb, oracle, gpu):
    import torch
    n = 1 << 20
    m = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 16, 0.2, 3)
    p = torch.empty(n, dtype=torch.float64, device=gpu)
    m.page_rank_device(p, epsilon=1e-15, max_iteration=400, rule="sumnorm")
    w = torch.from_numpy(box_muller_w(oracle, 3, n)).to(gpu)
    return m, p, w


@pytest.mark.gpu
def test_adjoint_2_20_against_fp64_cpu_gradient(graph_cfg5, gpu):
    """FP32 transposed products within 1e-4 (FP64: 1e-8) of the FP64 CPU
    gradient, relative to the largest gradient entry."""
    import torch
    m, p, w = graph_cfg5
    A = csr_scipy(m)
    gv_ref, gt_ref, _ = adj.adjoint(A, p.cpu().numpy(), w.cpu().numpy(), 0.85, tol=1e-14)
    for prec, tol, eps in (("f32", 1e-4, 1e-7), ("f64", 1e-8, 1e-13)):
        gv = torch.zeros(m.n_entries(), dtype=torch.float64, device=gpu)
        gt = torch.zeros(m.get_n_col(), dtype=torch.float64, device=gpu)
        m.page_rank_adjoint(p, w, gv, gt, precision=prec, epsilon=eps)
        err_v = np.abs(gv.cpu().numpy() - gv_ref).max() / np.abs(gv_ref).max()
        err_t = np.abs(gt.cpu().numpy() - gt_ref).max() / np.abs(gt_ref).max()
        assert err_v <= tol and err_t <= tol, (prec, err_v, err_t)


@pytest.mark.gpu
def test_adjoint_2_20_central_differences(graph_cfg5, oracle, gpu):
    """Central differences of the FP64 HIP forward on 64 seeded edges
    (h = 1e-3 |a|) against the adjoint gradient."""
    import torch
    m, p, w = graph_cfg5
    ne = m.n_entries()
    vals = torch.from_numpy(m.arrays()["values"][:ne].copy()).to(gpu)
    gv = torch.zeros(ne, dtype=torch.float64, device=gpu)
    m.page_rank_adjoint(p, w, gv, None, precision="f64", epsilon=1e-13)
    g = gv.cpu().numpy()
    edges = (oracle.uniform_vector(5, 64) * ne).astype(np.int64)
    q = torch.empty_like(p)

    def L(values):
        m.page_rank_device(q, epsilon=1e-15, max_iteration=400, rule="sumnorm", values=values)
        return float((w * q).sum().item())

    scale = np.abs(g).max()
    for e in edges:
        a = float(vals[e].item())
        h = 1e-3 * abs(a)
        vp, vm = vals.clone(), vals.clone()
        vp[e] += h
        vm[e] -= h
        fd = (L(vp) - L(vm)) / (2 * h)
        assert abs(fd - g[e]) <= 1e-4 * abs(g[e]) + 1e-6 * scale, (int(e), fd, g[e])


@pytest.mark.gpu
def test_autograd_function(tb, oracle, gpu):
    """torch.autograd through page_rank: values.grad and teleport.grad equal
    the FP64 CPU gradient."""
    import torch
    from tbsla_amd.autograd import page_rank
    n = 2000
    m = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 12, 0.3, 1002)
    ne = m.n_entries()
    vals = torch.from_numpy(m.arrays()["values"][:ne].copy()).to(gpu).requires_grad_(True)
    v = oracle.uniform_vector(11, n) + 0.1
    v /= v.sum()
    tele = torch.from_numpy(v).to(gpu).requires_grad_(True)
    w = torch.from_numpy(box_muller_w(oracle, 3, n)).to(gpu)
    p = page_rank(m, vals, tele, adjoint_precision="f64", adjoint_epsilon=1e-14)
    (w * p).sum().backward()
    A = csr_scipy(m)
    gv_ref, gt_ref, _ = adj.adjoint(A, p.detach().cpu().numpy(), w.cpu().numpy(), 0.85, v)
    assert np.abs(vals.grad.cpu().numpy() - gv_ref).max() <= 1e-9 * np.abs(gv_ref).max()
    assert np.abs(tele.grad.cpu().numpy() - gt_ref).max() <= 1e-9 * np.abs(gt_ref).max()
