"""FP64 CPU gradient of PageRank -- TEST INFRASTRUCTURE ONLY (the oracle of
the adjoint path; the product never imports it).

The reference has no adjoint.  The forward quantity is the fixed point of the
reference iteration (src/cpp/matrix/Matrix.cpp:13-61), generalised to a
teleport distribution v:

    p = M p / (1^T M p),  M = beta*A + (1-beta) v 1^T,  1^T p = 1.

For L(p) with w = dL/dp, s = 1^T M p and the adjoint fixed point
lambda = M^T (lambda - (p^T lambda) 1) / s + w:

    g = (lambda - (p^T lambda) 1) / s
    dL/dA_ij = beta * g_i * p_j,    dL/dv_k = (1-beta) * (1^T p) * g_k.

Pinning (tests/test_oracle.py): at tiny n the iterative gradient equals the
dense closed form (an exact linear solve of the adjoint system) and central
finite differences of the dense Perron-vector forward (numpy eig), so a
dropped projection, a transposed operand or a wrong sign fails.
"""
from __future__ import annotations

import numpy as np


def csr_matrix(n, rowptr, colidx, values):
    import scipy.sparse as sp
    return sp.csr_matrix((np.asarray(values, dtype=np.float64), np.asarray(colidx), np.asarray(rowptr)), shape=(n, n))


def fixed_point(A, beta=0.85, v=None, tol=1e-15, max_iteration=5000):
    """Sum-normalised Perron vector of M by power iteration (FP64)."""
    n = A.shape[0]
    v = np.full(n, 1.0 / n) if v is None else np.asarray(v, dtype=np.float64)
    p = np.full(n, 1.0 / n)
    for _ in range(max_iteration):
        q = beta * (A @ p) + (1.0 - beta) * v * p.sum()
        q /= q.sum()
        d = np.abs(q - p).sum()
        p = q
        if d <= tol:
            break
    return p


def adjoint(A, p, w, beta=0.85, v=None, tol=1e-15, max_iteration=5000):
    """(grad_values in CSR entry order of A, grad_teleport, iterations)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    v = np.full(n, 1.0 / n) if v is None else np.asarray(v, dtype=np.float64)
    sp_ = p.sum()
    s = beta * (A @ p).sum() + (1.0 - beta) * v.sum() * sp_
    AT = A.T.tocsr()
    lam = np.array(w, dtype=np.float64)
    it = 0
    for it in range(1, max_iteration + 1):
        u = (lam - p @ lam) / s
        new = beta * (AT @ u) + (1.0 - beta) * (v @ u) + w
        d = np.abs(new - lam).sum()
        lam = new
        if d <= tol * np.abs(lam).sum():
            break
    g = (lam - p @ lam) / s
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    gv = beta * g[rows] * p[A.indices]
    gt = (1.0 - beta) * sp_ * g
    return gv, gt, it


def dense_fixed_point(Ad, beta=0.85, v=None):
    """Perron vector of the dense M (numpy eig), sum 1."""
    n = Ad.shape[0]
    v = np.full(n, 1.0 / n) if v is None else np.asarray(v, dtype=np.float64)
    M = beta * Ad + (1.0 - beta) * np.outer(v, np.ones(n))
    ev, V = np.linalg.eig(M)
    k = int(np.argmax(ev.real))
    p = np.real(V[:, k])
    return p / p.sum()


def dense_adjoint(Ad, w, beta=0.85, v=None):
    """Closed form at tiny n: lambda = (I - J^T)^{-1} w with
    J = (I - p 1^T) M / s; returns dense dL/dA (n x n) and dL/dv."""
    n = Ad.shape[0]
    v = np.full(n, 1.0 / n) if v is None else np.asarray(v, dtype=np.float64)
    p = dense_fixed_point(Ad, beta, v)
    M = beta * Ad + (1.0 - beta) * np.outer(v, np.ones(n))
    s = (M @ p).sum()
    J = (np.eye(n) - np.outer(p, np.ones(n))) @ M / s
    lam = np.linalg.solve(np.eye(n) - J.T, w)
    g = (lam - p @ lam) / s
    return beta * np.outer(g, p), (1.0 - beta) * p.sum() * g, p
