"""tbsla_hip_page_rank_noalloc, the PageRank entry point that allocates no
device memory (its buffers are slices of a static scratch buffer per GPU),
against tbsla_hip_page_rank with the same fixed-order tile products: equal
bit for bit, on two calls in a row (the second on a smaller graph, which
must not see the first call's leftovers), uniform and personalised; and its
refusals, which happen before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EINVAL = -1


def params(tb, **kw):
    from tbsla_amd import _lib
    p = _lib.PageRankParams(epsilon=1e-12, beta=0.85, max_iteration=400, precision=_lib.F64, rule=0,
                            check_every=8, anchor_tau=0.0, nodes=None, n_nodes=0, gather=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def run(fn, m, p, out):
    from tbsla_amd import _lib
    rep = _lib.PageRankReport()
    dst = ctypes.cast(ctypes.c_void_p(out.data_ptr()), ctypes.POINTER(ctypes.c_double))
    rc = fn(m._h, ctypes.byref(p), dst, out.numel(), None, ctypes.byref(rep))
    return rc, rep


def test_noalloc_matches_allocating_entry_point(tb, gpu):
    import torch
    from tbsla_amd import _lib
    lib = _lib.lib
    for n, seed in ((300_000, 7), (40_000, 3)):  # the second call is smaller
        print(f"noalloc n={n}", flush=True)
        m = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 8, 0.5, seed)
        m.prepare("f64")
        for rule in (0, 1):
            nodes = np.array([5, n // 2, n - 1], dtype=np.int32)
            for personal in (False, True):
                kw = {"rule": rule}
                if personal:
                    kw.update(nodes=nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n_nodes=3)
                ref = torch.full((n,), -1.0, dtype=torch.float64, device=gpu)
                out = torch.full((n,), -2.0, dtype=torch.float64, device=gpu)
                rc, r0 = run(lib.tbsla_hip_page_rank, m, params(tb, **kw), ref)
                assert rc == 0, lib.tbsla_last_error()
                rc, r1 = run(lib.tbsla_hip_page_rank_noalloc, m, params(tb, **kw), out)
                assert rc == 0, lib.tbsla_last_error()
                torch.cuda.synchronize()
                assert r1.iterations == r0.iterations and r1.last_error == r0.last_error, (n, rule, personal)
                assert torch.equal(out, ref), (n, rule, personal)


def test_noalloc_refusals(tb, gpu):
    import torch
    from tbsla_amd import _lib
    lib = _lib.lib
    n = 20_000
    m = tb.MatrixCSR().fill_cqmat_stochastic(n, n, 8, 0.5, 1)
    out = torch.zeros(n, dtype=torch.float64, device=gpu)
    rc, _ = run(lib.tbsla_hip_page_rank_noalloc, m, params(tb), out)
    assert rc == EINVAL and b"tbsla_hip_prepare" in lib.tbsla_last_error()  # mirror not built
    m.prepare("f64")
    for kw in ({"precision": _lib.F32}, {"gather": 2}, {"rule": 5}):
        rc, _ = run(lib.tbsla_hip_page_rank_noalloc, m, params(tb, **kw), out)
        assert rc == EINVAL, kw
    short = out[: n - 1]
    rc, _ = run(lib.tbsla_hip_page_rank_noalloc, m, params(tb), short)
    assert rc == EINVAL and b"PageRank writes" in lib.tbsla_last_error()
    # beyond the scratch buffer: refused before the mirror is even looked at
    big = 3_000_000
    g = tb.MatrixCSR().fill_cdiag(big, big, 1)
    ob = torch.zeros(big, dtype=torch.float64, device=gpu)
    rc, _ = run(lib.tbsla_hip_page_rank_noalloc, g, params(tb), ob)
    assert rc == EINVAL and b"bytes of scratch" in lib.tbsla_last_error()
