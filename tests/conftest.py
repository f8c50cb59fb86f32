import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: full-size workload (tens of seconds)")


@pytest.fixture(scope="session")
def digests():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "golden_small.npz")))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    if not o.Oracle.available():
        o.build(with_reference=False)
    assert o.Oracle.available(), "oracle/liborc.so missing (make -C oracle)"
    return o.Oracle


@pytest.fixture(scope="session")
def reference():
    from oracle import oracle as o
    if not o.Reference.available():
        pytest.skip("oracle/_ref/libtbsla_ref.so not built (needs the reference tree)")
    return o.Reference


@pytest.fixture(scope="session")
def tb():
    import tbsla_amd
    return tbsla_amd


@pytest.fixture(scope="session")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


class single_thread_omp:
    """Run the OpenMP regions started from this thread with one thread, so
    the reference's COO/SCOO vector reductions (MatrixCOO.cpp:109-112,
    MatrixSCOO.cpp:116-119) take their sequential order."""

    def __enter__(self):
        import ctypes
        self.gomp = ctypes.CDLL("libgomp.so.1")
        self.prev = self.gomp.omp_get_max_threads()
        self.gomp.omp_set_num_threads(1)
        return self

    def __exit__(self, *exc):
        self.gomp.omp_set_num_threads(self.prev)


def grid_q():
    """q values of the reference test grid (tests/test_spmv_cqmat.cpp:155)."""
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
FMTS = ["COO", "SCOO", "CSR", "ELL", "DENSE"]
INFO_KEYS = ["n_row", "n_col", "ln_row", "ln_col", "f_row", "f_col", "pr", "pc", "NR", "NC"]
