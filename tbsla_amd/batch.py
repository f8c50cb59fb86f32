"""Batched personalised PageRank over the C ABI (tbsla_batch_*).

    b = PageRankBatch("f32")
    b.add_graph(g1, [[17]])                 # one teleport set: block-diagonal product
    b.add_graph(big, [[n0], [n1], ...])     # many sets: SpMM, 64 per group
    out, rep, its = b.run(epsilon=1e-6, rule="sumnorm")
    p_k = out[off[k]:off[k+1]]              # off = b.layout()

Each teleport set follows Matrix::personalized_page_rank (Matrix.cpp:63-114)
and stops on its own; the iteration runs as a replayed hipGraph of one
persistent work-queue kernel plus per-set reductions.

Also here: the definition of the 256-graph + 2^22 x 64 benchmark batch
(graph sizes and teleport nodes from a SplitMix64 counter generator).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check, lib
from .matrix import MatrixCSR, _stream_handle, dtype_code

_M64 = (1 << 64) - 1


def _splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & _M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def u01(seed: int, i: int) -> float:
    """(splitmix64(splitmix64(seed) ^ i) >> 11) * 2^-53: the generator of the
    value overlay and of x (same definition as the C++ generators)."""
    return (_splitmix64(_splitmix64(seed) ^ i) >> 11) * 2.0 ** -53


def u01_array(seed: int, first: int, n: int) -> np.ndarray:
    """u01(seed, i) for i in [first, first + n), vectorised."""
    i = np.arange(first, first + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(_splitmix64(seed)) ^ i) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def normal_array(seed: int, n: int) -> np.ndarray:
    """N(0,1) by Box-Muller: sqrt(-2 ln(1 - u01(seed, 2i))) cos(2 pi u01(seed, 2i+1))."""
    u = u01_array(seed, 0, 2 * n)
    return np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])


def personalised_workload(n_graphs: int = 256, big_n: int = 1 << 22, big_rhs: int = 64):
    """The batch: graph g (seed 1000+g) has n_g = floor(2^(10 + 8 u01(seed, 0)) + 0.5)
    nodes, c = 12, q = 0.3 and one teleport node floor(u01(seed, 1) * n_g);
    the big graph has n = big_n, c = 16, q = 0.3, seed 1256 and teleport
    nodes floor(u01(1256, r) * n), r < big_rhs.  Returns a list of
    (n, c, q, seed, [node sets])."""
    out = []
    for g in range(n_graphs):
        seed = 1000 + g
        n = int(math.floor(2.0 ** (10 + 8 * u01(seed, 0)) + 0.5))
        out.append((n, 12, 0.3, seed, [[int(u01(seed, 1) * n)]]))
    if big_rhs > 0:
        out.append((big_n, 16, 0.3, 1256, [[int(u01(1256, r) * big_n)] for r in range(big_rhs)]))
    return out


class PageRankBatch:
    def __init__(self, precision="f32"):
        self._h = lib.tbsla_batch_create(dtype_code(precision))
        if not self._h:
            raise _lib.TbslaError(-1, lib.tbsla_last_error().decode(errors="replace"))
        self._keep = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:  # module globals are cleared at interpreter exit
            lib.tbsla_batch_destroy(h)
            self._h = None

    def add_graph(self, matrix: MatrixCSR, node_sets):
        """node_sets: one list of nodes per teleport set (an int is a one-node set)."""
        sets = [[s] if isinstance(s, (int, np.integer)) else list(s) for s in node_sets]
        ptr = np.zeros(len(sets) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(s) for s in sets])
        nodes = np.array([v for s in sets for v in s] or [0], dtype=np.int32)
        check(lib.tbsla_batch_add_graph(self._h, matrix._h, _lib.iptr(ptr), _lib.iptr(nodes), len(sets)))
        self._keep.append(matrix)  # read again by prepare()
        return self

    def prepare(self):
        check(lib.tbsla_batch_prepare(self._h))
        self._keep = []
        return self

    def n_rhs(self) -> int:
        return lib.tbsla_batch_n_rhs(self._h)

    def layout(self) -> np.ndarray:
        o = np.zeros(self.n_rhs() + 1, dtype=np.int64)
        check(lib.tbsla_batch_layout(self._h, o.ctypes.data_as(C.POINTER(C.c_long)), o.size))
        return o

    def run(self, out=None, epsilon=1e-10, beta=0.85, max_iteration=1000, rule="reference", check_every=16,
            graph=True, autotune=True, stream=None):
        """Returns (out, report, per-set iteration counts); out is a float64
        CUDA tensor of layout()[-1] values (allocated when None)."""
        import torch
        off = self.layout()
        if out is None:
            out = torch.empty(int(off[-1]), dtype=torch.float64, device="cuda")
        if not (out.is_cuda and out.dtype == torch.float64 and out.is_contiguous()):
            raise TypeError("out must be a contiguous float64 CUDA tensor")
        p = _lib.BatchParams()
        p.epsilon, p.beta, p.max_iteration = epsilon, beta, max_iteration
        p.rule = {"reference": 0, "sumnorm": 1}[rule] if isinstance(rule, str) else int(rule)
        p.check_every, p.graph_capture, p.autotune = check_every, int(bool(graph)), int(bool(autotune))
        its = np.zeros(max(1, self.n_rhs()), dtype=np.int32)
        rep = _lib.BatchReport()
        check(lib.tbsla_batch_run(self._h, C.byref(p), out.data_ptr(), out.numel(), _lib.iptr(its), its.size,
                                  _stream_handle(stream), C.byref(rep)))
        self._keep = []
        report = {"iterations": rep.iterations, "seconds": rep.seconds, "tune_seconds": rep.tune_seconds,
                  "tune_iterations": rep.tune_iterations, "tune_cached": bool(rep.tune_cached),
                  "bytes_per_iteration": rep.bytes_per_iteration,
                  "config": {"wg_per_cu": rep.wg_per_cu, "unroll": rep.unroll, "xcd_map": rep.xcd_map,
                             "tile_nnz": rep.tile_nnz, "spmm_rows": rep.spmm_rows}}
        return out, report, its[: self.n_rhs()]
