"""tbsla_amd -- MI355X (gfx950) build of the TBSLA SpMV and PageRank path.

The product is libtbsla_hip.so (C ABI: include/tbsla_hip.h and the
reference's C interface include/tbsla_cpp_c.h; C++ API: include/tbsla/cpp).
This package holds its sources (csrc/), the build recipe (build.py) and the
Python mirror of the reference classes (matrix.py).  The library is loaded on
first use of a matrix class, so ``python -m tbsla_amd.build`` works before it
exists.
"""
_MATRIX = ["FORMATS", "Matrix", "MatrixCOO", "MatrixSCOO", "MatrixCSR", "MatrixELL", "MatrixDENSE"]
_BATCH = ["PageRankBatch", "personalised_workload", "u01", "u01_array", "normal_array"]
__all__ = _MATRIX + _BATCH


def __getattr__(name):
    if name in _MATRIX:
        from . import matrix
        return getattr(matrix, name)
    if name in _BATCH:
        from . import batch
        return getattr(batch, name)
    raise AttributeError(name)
