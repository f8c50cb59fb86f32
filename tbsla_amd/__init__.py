"""tbsla_amd -- MI355X (gfx950) build of the TBSLA SpMV and PageRank path.

The product is libtbsla_hip.so (C ABI: include/tbsla_hip.h and the
reference's C interface include/tbsla_cpp_c.h; C++ API: include/tbsla/cpp).
This package holds its sources (csrc/), the build recipe (build.py) and the
Python mirror of the reference classes (matrix.py).  The library is loaded on
first use of a matrix class, so ``python -m tbsla_amd.build`` works before it
exists.
"""
__all__ = ["FORMATS", "Matrix", "MatrixCOO", "MatrixSCOO", "MatrixCSR", "MatrixELL", "MatrixDENSE"]


def __getattr__(name):
    if name in __all__:
        from . import matrix
        return getattr(matrix, name)
    raise AttributeError(name)
