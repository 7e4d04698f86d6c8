"""syngular_amd: the matrix-product core of Syngular (syngular.tensor) on MI355X.

``from syngular_amd.tensor import MatrixProductState, MatrixProductOperator``
and ``syngular_amd.mul`` mirror ``syngular.tensor`` / ``syngular.mul``; the
cores live in GPU memory and every arithmetic step runs in the gfx950 HIP
kernels of ``libsyngular_amd.so`` (C ABI: include/syngular_amd.h).
"""
from .tensor import MatrixProductOperator, MatrixProductState, mul

__all__ = ["MatrixProductState", "MatrixProductOperator", "mul"]
__version__ = "0.1.0"
