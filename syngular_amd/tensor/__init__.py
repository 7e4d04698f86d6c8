from .matrix_product_state import MatrixProductState
from .matrix_product_operator import MatrixProductOperator
from .utils import mul

__all__ = ["MatrixProductState", "MatrixProductOperator", "mul"]
