"""syngular.mul (tensor/utils.py:8-99) on device-resident operands."""
from __future__ import annotations

from .matrix_product_operator import MatrixProductOperator
from .matrix_product_state import MatrixProductState


def mul(op1, op2, mode="standard"):
    """Standard mode of tensor/utils.py:27-73: MPS x MPO and MPO x MPS are the
    MPO applied to the MPS, MPS x MPS is the overlap, MPO x MPO is op2 @ op1;
    each product finishes with >> min_bond.  The other modes are empty stubs in
    the reference (:82-99) and return None here as well."""
    n = op1.sites_number
    m = op2.sites_number
    if n != m:
        raise Exception("both operator do not have the same number of sites")
    if mode != "standard":
        return None
    if isinstance(op1, MatrixProductState) and isinstance(op2, MatrixProductOperator):
        if not op1.decomposed or not op2.decomposed:
            raise Exception("Operators and States must be decomposed")
        return op2 @ op1
    if isinstance(op1, MatrixProductOperator) and isinstance(op2, MatrixProductState):
        if not op1.decomposed or not op2.decomposed:
            raise Exception("Operators and States must be decomposed")
        return op1 @ op2
    if isinstance(op1, MatrixProductState) and isinstance(op2, MatrixProductState):
        return op1 | op2
    if isinstance(op1, MatrixProductOperator) and isinstance(op2, MatrixProductOperator):
        return op2 @ op1
    raise Exception("`syn.mul` should be provided MatrixProductState or MatrixProductOperator objects only")
