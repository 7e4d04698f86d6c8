"""Qubit-register simulation on matrix product states (syngular.quantum):
``Qbit`` and ``Circuit`` over the device ``MatrixProductState.apply``."""
from .circuit import Circuit
from .qbit import Qbit

__all__ = ["Circuit", "Qbit"]
