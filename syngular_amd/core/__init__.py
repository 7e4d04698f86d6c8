"""Tensor-train layers over the device matrix products (syngular.core)."""
from .model import Layer, Linear, Model, Output, matmul_batch

__all__ = ["Model", "Layer", "Linear", "Output", "matmul_batch"]
