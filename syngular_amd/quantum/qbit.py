"""Qubit register held as a matrix product state on the device
(drop-in for syngular.quantum.Qbit, quantum/qbit.py:9-145).

The register of ``size`` qubits is a bond-2 MPS; gates go through
``MatrixProductState.apply``, which keeps every bond at 2 (the reference's
QR split).  Behaviour follows the reference, including:

* ``qbit @ (G, i)`` applies G on sites i.. (one- or multi-site gate);
* ``qbit @ (G, a, b)`` for a two-site gate on qubits a and b: when they are
  not neighbours, SWAPs bring qubit min(a, b) next to max(a, b), one more
  SWAP puts the control (a) first when a > b, the gate acts on the pair, and
  the SWAPs are undone in the reference's order (qbit.py:37-58);
* ``qbit @ G`` with a bare tensor goes through a bond-less
  MatrixProductOperator and ``@`` (qbit.py:59-61), which the reference's
  ``@`` rejects for bond-less operators (``min`` of an empty bond shape);
* the reference's console prints are dropped.
"""
from __future__ import annotations

import numpy as np
import torch

from ..tensor import MatrixProductOperator, MatrixProductState
from . import gate


class Qbit:
    VERBOSE = 0
    LSB = True

    def __init__(self, size, init=True):
        # quantum/qbit.py:14-28: |0...0> as bond-2 cores whose first slice
        # is the identity (the last core holds the single 1)
        self.size = size
        self.dim = 2 ** self.size
        self.state = None
        if init:
            self.state = MatrixProductState.zeros((2,) * size, (2,) * (size - 1)).decompose()
            self.state.real_parameters_number = self.dim
            eye = torch.eye(2, dtype=torch.float64, device=self.state.sites[0].device)
            for idx in range(self.state.sites_number - 1):
                self.state.sites[idx][0].copy_(eye)
            self.state.sites[-1][0][0].fill_(1.0)

    def __matmul__(self, operator):
        # quantum/qbit.py:30-61
        if isinstance(operator, tuple):
            if len(operator) == 2:
                return Qbit.from_mps(self.state.apply(*operator))
            min_index = min(operator[1], operator[2])
            max_index = max(operator[1], operator[2])
            qbit = self
            if abs(min_index - max_index) != 1:
                qbit = self.swap_in(min_index, max_index - 1)
                if min_index != operator[1]:
                    qbit @= (gate.SWAP, max_index - 1)
            qbit = Qbit.from_mps(qbit.state.apply(operator[0], max_index - 1))
            if abs(min_index - max_index) != 1:
                qbit = qbit.swap_out(min_index, max_index - 1)
            return qbit
        op = MatrixProductOperator(operator, bond_shape=()).decompose()
        return Qbit.from_mps(op @ self.state)

    def __imatmul__(self, operator):
        return self @ operator

    def swap(self, idx1, idx2):
        # quantum/qbit.py:66-69
        self = self.swap_in(idx1, idx2)
        self = self.swap_out(idx1, idx2)
        return self

    def swap_in(self, idx1, idx2):
        # quantum/qbit.py:71-86: SWAP at min .. max-1
        for i in range(min(idx1, idx2), max(idx1, idx2)):
            self @= (gate.SWAP, i)
        return self

    def swap_out(self, idx1, idx2):
        # quantum/qbit.py:88-103: SWAP at max .. min, descending
        for i in range(max(idx1, idx2), min(idx1, idx2) - 1, -1):
            self @= (gate.SWAP, i)
        return self

    def apply(self, operator):
        # quantum/qbit.py:105-107
        op = MatrixProductOperator(operator, bond_shape=()).decompose()
        return Qbit.from_mps(op @ self.state)

    def to_tensor(self):
        return self.state.to_tensor().reshape(self.dim)

    def to_binary(self):
        # quantum/qbit.py:112-131: index of the amplitude equal to 1 after an
        # integer cast, as a bit string (reversed when LSB)
        tensor = self.to_tensor().astype(int)
        bits = bin(np.where(tensor == 1)[0][0])[2:].zfill(self.size)
        return bits[::-1] if Qbit.LSB else bits

    @staticmethod
    def from_mps(mps):
        qbit = Qbit(mps.sites_number, init=False)
        qbit.state = mps
        return qbit

    @staticmethod
    def from_binary(bin):
        # quantum/qbit.py:138-145
        size = len(bin)
        qbit = Qbit(size)
        for i in range(size):
            if bin[i] == "1":
                qbit @= (gate.X, i)
        return qbit
