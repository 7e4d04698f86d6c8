"""Oracle restatement of the reference's qubit register (quantum/qbit.py,
quantum/circuit.py, quantum/gate.py) over the oracle MPS (reference.py).

Test infrastructure only (see oracle/__init__.py).  Pinned in
tests/test_oracle.py against a dense state-vector simulation of the same
circuits (Bell pair, Bernstein-Vazirani, SWAP networks, Toffoli).
"""
from __future__ import annotations

import numpy as np

from .reference import MPO, MPS

# quantum/gate.py:1-61, [out, in] matrices; two- and three-qubit gates as
# (2,)*2k leg tensors, qubit 0 the most significant bit
I = np.eye(2)
X = np.array([[0.0, 1.0], [1.0, 0.0]])
Z = np.array([[1.0, 0.0], [0.0, -1.0]])
H = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
CX = np.eye(4)[[0, 1, 3, 2]].reshape((2, 2, 2, 2))
SWAP = np.eye(4)[[0, 2, 1, 3]].reshape((2, 2, 2, 2))
TOFFOLI = np.eye(8)[[0, 1, 2, 3, 4, 5, 7, 6]].reshape((2,) * 6)


class Qbit:
    LSB = True

    def __init__(self, size, init=True):
        # quantum/qbit.py:14-28
        self.size = size
        self.dim = 2 ** size
        self.state = None
        if init:
            self.state = MPS.zeros((2,) * size, (2,) * (size - 1))
            for idx in range(self.state.sites_number - 1):
                self.state.sites[idx][0] = I
            self.state.sites[-1][0][0] = 1.0

    def __matmul__(self, operator):
        # quantum/qbit.py:30-61
        if isinstance(operator, tuple):
            if len(operator) == 2:
                return Qbit.from_mps(self.state.apply(*operator))
            lo, hi = min(operator[1], operator[2]), max(operator[1], operator[2])
            q = self
            if hi - lo != 1:
                q = self.swap_in(lo, hi - 1)
                if lo != operator[1]:
                    q = q @ (SWAP, hi - 1)
            q = Qbit.from_mps(q.state.apply(operator[0], hi - 1))
            if hi - lo != 1:
                q = q.swap_out(lo, hi - 1)
            return q
        return Qbit.from_mps(MPO(operator, bond_shape=()).decompose() @ self.state)

    def swap(self, a, b):
        return self.swap_in(a, b).swap_out(a, b)

    def swap_in(self, a, b):
        # quantum/qbit.py:71-86
        q = self
        for i in range(min(a, b), max(a, b)):
            q = q @ (SWAP, i)
        return q

    def swap_out(self, a, b):
        # quantum/qbit.py:88-103
        q = self
        for i in range(max(a, b), min(a, b) - 1, -1):
            q = q @ (SWAP, i)
        return q

    def to_tensor(self):
        return self.state.to_tensor().reshape(self.dim)

    def to_binary(self):
        t = self.to_tensor().astype(int)
        bits = bin(np.where(t == 1)[0][0])[2:].zfill(self.size)
        return bits[::-1] if Qbit.LSB else bits

    @staticmethod
    def from_mps(mps):
        q = Qbit(mps.sites_number, init=False)
        q.state = mps
        return q

    @staticmethod
    def from_binary(bits):
        q = Qbit(len(bits))
        for i, b in enumerate(bits):
            if b == "1":
                q = q @ (X, i)
        return q


def run_circuit(size, structure):
    """quantum/circuit.py:20-39: ground state, then every step in order."""
    q = Qbit(size)
    for op in structure:
        q = q @ op
    return q
