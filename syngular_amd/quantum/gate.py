"""Gate tensors in the [outputs..., inputs...] convention of
``MatrixProductState.apply`` (quantum/gate.py): a k-qubit gate is the
(2^k x 2^k) matrix G with out = G @ in, reshaped to 2k legs of dimension 2.

The device path computes in float64, so the complex phase gates (Y, T, S)
are kept here for completeness but are rejected by ``apply``.
"""
import numpy as np


def _permutation(perm):
    """Permutation matrix with G[perm[i], i] = 1 (basis state i -> perm[i])."""
    G = np.zeros((len(perm), len(perm)))
    G[perm, np.arange(len(perm))] = 1.0
    return G


def _legs(G):
    k = int(round(np.log2(G.shape[0])))
    return G.reshape((2,) * (2 * k))


I = np.eye(2)
X = _permutation([1, 0])
Y = np.array([[0.0, -1.0j], [1.0j, 0.0]])
Z = np.diag([1.0, -1.0])
H = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
T = np.diag([1.0, np.exp(0.25j * np.pi)])
S = np.diag([1.0, np.exp(0.5j * np.pi)])

# two- and three-qubit gates; qubit 0 is the most significant bit of the
# basis index (the first leg)
CX = _legs(_permutation([0, 1, 3, 2]))                      # flip bit 1 when bit 0 is set
SWAP = _legs(_permutation([0, 2, 1, 3]))                    # exchange bits 0 and 1
TOFFOLI = _legs(_permutation([0, 1, 2, 3, 4, 5, 7, 6]))     # flip bit 2 when bits 0 and 1 are set

__all__ = ["I", "X", "Y", "Z", "H", "T", "S", "CX", "SWAP", "TOFFOLI"]
