"""Oracle restatement of the DMRG environment blocks
(variational/dmrg.py:65-112) and of the two-site restore split
(variational/dmrg.py:219-239).

Test infrastructure only (see oracle/__init__.py).  Pinned in
tests/test_oracle.py: for every bond k the left block of sites < k and the
right block of sites >= k contract to the dense bilinear form
sum_{s,t} x[s] W[t -> s] x[t] of the dense state x and operator W, and the
restore split reproduces the dense projector closed form.
"""
from __future__ import annotations

import numpy as np


def _sandwich(X, W):
    """Z[a, c, e, b, d, f] = sum_{s,t} X[a,s,b] W[c,t,s,d] X[e,t,f]: the ket's
    physical leg meets the operator's output leg, the bra's its input leg
    (variational/dmrg.py:75-79, :96-100)."""
    return np.einsum("asb,ctsd,etf->acebdf", X, W, X, optimize=True)


def right_blocks(operator_sites, state_sites):
    """variational/dmrg.py:65-86: blocks[k] (k = n-1 .. 2), indices
    [ket, operator, bra] of bond k's left side; the last one contracts no
    previous block (its right bonds are summed over)."""
    n = len(operator_sites)
    blocks = [None] * n
    for k in range(n - 1, 1, -1):
        Z = _sandwich(state_sites[k], operator_sites[k])
        if k == n - 1:
            blocks[k] = Z.sum(axis=(3, 4, 5))
        else:
            blocks[k] = np.einsum("acebdf,bdf->ace", Z, blocks[k + 1], optimize=True)
    return blocks


def left_blocks(operator_sites, state_sites):
    """variational/dmrg.py:89-110: blocks[k] (k = 0 .. n-3), indices [ket,
    operator, bra] of bond k+1's right side; the first one contracts no
    previous block (its left bonds are summed over)."""
    n = len(operator_sites)
    blocks = [None] * n
    for k in range(n - 2):
        Z = _sandwich(state_sites[k], operator_sites[k])
        if k == 0:
            blocks[k] = Z.sum(axis=(0, 1, 2))
        else:
            blocks[k] = np.einsum("acebdf,ace->bdf", Z, blocks[k - 1], optimize=True)
    return blocks


def restore(shapes, site, index):
    """variational/dmrg.py:219-239: split a two-site tensor (l, d, d', r)
    back into sites `index` and `index + 1` of the given shapes by a complete
    QR of its (l d, d' r) unfolding, keeping the first rank columns."""
    if index >= len(shapes):
        raise Exception("site index for restoration is out of range (between 0 & n-1)")
    L = np.reshape(site, (site.shape[0] * site.shape[1], -1))
    Q, R = np.linalg.qr(L, mode="complete")
    rank = shapes[index][2]
    return np.reshape(Q[:, :rank], shapes[index]), np.reshape(R[:rank, :], shapes[index + 1])
