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
    previous block (its right bonds are summed over: a block of ones)."""
    n = len(operator_sites)
    blocks = [None] * n
    for k in range(n - 1, 1, -1):
        X, W = state_sites[k], operator_sites[k]
        R = blocks[k + 1] if k != n - 1 else np.ones((X.shape[2], W.shape[3], X.shape[2]))
        # pairwise, cheapest first: (e,t,f)(b,d,f) -> (e,t,b,d); with W -> (e,b,c,s); with X -> (a,e,c)
        T = np.tensordot(X, R, axes=([2], [2]))
        T = np.tensordot(T, W, axes=([1, 3], [1, 3]))
        blocks[k] = np.tensordot(X, T, axes=([1, 2], [3, 1])).transpose(0, 2, 1)
    return blocks


def left_blocks(operator_sites, state_sites):
    """variational/dmrg.py:89-110: blocks[k] (k = 0 .. n-3), indices [ket,
    operator, bra] of bond k+1's right side; the first one contracts no
    previous block (its left bonds are summed over: a block of ones)."""
    n = len(operator_sites)
    blocks = [None] * n
    for k in range(n - 2):
        X, W = state_sites[k], operator_sites[k]
        L = blocks[k - 1] if k != 0 else np.ones((X.shape[0], W.shape[0], X.shape[0]))
        # pairwise: (a,c,e)(a,s,b) -> (c,e,s,b); with W -> (e,b,t,d); with X -> (b,d,f)
        T = np.tensordot(L, X, axes=([0], [0]))
        T = np.tensordot(T, W, axes=([0, 2], [0, 2]))
        blocks[k] = np.tensordot(T, X, axes=([0, 2], [0, 1]))
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
