"""DMRG environment blocks and the two-site restore split on the device
(drop-in for the block builders of syngular.variational.DMRG,
variational/dmrg.py:65-112, and DMRG.__restore, :219-239).

A block is one site of the sandwich <X| W |X> with the neighbouring block
folded in; indices [ket bond, operator bond, bra bond].  The ket's physical
leg meets the operator's output leg and the bra's its input leg
(variational/dmrg.py:75-79).  Per site, on the gfx950 kernels:

1. Z[(c,e), s, (d,f)] = sum_t W[c,t,s,d] X[e,t,f]: the MPO x MPS apply kernel
   with W as the chain core (W[c, t, (s,d)]) and the bra X as a one-output
   operator core (X[e, t, 0, f]), so the bond pairs come out (operator, bra)
   and every later product is a plain GEMM in the reference's index order;
2. right block: U[(c,e) s, b] = Z[(c,e) s, (d,f)] R[b, (d,f)]^T, then
   R'[a, (c,e)] = X[a, (s,b)] U[(c,e), (s,b)]^T;
   left block: U[a, s (d,f)] = L[a, (c,e)] Z[(c,e), s (d,f)], then
   L'[b, (d,f)] = X[(a,s), b]^T U[(a,s), (d,f)].

The outermost blocks contract no neighbour: their free outer bonds are
summed, as the reference's contraction without that operand does, which is
a neighbour block of ones.  The reference's sweep (``solve``) is not part of
this path (see DMRG.solve).
"""
from __future__ import annotations

import torch

from .. import linalg, ops

F64 = torch.float64


def _sandwich_core(X, W):
    """Z[(c,e), s, (d,f)] = sum_t W[c,t,s,d] X[e,t,f] as ((c,e) s, (d,f))."""
    c, t, s, d = W.shape
    e, t2, f = X.shape
    if t != t2 or s != t:
        raise ValueError(f"operands could not be broadcast: operator legs {tuple(W.shape)} vs state {tuple(X.shape)}")
    Wc = W.contiguous().reshape(c, t, s * d)
    Xo = X.contiguous().reshape(e, t, 1, f)
    Z = ops.apply_many([(Xo, Wc)])[0]                # (c*e, 1, s*d*f)
    return Z.reshape(c * e * s, d * f)


def right_blocks(operator, state):
    """variational/dmrg.py:65-86: list of n entries, blocks[k] for
    k = n-1 .. 2 as (ket, operator, bra) device tensors, None elsewhere."""
    n = operator.sites_number
    blocks = [None] * n
    for k in range(n - 1, 1, -1):
        X, W = state.sites[k], operator.sites[k]
        a, s, b = X.shape
        c, d = W.shape[0], W.shape[3]
        Z = _sandwich_core(X, W)
        if k == n - 1:
            R = torch.ones((b, d * b), dtype=F64, device=X.device)
        else:
            R = blocks[k + 1].reshape(b, d * b)
        U = ops.gemm(Z, R.t())                                     # ((c,e) s, b)
        E = ops.gemm(X.reshape(a, s * b), U.reshape(c * a, s * b).t())
        blocks[k] = E.reshape(a, c, a)
    return blocks


def left_blocks(operator, state):
    """variational/dmrg.py:89-110: list of n entries, blocks[k] for
    k = 0 .. n-3 as (ket, operator, bra) device tensors, None elsewhere."""
    n = operator.sites_number
    blocks = [None] * n
    for k in range(n - 2):
        X, W = state.sites[k], operator.sites[k]
        a, s, b = X.shape
        c, d = W.shape[0], W.shape[3]
        Z = _sandwich_core(X, W)
        if k == 0:
            L = torch.ones((a, c * a), dtype=F64, device=X.device)
        else:
            L = blocks[k - 1].reshape(a, c * a)
        U = ops.gemm(L, Z.reshape(c * a, s * d * b))             # (a, s (d,f))
        E = ops.gemm(X.reshape(a * s, b).t(), U.reshape(a * s, d * b))
        blocks[k] = E.reshape(b, d, b)
    return blocks


def restore(state, site, index):
    """variational/dmrg.py:219-239: split a two-site tensor (l, d, d', r) into
    sites index and index+1 of `state`'s shapes: complete Householder QR of
    the (l d, d' r) unfolding, the first rank columns of Q and rows of R."""
    if index >= state.sites_number:
        raise Exception("site index for restoration is out of range (between 0 & n-1)")
    site = site if isinstance(site, torch.Tensor) else torch.as_tensor(site, dtype=F64, device=ops.device())
    site = site.to(dtype=F64).contiguous()
    L = site.reshape(site.shape[0] * site.shape[1], -1)
    rank = state.shape[index][2]
    Q, S = linalg.qr_keep(L, rank)
    return linalg.reshape(Q, state.shape[index]), linalg.reshape(S, state.shape[index + 1])


class DMRG:
    """The reference's DMRG namespace (variational/dmrg.py:11).  The block
    builders and the restore split keep the reference's (name-mangled)
    names, so ``DMRG._DMRG__left_blocks(operator=..., state=...)`` works as
    it does there; the module functions are the public spelling."""

    @staticmethod
    def __right_blocks(operator, state):
        return right_blocks(operator, state)

    @staticmethod
    def __left_blocks(operator, state):
        return left_blocks(operator, state)

    @staticmethod
    def __restore(state, site, index):
        return restore(state, site, index)

    @staticmethod
    def solve(operator, optimizer=None):
        """Not on the device path: the reference's sweep (variational/dmrg.py:15-62,
        :114-216) feeds the Krylov basis that its Lanczos.fit returns back as
        the new two-site tensor (variational/optimizer.py:10-58), so it does
        not compute a ground state."""
        raise NotImplementedError("DMRG.solve is not part of the device path; "
                                  "left_blocks / right_blocks / restore are")
