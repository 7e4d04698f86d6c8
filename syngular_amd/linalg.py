"""Device sweeps over matrix-product cores, shared by MPS and MPO.

A core is a float64 device tensor (l, *mid, r): mid = (d,) for an MPS core and
(d_in, d_out) for an MPO core.  Its left unfolding is (l*prod(mid)) x r, its
right unfolding l x (prod(mid)*r); both are views of the contiguous core.
Every arithmetic step runs in the gfx950 kernels of libsyngular_amd.so
(Householder QR, MFMA GEMM); torch only allocates and reshapes.
"""
from __future__ import annotations

import math

import torch

from . import ops

F64 = torch.float64


def reshape(t, shape):
    """Row-major reshape that raises ValueError (as numpy.reshape does) when
    the sizes do not match."""
    shape = tuple(int(s) for s in shape)
    if math.prod(shape) != t.numel():
        raise ValueError(f"cannot reshape array of size {t.numel()} into shape {shape}")
    return t.contiguous().reshape(shape)


def left_unfold(core, shape):
    """Reference left matricization: (shape[1]*...*shape[0], -1)."""
    rows = math.prod(shape[:-1])
    if rows == 0 or core.numel() % rows:
        raise ValueError(f"cannot reshape array of size {core.numel()} into shape ({rows},newaxis)")
    return core.contiguous().reshape(rows, -1)


def right_unfold(core, shape):
    """Reference right matricization: (-1, shape[1]*...*shape[-1])."""
    cols = math.prod(shape[1:])
    if cols == 0 or core.numel() % cols:
        raise ValueError(f"cannot reshape array of size {core.numel()} into shape (newaxis,{cols})")
    return core.contiguous().reshape(-1, cols)


def qr_keep(L, keep):
    """First `keep` columns of the complete Householder Q of L and the matching
    rows of R (numpy.linalg.qr(L, 'complete') then Q[:, :keep], R[:keep]).
    Returns (Q (m x kq), S (kq x n)) with kq = min(keep, m)."""
    Q = ops.householder_q(L, keep)
    S = ops.gemm(Q.t(), L)
    return Q, S


def qr_truncate_sweep(sites, shape, dim):
    """Reference QR truncation (tensor/matrix_product_state.py:439-466 and
    tensor/matrix_product_operator.py:551-578): for each bond keep the span of
    the first `dim` columns of the left unfolding.  `sites` and `shape` are
    rewritten in place; returns the parameter count."""
    params = 0
    for k in range(len(sites) - 1):
        L = left_unfold(sites[k], shape[k])
        Rm = right_unfold(sites[k + 1], shape[k + 1])
        Q, S = qr_keep(L, dim)
        W = ops.gemm(S, Rm)
        shape[k] = tuple(shape[k][:-1]) + (dim,)
        shape[k + 1] = (dim,) + tuple(shape[k + 1][1:])
        sites[k] = reshape(Q, shape[k])
        sites[k + 1] = reshape(W, shape[k + 1])
        params += sites[k].numel()
    return params + sites[-1].numel()


def left_orthonormalize(sites, shape):
    """Reduced-QR left sweep with shapes kept (tensor/matrix_product_state.py:554-566,
    tensor/matrix_product_operator.py:673-694)."""
    for k in range(len(sites) - 1):
        L = left_unfold(sites[k], shape[k])
        Rm = right_unfold(sites[k + 1], shape[k + 1])
        m, r = L.shape
        U, B = qr_keep(L, r)
        sites[k] = reshape(U, shape[k])
        sites[k + 1] = reshape(ops.gemm(B, Rm), shape[k + 1])


def right_orthonormalize(sites, shape):
    """Reduced-QR right sweep (tensor/matrix_product_state.py:568-580,
    tensor/matrix_product_operator.py:696-719): V, U = qr(R.T);
    site k = V.T, site k-1 = L @ U.T."""
    for k in range(len(sites) - 1, 0, -1):
        Rm = right_unfold(sites[k], shape[k])
        Lm = left_unfold(sites[k - 1], shape[k - 1])
        l = Rm.shape[0]
        Vt = ops.householder_q(Rm.t(), l, q_trans=True)   # (kq x d*r) = V^T
        RQ = ops.gemm(Rm, Vt.t())                          # R V = U^T (l x kq)
        sites[k] = reshape(Vt, shape[k])
        sites[k - 1] = reshape(ops.gemm(Lm, RQ), shape[k - 1])


def decompose_left(T, shape):
    """Left QR-truncation TT of a dense device tensor
    (tensor/matrix_product_state.py:298-322, matrix_product_operator.py:428-453).
    Returns the site list and the parameter count."""
    n = len(shape)
    sites = [None] * n
    params = 0
    for k in range(n - 1):
        rows = math.prod(shape[k][:-1])
        L = reshape(T, (rows, T.numel() // rows)) if T.numel() % rows == 0 else None
        if L is None:
            raise ValueError(f"cannot reshape array of size {T.numel()} into shape ({rows},newaxis)")
        Q, T = qr_keep(L, shape[k][-1])
        sites[k] = reshape(Q, shape[k])
        params += sites[k].numel()
    sites[-1] = reshape(T, shape[-1])
    return sites, params + sites[-1].numel()


def decompose_right(T, shape):
    """Right-mode decomposition of an MPS (tensor/matrix_product_state.py:324-347):
    QR of the transposed right unfolding, keep shape[k][0]."""
    n = len(shape)
    sites = [None] * n
    params = 0
    for k in range(n - 1, 0, -1):
        cols = math.prod(shape[k][1:])
        if T.numel() % cols:
            raise ValueError(f"cannot reshape array of size {T.numel()} into shape (newaxis,{cols})")
        Rm = T.contiguous().reshape(-1, cols)            # (rest x d*rrank)
        rank = shape[k][0]
        Qt = ops.householder_q(Rm.t(), rank, q_trans=True)   # (kq x d*rrank)
        sites[k] = reshape(Qt, shape[k])
        T = ops.gemm(Rm, Qt.t())                          # R[:rank].T = Rm @ Q  (rest x kq)
        params += sites[k].numel()
    sites[0] = reshape(T, shape[0])
    return sites, params + sites[0].numel()


def overlap(xs, ys):
    """Bilinear <X|Y> (tensor/matrix_product_state.py:116-129) as a streamed
    transfer-matrix sweep of two GEMMs per site:
        T = E^T X_k  (right unfolding of X_k),  E' = T^T Y_k (left unfolding)."""
    if xs[0].shape[0] != 1 or ys[0].shape[0] != 1 or xs[-1].shape[-1] != 1 or ys[-1].shape[-1] != 1:
        raise ValueError("cannot reshape array: chain boundary bonds must be 1")
    dev = xs[0].device
    E = torch.ones((1, 1), dtype=F64, device=dev)
    for X, Y in zip(xs, ys):
        if X.shape[1:-1] != Y.shape[1:-1]:
            raise ValueError("operands could not be broadcast together: physical dimensions differ")
        lx, rx = X.shape[0], X.shape[-1]
        ly, ry = Y.shape[0], Y.shape[-1]
        if E.shape != (lx, ly):
            raise ValueError("operands could not be broadcast together: bond dimensions differ")
        Xr = X.contiguous().reshape(lx, -1)
        T = ops.gemm(E.t(), Xr)                  # (ly x mid*rx)
        Tl = T.reshape(-1, rx)                   # ((ly*mid) x rx)
        Yl = Y.contiguous().reshape(-1, ry)      # ((ly*mid) x ry)
        E = ops.gemm(Tl.t(), Yl)                 # (rx x ry)
    return E


def dense(sites):
    """Contract a chain to its dense tensor (l=r=1), layout (mid_0, mid_1, ...)."""
    T = sites[0].contiguous().reshape(-1, sites[0].shape[-1])
    for s in sites[1:]:
        R = s.contiguous().reshape(s.shape[0], -1)
        T = ops.gemm(T, R).reshape(-1, s.shape[-1])
    mids = []
    for s in sites:
        mids += list(s.shape[1:-1])
    return T.reshape([sites[0].shape[0]] + mids + [sites[-1].shape[-1]])


def row_product(mats):
    """Product of a chain of (device) matrices, left to right."""
    v = mats[0]
    for M in mats[1:]:
        v = ops.gemm(v, M)
    return v


def frob_norm2(t):
    """||t||_F^2 as a (1 x 1) device tensor (one GEMM)."""
    flat = t.contiguous().reshape(1, -1)
    return ops.gemm(flat, flat.t())
