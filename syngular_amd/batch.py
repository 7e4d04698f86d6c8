"""Grouped apply + rounding of many independent chains of uneven shape.

A batch holds J chains (MPS) on the device.  Cores are not torch tensors of
their own: each lives at a device address inside an arena tensor, and the
batch keeps three host tables

    ptr[j, k]   device address of core k of chain j   (row-major (b_k, d_k, b_k+1))
    bond[j, k]  bond dimension b_k, k = 0..n_j         (b_0 = b_n = 1 normally)
    phys[j, k]  physical dimension d_k

so that every step of a sweep builds its kernel item tables with vectorised
numpy over all chains at once and issues ONE grouped launch per kernel,
whatever the mix of shapes.  The sweeps run in lock-step: step k of the left
sweep touches core k of every chain with more than k+1 sites; step t of the
right sweep touches core n_j-1-t of every chain, so chains of different
length finish at different steps but share every launch.

The algorithm per chain is the one of ``rounding.tt_round`` (left QR sweep,
then right Gram + one-sided Jacobi SVD sweep with truncation on the device),
so each chain's result is independent of what else is in the batch: the
kernels treat every item on its own, and a batch of one chain gives the same
bits as the same chain in a batch of thousands.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, autotune, ops
from ._lib import call

F64 = torch.float64
ALIGN = 32                      # doubles (256 B) between sub-allocations
# Gram matrices up to this size: one work-group (one wave) in LDS; larger
# ones take the block Jacobi.  Past 16 the LDS kernel's rounds are bound by
# LDS bandwidth (every round rewrites the whole matrix and its rotations),
# while the block method confines the rounds to 32 x 32 pair tiles and
# updates the rest on the MFMA: orders 17..64 there took the batch of 4096
# chains from 1255 to 1135 ms per step.
LDS_JACOBI_MAX = 16
CHOLQR_MIN_COLS = 96            # tall unfoldings with this many columns (or too big for LDS): CholeskyQR2
# Jacobi orthogonality thresholds, fixed per path (not per launch) so that a
# chain's rotations do not depend on the other chains of its batch: the
# rounding bound length * eps of the largest problem a path may take (the
# LDS kernel takes up to 64).
TOL_LDS = ops.jacobi_tol(64)
TOL_BLOCKED = ops.jacobi_tol(2048)
# Stopping threshold of the blocked Jacobi for the forms that work to an
# error budget of 1e-6 (the mixed form, and served jobs that ask for it):
# rotations stop once every |G_pq| <= 1e-8 sqrt(G_pp G_qq).  Eigenvalues then
# carry relative errors of order 1e-16 and the retained subspace moves by
# angles of order 1e-8 / (relative gap); the FP64 form keeps TOL_BLOCKED.
TOL_BUDGET = 1e-8


# --------------------------------------------------------------------------
# arenas
# --------------------------------------------------------------------------
def _round_up(x, a=ALIGN):
    return (x + a - 1) // a * a


def _empty(n, dtype=F64, zero=False):
    """Device tensor of n elements (every device allocation of this module
    goes through here)."""
    f = torch.zeros if zero else torch.empty
    return f(max(int(n), 1), dtype=dtype, device=ops.device())


def _h2d(host):
    """Host tensor -> device copy through pinned memory, asynchronous on the
    current stream."""
    return host.pin_memory().to(ops.device(), non_blocking=True)


def arena(sizes, dtype=F64):
    """One device tensor holding len(sizes) sub-buffers of the given element
    counts; returns (tensor, addresses (uint64 array))."""
    sizes = np.asarray(sizes, dtype=np.int64)
    padded = _round_up(np.maximum(sizes, 1))
    off = np.zeros(len(sizes), dtype=np.int64)
    if len(sizes):
        off[1:] = np.cumsum(padded)[:-1]
    total = int(padded.sum()) if len(sizes) else 1
    t = _empty(total, dtype)
    esz = t.element_size()
    return t, (np.uint64(t.data_ptr()) + (off * esz).astype(np.uint64))


def _vec(v, dtype, cnt):
    """v (a scalar or an array of cnt values) as an array of cnt values."""
    a = np.asarray(v, dtype=dtype)
    return a if a.ndim else np.full(cnt, a, dtype=dtype)


def _blocks(nel):
    return (np.asarray(nel, dtype=np.int64) + _lib.APPLY_BLOCK - 1) // _lib.APPLY_BLOCK


def _starts(nblocks):
    nblocks = np.asarray(nblocks, dtype=np.int64)
    st = np.zeros(len(nblocks), dtype=np.int64)
    if len(nblocks):
        st[1:] = np.cumsum(nblocks)[:-1]
    return st, int(nblocks.sum())


# --------------------------------------------------------------------------
# grouped launches from numpy tables
# --------------------------------------------------------------------------
def gemm_table(A, B, C, m, n, k, lda, ldb, ldc, ta=0, tb=0, alpha=1.0, beta=0.0, tri=0, mixed=False):
    """C_i = alpha op(A_i) op(B_i) + beta C_i for every i, one launch.  All
    arguments are arrays (or scalars broadcast) over the problems.  mixed:
    the products on the FP32 matrix cores (FP64 operands and result)."""
    m = np.asarray(m, dtype=np.int64)
    cnt = len(m)
    if cnt == 0:
        return
    n, k = _vec(n, np.int64, cnt), _vec(k, np.int64, cnt)
    live = (m > 0) & (n > 0)
    if not live.all():
        sel = np.nonzero(live)[0]
        return gemm_table(*(_vec(v, np.uint64, cnt)[sel] for v in (A, B, C)), m[sel], n[sel], k[sel],
                          *(_vec(v, np.int64, cnt)[sel] for v in (lda, ldb, ldc, ta, tb)),
                          _vec(alpha, np.float64, cnt)[sel], _vec(beta, np.float64, cnt)[sel],
                          _vec(tri, np.int64, cnt)[sel], mixed)
    # one tile configuration for the launch, from its shape class (autotune)
    tile = autotune.tile(m, n, k, 2 if isinstance(mixed, str) else int(bool(mixed)))
    bm = 128 if tile else 64
    tn = (n + bm - 1) // bm
    tm = (m + bm - 1) // bm
    st, total = _starts(tm * tn)
    # field assignment broadcasts scalars and casts arrays to the field type
    it = np.zeros(cnt, dtype=_lib.GEMM_ITEM)
    it["A"], it["B"], it["C"] = A, B, C
    it["m"], it["n"], it["k"] = m, n, k
    it["lda"], it["ldb"], it["ldc"] = lda, ldb, ldc
    it["alpha"], it["beta"] = alpha, beta
    it["transA"], it["transB"] = ta, tb
    it["block_start"], it["tiles_n"] = st, tn
    it["tri"] = tri
    dev = ops._upload(it)
    if ops._profiling[0]:
        call("syn_prof_note_flops", float((2.0 * m * n * k).sum()))
    call(ops.gemm_fn(mixed), dev.data_ptr(), cnt, total, tile, ops._stream())


def _launch_blocks(fn, it, nel):
    """Fill block_start from per-item element counts and launch fn."""
    if len(it) == 0:
        return
    st, total = _starts(_blocks(nel))
    it["block_start"] = st
    if total == 0:
        return
    dev = ops._upload(it)
    call(fn, dev.data_ptr(), len(it), total, ops._stream())


def eye_table(x, n):
    it = np.zeros(len(x), dtype=_lib.EYE_ITEM)
    it["x"], it["n"], it["ld"] = x, n, n
    _launch_blocks("syn_eye_f64", it, np.asarray(n, dtype=np.int64) ** 2)


def copy_table(src, dst, n):
    it = np.zeros(len(src), dtype=_lib.COPY_ITEM)
    it["src"], it["dst"], it["n"] = src, dst, n
    _launch_blocks("syn_copy_f64", it, n)


def rowscale_table(x, rows, cols, ld, s, mode):
    it = np.zeros(len(x), dtype=_lib.ROWSCALE_ITEM)
    it["x"], it["rows"], it["cols"], it["ld"], it["s"], it["mode"] = x, rows, cols, ld, s, mode
    _launch_blocks("syn_rowscale_f64", it, np.asarray(rows, dtype=np.int64) * np.asarray(cols, dtype=np.int64))


def vecnorm_table(x, rows, cols, ld, axis, out):
    """out_i = column norms (axis 0) or row norms (axis 1) of every
    row-major rows x cols matrix x_i (leading dimension ld)."""
    cnt = len(x)
    if cnt == 0:
        return
    rows, cols = (np.asarray(v, dtype=np.int64) for v in (rows, cols))
    axis = np.broadcast_to(np.asarray(axis, dtype=np.int32), (cnt,))
    it = np.zeros(cnt, dtype=_lib.VECNORM_ITEM)
    it["x"], it["rows"], it["cols"], it["ld"], it["axis"], it["out"] = x, rows, cols, ld, axis, out
    st, total = _starts(np.where(axis == 0, (cols + 63) // 64, (rows + 3) // 4))
    it["block_start"] = st
    if total:
        dev = ops._upload(it)
        call("syn_vecnorms_f64", dev.data_ptr(), cnt, total, ops._stream())


def apply_table(x, o, p, chil, din, chir, dl, dout, dr):
    it = np.zeros(len(x), dtype=_lib.APPLY_ITEM)
    it["x"], it["o"], it["p"] = x, o, p
    it["chil"], it["din"], it["chir"], it["dl"], it["dout"], it["dr"] = chil, din, chir, dl, dout, dr
    nel = (np.asarray(chil) * np.asarray(dl) * np.asarray(dout) * np.asarray(chir) * np.asarray(dr)).astype(np.int64)
    if ops._profiling[0] and len(it):
        call("syn_prof_note_flops", float(8.0 * nel.sum()))
    _launch_blocks("syn_mpo_apply_f64", it, nel)


# LDS classes (doubles) of the grouped Householder QR: 16, 48 and 96 KB;
# matrices above the last work from their workspace
QR_LDS_CLASSES = (2048, 6144, 12288)


def qr_table(a, m, n, q, work, tau):
    """Householder QR of every row-major m x n (m > n) matrix: q (m x n) gets
    the thin orthogonal factor."""
    it = np.zeros(len(a), dtype=_lib.QR_ITEM)
    it["a"], it["rs"], it["cs"] = a, n, 1
    it["m"], it["n"], it["keep"] = m, n, n
    it["q"], it["ldq"], it["q_trans"] = q, n, 0
    it["work"], it["tau"] = work, tau
    # one launch per LDS class of the matrices, each with the dynamic LDS its
    # largest matrix needs: small matrices do not hold a compute unit's LDS
    # for the large ones
    need = np.asarray(m, dtype=np.int64) * np.asarray(n, dtype=np.int64)
    cls = np.searchsorted(QR_LDS_CLASSES, need, side="left")
    for c in range(len(QR_LDS_CLASSES) + 1):
        sel = np.nonzero(cls == c)[0]
        if len(sel) == 0:
            continue
        dev = ops._upload(it[sel])
        lds = int(need[sel].max()) if c < len(QR_LDS_CLASSES) else 0
        call("syn_qr_sized_f64", dev.data_ptr(), len(sel), lds, ops._stream())


# size classes (matrix order) of the LDS Jacobi launches
JACOBI_LDS_CLASSES = (16, 32, 64)


def jacobi_table(x, g, rt, s):
    """LDS two-sided Jacobi of every symmetric g x g Gram matrix (g <= 64,
    one work-group each); x is read only."""
    it = np.zeros(len(x), dtype=_lib.JACOBI_ITEM)
    it["x"], it["ldx"], it["nv"], it["len"] = x, g, g, g
    it["rt"], it["ldr"], it["s"], it["sweeps"] = rt, g, s, 0
    it["gram"] = 1
    # one launch per size class: the launch's LDS follows its largest matrix,
    # so small matrices launched apart keep more work-groups on a compute unit
    g = np.asarray(g, dtype=np.int64)
    if len(g) and g.max() > JACOBI_LDS_CLASSES[-1]:
        raise ValueError(f"LDS Jacobi takes order <= {JACOBI_LDS_CLASSES[-1]}, got {g.max()}")
    cls = np.searchsorted(JACOBI_LDS_CLASSES, g, side="left")
    for c in range(len(JACOBI_LDS_CLASSES)):
        sel = np.nonzero(cls == c)[0]
        if len(sel) == 0:
            continue
        dev = ops._upload(it[sel])
        call("syn_jacobi_gram_f64", dev.data_ptr(), len(sel), int(g[sel].max()), TOL_LDS, ops.JACOBI_SWEEPS,
             ops._stream())


def jacobi_blocked_table(x, g, rt, s, gram=1):
    """Block one-sided Jacobi of many larger g x g problems, one batched call."""
    cnt = len(x)
    if cnt == 0:
        return 0
    it = np.zeros(cnt, dtype=_lib.BJAC_ITEM)
    it["x"], it["ldx"], it["nv"], it["len"] = x, g, g, g
    it["rt"], it["ldr"], it["s"] = rt, g, s
    it["gram"] = gram
    g = np.ascontiguousarray(g, dtype=np.int64)
    lib = _lib.load()
    wsz = int(lib.syn_jacobi_blocked_many_workspace(g.ctypes.data_as(ctypes.c_void_p),
                                                    g.ctypes.data_as(ctypes.c_void_p), cnt))
    work = _empty(wsz)
    counters = _empty(cnt, torch.int32, zero=True)
    dev_items = _empty(it.nbytes, torch.uint8)
    sweeps = ctypes.c_int(0)
    call("syn_jacobi_blocked_many_f64", dev_items.data_ptr(), it.ctypes.data_as(ctypes.c_void_p), cnt,
         work.data_ptr(), counters.data_ptr(), TOL_BLOCKED, ops.JACOBI_SWEEPS, ctypes.byref(sweeps),
         ops._stream())
    return sweeps.value


def diagmax_table(g, n, ld, out, dropped):
    it = np.zeros(len(g), dtype=_lib.DIAGMAX_ITEM)
    it["g"], it["n"], it["ld"], it["out"], it["dropped"] = g, n, ld, out, dropped
    if len(it):
        dev = ops._upload(it)
        call("syn_diagmax_f64", dev.data_ptr(), len(it), ops._stream())


def chol_leaf_table(g, l, x, ld, n, maxdiag, tolrel, dropped):
    it = np.zeros(len(g), dtype=_lib.CHOL_ITEM)
    it["g"], it["ldg"], it["l"], it["ldl"], it["x"], it["ldx"] = g, ld, l, ld, x, ld
    it["n"], it["maxdiag"], it["tolrel"], it["dropped"] = n, maxdiag, tolrel, dropped
    if len(it):
        dev = ops._upload(it)
        call("syn_chol_leaf_f64", dev.data_ptr(), len(it), ops._stream())


CHOL_LEAF = 64


def _chol_inv_many(G, L, X, ld, n, T, maxdiag, tolrel, dropped):
    """Recursive Cholesky with inverse for many problems at once:
    G = L L^T and X = inv(L~) (pivot-drop convention of syn_chol_leaf_f64) on
    the n x n blocks at G/L/X (leading dimension ld, L and X zero on entry):
        (L11, X11) = chol_inv(G11);  L21 = G21 X11^T;  G22 -= L21 L21^T;
        (L22, X22) = chol_inv(G22);  X21 = -X22 (L21 X11).
    Problems of different size share every launch of their level; blocks of
    at most 64 are factored by the leaf kernel.  T: per-problem scratch of at
    least n1 * n2 doubles."""
    leaf = n <= CHOL_LEAF
    if leaf.any():
        chol_leaf_table(G[leaf], L[leaf], X[leaf], ld[leaf], n[leaf], maxdiag[leaf], tolrel[leaf], dropped[leaf])
    nl = ~leaf
    if not nl.any():
        return
    G, L, X, ld, n, T, maxdiag, tolrel, dropped = (v[nl] for v in (G, L, X, ld, n, T, maxdiag, tolrel, dropped))
    n1 = ((n // CHOL_LEAF + 1) // 2) * CHOL_LEAF
    n2 = n - n1
    _chol_inv_many(G, L, X, ld, n1, T, maxdiag, tolrel, dropped)
    e = np.uint64(8)
    down = (n1 * ld).astype(np.uint64) * e            # row offset n1
    diag = (n1 * ld + n1).astype(np.uint64) * e       # block (n1, n1)
    L21 = L + down
    # L21 = G21 X11^T (op(B) = X11^T upper)
    gemm_table(G + down, X, L21, n2, n1, n1, ld, ld, ld, tb=1, tri=8)
    # G22 -= L21 L21^T, lower triangle only
    gemm_table(L21, L21, G + diag, n2, n2, n1, ld, ld, ld, tb=1, alpha=-1.0, beta=1.0, tri=1)
    _chol_inv_many(G + diag, L + diag, X + diag, ld, n2, T, maxdiag, tolrel, dropped)
    # T = L21 X11 (op(B) = X11 lower);  X21 = -X22 T (op(A) = X22 lower)
    gemm_table(L21, X, T, n2, n1, n1, ld, ld, n1, tri=16)
    gemm_table(X + diag, T, X + down, n2, n1, n2, ld, n1, ld, alpha=-1.0, tri=2)


def cholqr_table(A, m, n, lda, hold, mixed=False):
    """One CholeskyQR pass on many tall row-major m x n matrices (A kept):
    G = A^T A, G = L L^T with X = inv(L~), Q = A X^T.  Returns the addresses
    of Q (m x n, ld n) and L (n x n lower, ld n); buffers go to `hold`.
    mixed: Q on the FP32 matrix cores (G and the Cholesky in FP64)."""
    cnt = len(A)
    gt, G = arena(n * n)
    lt, Lp = arena(n * n)
    xt, Xp = arena(n * n)
    lt.zero_()
    xt.zero_()
    # recursion scratch: every level needs at most n1 * n2 <= n^2 / 2 + 4096
    tt, Tp = arena(n * n // 2 + CHOL_LEAF * CHOL_LEAF)
    st, sp = arena(np.ones(cnt, np.int64))
    dt_ = _empty(cnt, torch.int32, zero=True)
    dp = np.uint64(dt_.data_ptr()) + np.uint64(4) * np.arange(cnt, dtype=np.uint64)
    qt, Q = arena(m * n)
    gemm_table(A, A, G, n, n, m, lda, lda, n, ta=1, tri=1)
    diagmax_table(G, n, n, sp, dp)
    tol = np.array([ops.cholqr_tol(int(a), int(b)) for a, b in zip(m, n)])
    _chol_inv_many(G, Lp, Xp, n.copy(), n.copy(), Tp, sp, tol, dp)
    gemm_table(A, Xp, Q, m, n, n, lda, n, n, tb=1, tri=8, mixed=mixed)
    hold += [gt, lt, xt, tt, st, dt_, qt]
    return Q, Lp


def cholqr2_table(A, m, n, lda, hold, mixed=False):
    """Reduced QR A = Q R of many tall matrices by two CholeskyQR passes;
    returns (Q (m x n, ld n), R (n x n upper, ld n)) addresses."""
    Q1, L1 = cholqr_table(A, m, n, lda, hold, mixed)
    Q, L2 = cholqr_table(Q1, m, n, n, hold, mixed)
    rt_, R = arena(n * n)
    # R = R2 R1 = L2^T L1^T, both upper
    gemm_table(L2, L1, R, n, n, n, n, n, n, ta=1, tb=1, tri=4 | 8)
    hold.append(rt_)
    return Q, R


# Orthogonality budget of the left sweep in TT-rounding (plain FP64), per
# chain: the same bound as the single-chain rounding (sum_k ||Q_k^T Q_k - I||_F
# <= 1e-10 keeps every singular value within 5e-11 relative).
REORTH_BUDGET = 1e-10


def orth_estimate_table(Q, m, n, hold):
    """Randomised ||Q_i^T Q_i - I||_F of many row-major m x n Q (ld n), from
    ops.REORTH_PROBES fixed-seed Gaussian columns (the first n_i rows of one
    shared probe matrix): ||Q^T Q X - X||_F / sqrt(p), the products on the
    device, the difference and norm of the small n x p results on the host."""
    p = ops.REORTH_PROBES
    nmax = int(np.max(n))
    xh = np.random.default_rng(0x5EED).standard_normal((nmax, p))
    X = _empty(nmax * p)
    X.copy_(torch.from_numpy(xh.reshape(-1)))
    xp = np.uint64(X.data_ptr())
    yt, Y = arena(m * p)
    wt, W = arena(n * p)
    gemm_table(Q, xp, Y, m, p, n, n, p, p)                 # Y = Q X
    gemm_table(Q, Y, W, n, p, m, n, p, p, ta=1)            # W = Q^T Y
    hold += [yt, wt, X]
    off = (W - W[0]) // 8
    wh = wt.cpu().numpy()
    return np.array([np.linalg.norm(wh[o:o + int(ni) * p].reshape(int(ni), p) - xh[:int(ni)]) / p ** 0.5
                     for o, ni in zip(off.astype(np.int64), n)])


def jacobi_gram_blocked_table(x, g, rt, s, tol=None):
    """Two-sided block Jacobi of many symmetric g x g Gram matrices (g > 16),
    one batched call; x is diagonalised in place.  tol: rotation threshold
    (default TOL_BLOCKED)."""
    cnt = len(x)
    if cnt == 0:
        return 0
    it = np.zeros(cnt, dtype=_lib.GBJ_ITEM)
    it["g"], it["ldg"], it["rt"], it["ldr"], it["s"], it["nv"] = x, g, rt, g, s, g
    g = np.ascontiguousarray(g, dtype=np.int64)
    lib = _lib.load()
    wsz = int(lib.syn_jacobi_gram_blocked_workspace(g.ctypes.data_as(ctypes.c_void_p), cnt))
    work = _empty(wsz)
    counters = _empty(cnt, torch.int32, zero=True)
    dev_items = _empty(it.nbytes, torch.uint8)
    sweeps = ctypes.c_int(0)
    call("syn_jacobi_gram_blocked_f64", dev_items.data_ptr(), it.ctypes.data_as(ctypes.c_void_p), cnt,
         work.data_ptr(), counters.data_ptr(), TOL_BLOCKED if tol is None else float(tol), ops.JACOBI_SWEEPS,
         ctypes.byref(sweeps), ops._stream())
    if JACOBI_LOG is not None:
        JACOBI_LOG.append((g.copy(), sweeps.value))
    return sweeps.value


# when a list: every blocked Gram Jacobi call appends (orders, sweeps)
JACOBI_LOG = None


def truncate_table(s, g, chi_max, cutoff, keep, disc):
    it = np.zeros(len(s), dtype=_lib.TRUNC_ITEM)
    it["s"], it["n"], it["chi_max"], it["cutoff"] = s, g, chi_max, cutoff
    it["keep"], it["discarded"] = keep, disc
    if len(it):
        dev = ops._upload(it)
        call("syn_truncate_f64", dev.data_ptr(), len(it), ops._stream())


# --------------------------------------------------------------------------
# the batch
# --------------------------------------------------------------------------
@dataclass
class BondRecord:
    """Per right-sweep step: chains touched, bond index, singular values
    (device addresses into an arena kept alive here), their count, kept rank
    and discarded weight (device)."""
    jobs: np.ndarray
    bond: np.ndarray
    s_ptr: np.ndarray
    g: np.ndarray
    kept: np.ndarray
    disc: torch.Tensor
    hold: tuple


class ChainBatch:
    """J chains of MPS cores on the device (see the module docstring)."""

    def __init__(self, n, bond, phys, ptr, hold):
        self.n = np.asarray(n, dtype=np.int64)
        self.bond = np.asarray(bond, dtype=np.int64)
        self.phys = np.asarray(phys, dtype=np.int64)
        self.ptr = np.asarray(ptr, dtype=np.uint64)
        self.hold = list(hold)

    @property
    def jobs(self):
        return len(self.n)

    @property
    def sites(self):
        return int(self.n.sum())

    # ---- construction ----------------------------------------------------
    @classmethod
    def from_numpy(cls, chains):
        """chains: list of lists of (b_k, d_k, b_k+1) float64 numpy cores."""
        J = len(chains)
        nmax = max((len(c) for c in chains), default=0)
        n = np.array([len(c) for c in chains], dtype=np.int64)
        bond = np.zeros((J, nmax + 1), dtype=np.int64)
        phys = np.zeros((J, nmax), dtype=np.int64)
        flat, sizes = [], []
        for j, c in enumerate(chains):
            for k, x in enumerate(c):
                if x.ndim != 3 or (k and x.shape[0] != c[k - 1].shape[2]):
                    raise ValueError(f"chain {j}: core {k} of shape {x.shape} does not continue the chain")
                bond[j, k], phys[j, k] = x.shape[0], x.shape[1]
                flat.append(np.ascontiguousarray(x, dtype=np.float64).reshape(-1))
                sizes.append(x.size)
            if len(c):
                bond[j, len(c)] = c[-1].shape[2]
        host_sizes = _round_up(np.maximum(np.asarray(sizes, dtype=np.int64), 1))
        buf = np.zeros(int(host_sizes.sum()) or 1, dtype=np.float64)
        off = np.zeros(len(sizes), dtype=np.int64)
        if len(sizes):
            off[1:] = np.cumsum(host_sizes)[:-1]
        for o, f in zip(off, flat):
            buf[o:o + f.size] = f
        t = _h2d(torch.from_numpy(buf))
        ptr = np.zeros((J, nmax), dtype=np.uint64)
        base = np.uint64(t.data_ptr())
        i = 0
        for j, c in enumerate(chains):
            for k in range(len(c)):
                ptr[j, k] = base + np.uint64(8 * off[i])
                i += 1
        return cls(n, bond, phys, ptr, [t])

    @classmethod
    def gaussian(cls, n, chi, d=2, seed=0):
        """Chains of n[j] sites with bonds (1, chi_j, ..., chi_j, 1) and
        cores ~ N(0, 1/chi_j), drawn on the device (torch's generator, seeded)
        so that thousands of chains need no host copy.  The arena holds the
        chains grouped by chi (one scaling per group); the tables keep the
        given chain order."""
        n = np.asarray(n, dtype=np.int64)
        chi = np.asarray(chi, dtype=np.int64)
        J = len(n)
        nmax = int(n.max()) if J else 0
        k = np.arange(nmax + 1)[None, :]
        bond = np.where(k == 0, 1, np.where(k < n[:, None], chi[:, None], np.where(k == n[:, None], 1, 0)))
        mask = np.arange(nmax)[None, :] < n[:, None]
        phys = np.where(mask, d, 0).astype(np.int64)
        order = np.argsort(chi, kind="stable")
        jj, kk = np.nonzero(mask[order])
        jj = order[jj]
        sizes = bond[jj, kk] * d * bond[jj, kk + 1]
        t, addr = arena(sizes)
        g = torch.Generator(device=t.device)
        g.manual_seed(int(seed))
        t.normal_(generator=g)
        base = np.uint64(t.data_ptr())
        for c in np.unique(chi):
            sel = np.nonzero(chi[jj] == c)[0]
            lo = int((addr[sel[0]] - base) // np.uint64(8))
            hi = int((addr[sel[-1]] - base) // np.uint64(8)) + int(sizes[sel[-1]])
            t[lo:hi].mul_(1.0 / np.sqrt(float(c)))
        ptr = np.zeros((J, nmax), dtype=np.uint64)
        ptr[jj, kk] = addr
        return cls(n, bond.astype(np.int64), phys, ptr, [t])

    def core_shape(self, j, k):
        return (int(self.bond[j, k]), int(self.phys[j, k]), int(self.bond[j, k + 1]))

    def to_numpy(self, jobs=None):
        """Copy chains back to the host (list of lists of numpy cores)."""
        jobs = range(self.jobs) if jobs is None else jobs
        src, sizes, shapes = [], [], []
        for j in jobs:
            for k in range(int(self.n[j])):
                sh = self.core_shape(j, k)
                src.append(self.ptr[j, k])
                sizes.append(int(np.prod(sh)))
                shapes.append(sh)
        t, dst = arena(sizes)
        copy_table(np.asarray(src, dtype=np.uint64), dst, np.asarray(sizes, dtype=np.int64))
        host = t.cpu().numpy()
        base = np.uint64(t.data_ptr())
        out, i = [], 0
        for j in jobs:
            cores = []
            for k in range(int(self.n[j])):
                o = int((dst[i] - base) // np.uint64(8))
                cores.append(host[o:o + sizes[i]].reshape(shapes[i]).copy())
                i += 1
            out.append(cores)
        return out

    def compact(self):
        """Copy every core into one fresh arena and drop the old ones."""
        J, nmax = self.ptr.shape
        mask = np.arange(nmax)[None, :] < self.n[:, None]
        jj, kk = np.nonzero(mask)
        sizes = self.bond[jj, kk] * self.phys[jj, kk] * self.bond[jj, kk + 1]
        t, dst = arena(sizes)
        copy_table(self.ptr[jj, kk], dst, sizes)
        self.ptr = self.ptr.copy()
        self.ptr[jj, kk] = dst
        self.hold = [t]

    def hbm_bytes(self):
        J, nmax = self.ptr.shape
        mask = np.arange(nmax)[None, :] < self.n[:, None]
        sz = self.bond[:, :-1] * self.phys * self.bond[:, 1:]
        return int(8 * (sz * mask).sum())

    def subset(self, jobs):
        """The chains `jobs` (tables trimmed to the longest of them; the
        arenas stay shared)."""
        jobs = np.asarray(jobs, dtype=np.int64)
        w = int(self.n[jobs].max()) if len(jobs) else 0
        return ChainBatch(self.n[jobs], self.bond[jobs, :w + 1], self.phys[jobs, :w], self.ptr[jobs, :w], self.hold)

    # ---- apply -----------------------------------------------------------
    def apply(self, ops_batch):
        """Materialised MPO x MPS product of every chain with its own MPO:
        P_k[a*Dl+c, o, b*Dr+e] = sum_s X_k[a,s,b] O_k[c,s,o,e], one launch.
        ops_batch is an OperatorBatch of the same chain lengths."""
        if not np.array_equal(ops_batch.n, self.n):
            raise ValueError("MPO and MPS batches differ in chain lengths")
        J = len(self.n)
        nmax = int(self.n.max()) if J else 0
        if self.ptr.shape[1] != nmax:
            self.bond, self.phys, self.ptr = self.bond[:, :nmax + 1], self.phys[:, :nmax], self.ptr[:, :nmax]
        if ops_batch.ptr.shape[1] != nmax:
            ops_batch = OperatorBatch(ops_batch.n, ops_batch.bond[:, :nmax + 1], ops_batch.din[:, :nmax],
                                      ops_batch.dout[:, :nmax], ops_batch.ptr[:, :nmax], ops_batch.hold)
        mask = np.arange(nmax)[None, :] < self.n[:, None]
        jj, kk = np.nonzero(mask)
        din = ops_batch.din[jj, kk]
        if not np.array_equal(din, self.phys[jj, kk]):
            raise ValueError("physical dimension mismatch between MPO inputs and MPS")
        chil, chir = self.bond[jj, kk], self.bond[jj, kk + 1]
        dl, dr = ops_batch.bond[jj, kk], ops_batch.bond[jj, kk + 1]
        dout = ops_batch.dout[jj, kk]
        t, p = arena(chil * dl * dout * chir * dr)
        apply_table(self.ptr[jj, kk], ops_batch.ptr[jj, kk], p, chil, din, chir, dl, dout, dr)
        ptr = np.zeros_like(self.ptr)
        ptr[jj, kk] = p
        bond = self.bond * ops_batch.bond
        phys = np.zeros_like(self.phys)
        phys[jj, kk] = dout
        return ChainBatch(self.n, bond, phys, ptr, [t, ops_batch.hold, self.hold])

    # ---- left sweep --------------------------------------------------------
    def left_orthonormalize(self, mixed=False, reorth_budget=None):
        """Left QR sweep over all chains in lock-step (in place on the tables;
        the input arenas are not written).  mixed: CholeskyQR's Q and the
        absorb products on the FP32 matrix cores.  reorth_budget (FP64 only):
        per chain, the total estimated loss of orthogonality
        sum_k ||Q_k^T Q_k - I||_F the sweep may keep; CholeskyQR skips its
        second pass on a problem while its chain stays within it.  None runs
        both passes everywhere."""
        J, nmax = self.ptr.shape
        ptr, bond = self.ptr.copy(), self.bond.copy()
        hold = list(self.hold)
        spent = np.zeros(J)
        for k in range(nmax - 1):
            a = np.nonzero(self.n - 1 > k)[0]
            if len(a) == 0:
                break
            l, d, r = bond[a, k], self.phys[a, k], bond[a, k + 1]
            m = l * d
            wide = m <= r
            tall = ~wide
            new_r = np.where(wide, m, r)
            # Q factors: identity for wide unfoldings, Householder (narrow
            # matrices that fit one work-group's LDS) or CholeskyQR2 (which
            # brings its own Q buffer) for tall ones
            hh = tall & (r < CHOLQR_MIN_COLS) & (m * r <= QR_LDS_CLASSES[-1])
            cq = tall & ~hh
            qt, q = arena(np.where(cq, 0, m * new_r))
            if wide.any():
                wi = wide.nonzero()[0]
                eye_table(q[wi], m[wi])
            S = ptr[a, k].copy()          # wide: S = L itself (m x r, ld r)
            lds = r.copy()
            sta = np.zeros(len(a), np.int64)   # 1: S holds L1 and the absorb takes L1^T
            if hh.any():
                # Householder, one work-group per matrix; S = Q^T L  (r x r)
                ti = hh.nonzero()[0]
                mt, rt = m[ti], r[ti]
                wk, wp = arena(mt * rt)
                tk, tp = arena(rt)
                qr_table(ptr[a[ti], k], mt, rt, q[ti], wp, tp)
                st, sp = arena(rt * rt)
                gemm_table(q[ti], ptr[a[ti], k], sp, rt, rt, mt, rt, rt, rt, ta=1)
                S[ti] = sp
                hold += [wk, tk, st]
            if cq.any():
                # CholeskyQR2 (grouped GEMMs + batched Cholesky leaves)
                ti = cq.nonzero()[0]
                if mixed or reorth_budget is None:
                    Qc, Rc = cholqr2_table(ptr[a[ti], k], m[ti], r[ti], r[ti], hold, mixed)
                    q[ti] = Qc
                    S[ti] = Rc
                else:
                    # first pass everywhere; the second only where the chain's
                    # orthogonality budget cannot take the first pass's loss
                    mt, rt = m[ti], r[ti]
                    Q1, L1 = cholqr_table(ptr[a[ti], k], mt, rt, rt, hold)
                    est = orth_estimate_table(Q1, mt, rt, hold)
                    keep = est <= reorth_budget - spent[a[ti]]
                    spent[a[ti][keep]] += est[keep]
                    q[ti[keep]] = Q1[keep]
                    S[ti[keep]] = L1[keep]
                    sta[ti[keep]] = 1
                    redo = ~keep
                    if redo.any():
                        mr, rr = mt[redo], rt[redo]
                        Q2, L2 = cholqr_table(Q1[redo], mr, rr, rr, hold)
                        rt_, R = arena(rr * rr)
                        gemm_table(L2, L1[redo], R, rr, rr, rr, rr, rr, rr, ta=1, tb=1, tri=4 | 8)  # R2 R1
                        hold.append(rt_)
                        q[ti[redo]] = Q2
                        S[ti[redo]] = R
            # absorb: next = S @ core_{k+1}  (new_r x d' b'')
            dn, bn = self.phys[a, k + 1], bond[a, k + 2]
            ncols = dn * bn
            nt, npk = arena(new_r * ncols)
            # S from a QR is upper triangular (tri 4); S = L of a wide unfolding is not
            gemm_table(S, ptr[a, k + 1], npk, new_r, ncols, r, lds, ncols, ncols, ta=sta, tri=np.where(wide, 0, 4),
                       mixed=mixed)
            hold += [qt, nt]
            ptr[a, k] = q
            ptr[a, k + 1] = npk
            bond[a, k + 1] = new_r
        out = ChainBatch(self.n, bond, self.phys, ptr, hold)
        return out

    # ---- right sweep -------------------------------------------------------
    def right_svd_sweep(self, chi_max, cutoff=0.0, records=None, mixed=False, jacobi_tol=None):
        """Right-to-left SVD truncation of left-orthonormal chains.
        chi_max, cutoff: scalars or per-chain arrays (chi_max 0 = no cap).
        mixed: the products with the singular vectors on the FP32 matrix
        cores (Gram matrices, Jacobi and truncation stay FP64).  jacobi_tol:
        the blocked Jacobi's rotation threshold (default TOL_BLOCKED)."""
        J, nmax = self.ptr.shape
        chi_max = np.broadcast_to(np.asarray(0 if chi_max is None else chi_max, dtype=np.int64), (J,))
        cutoff = np.broadcast_to(np.asarray(cutoff, dtype=np.float64), (J,))
        ptr, bond = self.ptr.copy(), self.bond.copy()
        hold = list(self.hold)
        for t in range(nmax - 1):
            a = np.nonzero(self.n - 1 - t >= 1)[0]
            if len(a) == 0:
                break
            k = self.n[a] - 1 - t
            l, d, r = bond[a, k], self.phys[a, k], bond[a, k + 1]
            ncol = d * r
            tall = l >= ncol
            g = np.where(tall, ncol, l)
            C = ptr[a, k]
            # Gram matrix of the short side
            gt, gp = arena(g * g)
            gemm_table(C, C, gp, g, g, np.where(tall, l, ncol), ncol, ncol, g,
                       ta=tall.astype(np.int64), tb=(~tall).astype(np.int64))
            rtt, rtp = arena(g * g)
            stt, sp = arena(g)
            small = g <= LDS_JACOBI_MAX
            if small.any():
                jacobi_table(gp[small], g[small], rtp[small], sp[small])
            if (~small).any():
                jacobi_gram_blocked_table(gp[~small], g[~small], rtp[~small], sp[~small], jacobi_tol)
            keep = _empty(len(a), torch.int32, zero=True)
            disc = _empty(len(a), F64, zero=True)
            kb, db = np.uint64(keep.data_ptr()), np.uint64(disc.data_ptr())
            idx = np.arange(len(a), dtype=np.uint64)
            truncate_table(sp, g, chi_max[a], cutoff[a], kb + 4 * idx, db + 8 * idx)
            cc = keep.cpu().numpy().astype(np.int64)
            # new core k and the factor M (l x cc) absorbed into core k-1
            wide = ~tall
            newc = rtp.copy()               # tall: newc = first cc rows of V^T (cc x ncol)
            Mp = np.zeros(len(a), dtype=np.uint64)
            hold_step = [gt, rtt, stt]
            if tall.any():
                ti = tall.nonzero()[0]
                mt, mpp = arena(l[ti] * cc[ti])
                # M = C V = C (V^T)^T  (l x cc)
                gemm_table(C[ti], rtp[ti], mpp, l[ti], cc[ti], ncol[ti], ncol[ti], g[ti], cc[ti], tb=1, mixed=mixed)
                if not mixed:
                    # retained s_i = column norms of U S (exact to eps s_max / s_i)
                    vecnorm_table(mpp, l[ti], cc[ti], cc[ti], 0, sp[ti])
                Mp[ti] = mpp
                hold_step.append(mt)
            if wide.any():
                wi = wide.nonzero()[0]
                nt, npp = arena(cc[wi] * ncol[wi])
                # newc = U^T C  (cc x ncol), then / s ; U^T *= s
                gemm_table(rtp[wi], C[wi], npp, cc[wi], ncol[wi], l[wi], g[wi], ncol[wi], ncol[wi], mixed=mixed)
                if not mixed:
                    # retained s_i = row norms of S V^T, before the scaling uses them
                    vecnorm_table(npp, cc[wi], ncol[wi], ncol[wi], 1, sp[wi])
                rowscale_table(np.concatenate([npp, rtp[wi]]), np.concatenate([cc[wi], cc[wi]]),
                               np.concatenate([ncol[wi], l[wi]]), np.concatenate([ncol[wi], g[wi]]),
                               np.concatenate([sp[wi], sp[wi]]),
                               np.concatenate([np.ones(len(wi), np.int32), np.zeros(len(wi), np.int32)]))
                newc[wi] = npp
                hold_step.append(nt)
            # core k-1 <- core k-1 (rows x l) @ M
            rows = bond[a, k - 1] * self.phys[a, k - 1]
            pt, pp = arena(rows * cc)
            # tall: M row-major (l x cc, ld cc); wide: M = (U S)^T, op(B) = rt rows^T with ld g
            gemm_table(ptr[a, k - 1], np.where(tall, Mp, rtp), pp, rows, cc, l, l,
                       np.where(tall, cc, g), cc, tb=wide.astype(np.int64), mixed=mixed)
            hold_step.append(pt)
            hold += hold_step
            ptr[a, k] = newc
            ptr[a, k - 1] = pp
            bond[a, k] = cc
            if records is not None:
                records.append(BondRecord(a, k, sp, g, cc, disc, (stt,)))
        return ChainBatch(self.n, bond, self.phys, ptr, hold)

    def tt_round(self, chi_max, cutoff=0.0, records=None, compact=True, precision="fp64", jacobi_tol=None):
        """TT-rounding of every chain; precision "fp64" or "mixed" (products
        on the FP32 matrix cores, Gram matrices, Cholesky, Jacobi and
        truncation in FP64, the blocked Jacobi stopped at TOL_BUDGET).
        jacobi_tol overrides the blocked Jacobi's threshold of either."""
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        mixed = precision == "mixed"
        if jacobi_tol is None:
            jacobi_tol = TOL_BUDGET if mixed else TOL_BLOCKED
        out = self.left_orthonormalize(mixed, None if mixed else REORTH_BUDGET).right_svd_sweep(
            chi_max, cutoff, records, mixed, jacobi_tol)
        if compact:
            out.compact()
        return out


class CapturedApplyLeft:
    """The materialised apply and the left sweep of one batch, captured once
    in a HIP graph and replayed per step.

    That part of a layer reads nothing back to the host when the CholeskyQR
    takes both passes (the mixed form always does; the plain form's
    orthogonality budget reads estimates back and is not captured): every
    kernel, item-table copy and arena of the sweep is fixed by the shapes and
    addresses of the inputs.  So one capture replaces the host's work of
    building and uploading the tables of every lock-step step (numpy over
    all chains per launch) by a graph launch.  The inputs X and Op must stay
    where they are (their addresses are in the tables); ``out`` (the
    left-orthonormal batch, in the graph's memory pool) is overwritten by
    every replay, and the right sweep only reads it.  The right sweep stays
    eager: its next shapes depend on the kept ranks decided on the device,
    and the block Jacobi tests convergence from the host."""

    def __init__(self, X, Op, mixed=True):
        self.mixed = bool(mixed)
        self.inputs = (X, Op)
        # one eager run first: kernel attributes and allocator pools settle,
        # and the bytes of its item tables size the staging pool
        ops.count_uploads(True)
        try:
            X.apply(Op).left_orthonormalize(self.mixed, None)
        finally:
            nbytes = ops.count_uploads(False)
        torch.cuda.synchronize()
        self.staging = ops.StagingPool(nbytes + 4096)
        self.graph = torch.cuda.CUDAGraph()
        ops._staging[0] = self.staging
        try:
            with torch.cuda.graph(self.graph):
                self.out = X.apply(Op).left_orthonormalize(self.mixed, None)
        finally:
            ops._staging[0] = None

    def __call__(self):
        """Replay on the current stream; returns the left-orthonormal batch."""
        self.graph.replay()
        return self.out

    def round(self, chi_max, cutoff=0.0, records=None, compact=True, jacobi_tol=None):
        """One layer: replay, then the eager right sweep (TT-rounding of the
        captured batch to chi_max with the cutoff; the blocked Jacobi's
        threshold as in ChainBatch.tt_round)."""
        if jacobi_tol is None:
            jacobi_tol = TOL_BUDGET if self.mixed else TOL_BLOCKED
        out = self().right_svd_sweep(chi_max, cutoff, records, self.mixed, jacobi_tol)
        if compact:
            out.compact()
        return out

    def close(self):
        """Release the graph, its memory pool and the staging pool now
        (an instantiated graph keeps device resources until it is reset)."""
        torch.cuda.synchronize()
        self.graph.reset()
        self.out = None
        self.staging = None
        self.inputs = None


class OperatorBatch:
    """J MPO chains (cores (D_k, din_k, dout_k, D_k+1)) on the device."""

    def __init__(self, n, bond, din, dout, ptr, hold):
        self.n = np.asarray(n, dtype=np.int64)
        self.bond, self.din, self.dout = (np.asarray(v, dtype=np.int64) for v in (bond, din, dout))
        self.ptr = np.asarray(ptr, dtype=np.uint64)
        self.hold = hold

    @classmethod
    def from_numpy(cls, chains):
        J = len(chains)
        nmax = max((len(c) for c in chains), default=0)
        n = np.array([len(c) for c in chains], dtype=np.int64)
        bond = np.zeros((J, nmax + 1), dtype=np.int64)
        din = np.zeros((J, nmax), dtype=np.int64)
        dout = np.zeros((J, nmax), dtype=np.int64)
        sizes, flat = [], []
        for j, c in enumerate(chains):
            for k, o in enumerate(c):
                if o.ndim != 4 or (k and o.shape[0] != c[k - 1].shape[3]):
                    raise ValueError(f"operator {j}: core {k} of shape {o.shape} does not continue the chain")
                bond[j, k], din[j, k], dout[j, k] = o.shape[0], o.shape[1], o.shape[2]
                flat.append(np.ascontiguousarray(o, dtype=np.float64).reshape(-1))
                sizes.append(o.size)
            if len(c):
                bond[j, len(c)] = c[-1].shape[3]
        padded = _round_up(np.maximum(np.asarray(sizes, dtype=np.int64), 1))
        buf = np.zeros(int(padded.sum()) or 1, dtype=np.float64)
        off = np.zeros(len(sizes), dtype=np.int64)
        if len(sizes):
            off[1:] = np.cumsum(padded)[:-1]
        for o, f in zip(off, flat):
            buf[o:o + f.size] = f
        t = _h2d(torch.from_numpy(buf))
        ptr = np.zeros((J, nmax), dtype=np.uint64)
        base = np.uint64(t.data_ptr())
        i = 0
        for j, c in enumerate(chains):
            for k in range(len(c)):
                ptr[j, k] = base + np.uint64(8 * off[i])
                i += 1
        return cls(n, bond, din, dout, ptr, [t])


def gate_pair_cores(U, d=2):
    """Operator-Schmidt split of two-site gates into MPO core pairs.

    U: (P, d*d, d*d) gates acting as y = U x on the pair.  The MPO cores
    store O[in, out] = U[out, in], the convention under which the product
    contracts the MPS physical index with the MPO input leg
    (tensor/matrix_product_operator.py:184), so applying them gives U x.
    O[i0,i1,o0,o1] regrouped as M[(i0,o0),(i1,o1)] = u diag(s) v^T is split
    into A = u sqrt(s) (1, d, d, d*d) and B = sqrt(s) v^T (d*d, d, d, 1)."""
    U = np.asarray(U, dtype=np.float64)
    P = U.shape[0]
    O = np.swapaxes(U, 1, 2).reshape(P, d, d, d, d)
    M = O.transpose(0, 1, 3, 2, 4).reshape(P, d * d, d * d)
    u, sv, vt = np.linalg.svd(M)
    sq = np.sqrt(sv)
    A = (u * sq[:, None, :]).reshape(P, 1, d, d, d * d)
    B = (sq[:, :, None] * vt).reshape(P, d * d, d, d, 1)
    return A, B


def gate_layer(n, offset, gates, d=2):
    """OperatorBatch of brickwork gate layers: chain j of n[j] sites carries
    gates on the pairs (offset[j] + 2p, offset[j] + 2p + 1) and the identity
    elsewhere.  gates: (sum_j npairs_j, d*d, d*d) in chain order, with
    npairs_j = (n[j] - offset[j]) // 2."""
    n = np.asarray(n, dtype=np.int64)
    offset = np.asarray(offset, dtype=np.int64)
    J = len(n)
    nmax = int(n.max()) if J else 0
    npairs = np.maximum((n - offset) // 2, 0)
    if gates.shape[0] != int(npairs.sum()):
        raise ValueError(f"expected {int(npairs.sum())} gates, got {gates.shape[0]}")
    A, B = gate_pair_cores(gates, d)
    stride = _round_up(d * d * d * d, 8)
    buf = np.zeros((J, max(nmax, 1), stride), dtype=np.float64)
    eye = np.eye(d).reshape(-1)
    mask = np.arange(nmax)[None, :] < n[:, None]
    buf[:, :nmax, : d * d][mask] = eye
    jp = np.repeat(np.arange(J), npairs)
    first = np.concatenate([np.arange(c) for c in npairs]) if len(npairs) else np.zeros(0, np.int64)
    sa = offset[jp] + 2 * first.astype(np.int64)
    buf[jp, sa, : d ** 4] = A.reshape(-1, d ** 4)
    buf[jp, sa + 1, : d ** 4] = B.reshape(-1, d ** 4)
    bond = np.ones((J, nmax + 1), dtype=np.int64)
    bond[jp, sa + 1] = d * d
    dd = np.full((J, nmax), d, dtype=np.int64) * mask
    t = _h2d(torch.from_numpy(buf.reshape(-1)))
    base = np.uint64(t.data_ptr())
    ptr = base + (8 * stride * (np.arange(J)[:, None] * max(nmax, 1) + np.arange(nmax)[None, :])).astype(np.uint64)
    return OperatorBatch(n, bond, dd, dd, ptr, [t])


def product_states(n, d=2, level=0):
    """Chains |level ... level> of n[j] sites: cores (1, d, 1) = e_level."""
    n = np.asarray(n, dtype=np.int64)
    J = len(n)
    nmax = int(n.max()) if J else 0
    stride = ALIGN // 4                     # 8 doubles per core
    buf = np.zeros((J, max(nmax, 1), stride), dtype=np.float64)
    buf[:, :, level] = 1.0
    t = _h2d(torch.from_numpy(buf.reshape(-1)))
    base = np.uint64(t.data_ptr())
    mask = np.arange(nmax)[None, :] < n[:, None]
    ptr = base + (8 * stride * (np.arange(J)[:, None] * max(nmax, 1) + np.arange(nmax)[None, :])).astype(np.uint64)
    bond = np.zeros((J, nmax + 1), dtype=np.int64)
    bond[np.arange(nmax + 1)[None, :] <= n[:, None]] = 1
    return ChainBatch(n, bond, np.full((J, nmax), d, dtype=np.int64) * mask, ptr, [t])


def concat(batches):
    """One ChainBatch holding the chains of several batches (tables only)."""
    batches = [b for b in batches if b.jobs]
    if not batches:
        return ChainBatch(np.zeros(0, np.int64), np.zeros((0, 1), np.int64), np.zeros((0, 0), np.int64),
                          np.zeros((0, 0), np.uint64), [])
    nmax = max(b.ptr.shape[1] for b in batches)

    def pad(x, w):
        return np.pad(x, ((0, 0), (0, w - x.shape[1])))
    return ChainBatch(np.concatenate([b.n for b in batches]),
                      np.concatenate([pad(b.bond, nmax + 1) for b in batches]),
                      np.concatenate([pad(b.phys, nmax) for b in batches]),
                      np.concatenate([pad(b.ptr, nmax) for b in batches]),
                      [b.hold for b in batches])


def overlap(A: ChainBatch, B: ChainBatch):
    """<A_j|B_j> for every chain pair (bilinear, no conjugation, as the
    reference's | at tensor/matrix_product_state.py:116-129), by lock-step
    transfer matrices: E <- (E^T A_k)^T-regrouped times B_k, two grouped
    GEMMs per site.  Returns a float64 device tensor (J,)."""
    if not np.array_equal(A.n, B.n) or not np.array_equal(A.phys, B.phys):
        raise ValueError("chain lengths or physical dimensions differ")
    J, nmax = A.ptr.shape
    if J == 0:
        return _empty(1)[:0]
    ends = np.concatenate([A.bond[:, 0], B.bond[:, 0], A.bond[np.arange(J), A.n], B.bond[np.arange(J), B.n]])
    if (ends != 1).any():
        raise ValueError("overlap needs edge bonds of dimension 1")
    hold = []
    et, e = arena(np.ones(J, np.int64))
    et.fill_(1.0)
    hold.append(et)
    for k in range(nmax):
        a = np.nonzero(A.n > k)[0]
        la, lb = A.bond[a, k], B.bond[a, k]
        s = A.phys[a, k]
        ra, rb = A.bond[a, k + 1], B.bond[a, k + 1]
        # Z = E^T A_k  (lb x s*ra)
        zt, z = arena(lb * s * ra)
        gemm_table(e[a], A.ptr[a, k], z, lb, s * ra, la, lb, s * ra, s * ra, ta=1)
        # E' = Z'^T B_k' with Z' (lb*s x ra), B' (lb*s x rb) -> (ra x rb)
        nt, ne = arena(ra * rb)
        gemm_table(z, B.ptr[a, k], ne, ra, rb, lb * s, ra, rb, rb, ta=1)
        e = e.copy()
        e[a] = ne
        hold += [zt, nt]
    out = _empty(J)
    copy_table(e, np.uint64(out.data_ptr()) + 8 * np.arange(J, dtype=np.uint64), np.ones(J, np.int64))
    return out


def apply_and_round(mps: ChainBatch, mpo: OperatorBatch, chi_max, cutoff=0.0, records=None, precision="fp64",
                    jacobi_tol=None):
    """One layer for every chain: materialised apply, then TT-rounding."""
    return mps.apply(mpo).tt_round(chi_max, cutoff, records, precision=precision, jacobi_tol=jacobi_tol)
