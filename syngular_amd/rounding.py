"""SVD rounding of an MPS on the device (an SVD form of ``>>`` added beside the reference's QR form).

The reference's live ``>>`` is QR truncation (tensor/matrix_product_state.py:432-468);
this module adds the textbook TT-rounding (Oseledets 2011, Alg. 2):

1. left sweep: QR of each left unfolding (CholeskyQR2 on the MFMA GEMM for
   96 columns or more, else Householder with one work-group per matrix) and
   R absorbed into the next core with the GEMM; an unfolding with no more
   rows than columns needs no factorisation (Q = I, R = L);
2. right sweep: at bond k the right unfolding C (m x n) of core k is
   decomposed through the Gram matrix of its short side (G = C^T C or C C^T,
   MFMA GEMM) by two-sided Jacobi (in LDS up to 64, block Jacobi beyond),
   the kept rank is chosen on the device (relative discarded weight <=
   cutoff, then <= chi_max, >= 1), and U*S is absorbed into core k-1 with
   the GEMM.

The Gram matrix decides the kept rank and the discarded weight (an
eigenvalue of G carries an absolute error of about eps * s_max^2, so the
tail sum is exact to a few eps of the total).  The retained singular values
themselves are read off the product that carries them in the plain form:
the column norms of C V = U S (tall C) or the row norms of U^T C = S V^T
(wide C), which are exact to about eps * s_max / s_i relative, where the
square root of an eigenvalue of G is only good to eps * (s_max / s_i)^2
(1e-6 at s_i = 1e-5 s_max, the smallest value a 1e-10 cutoff keeps).  The
mixed form keeps the FP64 Gram values: its products round to FP32.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import ops
from .linalg import reshape

F64 = torch.float64


@dataclass
class RoundingReport:
    """Per bond k (between core k-1 and k, k = 1..N-1): all singular values at
    that step, the kept count and the discarded weight sum_{i>=c} s_i^2."""
    svals: list = field(default_factory=list)
    kept: list = field(default_factory=list)
    discarded: list = field(default_factory=list)


# Unfoldings with at least this many columns take the CholeskyQR2 path (all
# MFMA GEMM work); smaller ones the one-work-group Householder kernel.
CHOLQR_MIN_COLS = 96
# Right-sweep Gram matrices up to this size are diagonalised in LDS by one
# work-group (two-sided Jacobi); larger ones by the two-sided block Jacobi
# spread over many work-groups (past 16 the LDS kernel's rounds are bound by
# LDS bandwidth and the block method is faster).
LDS_JACOBI_MAX = 16
# Orthogonality budget of the left sweep inside TT-rounding (plain FP64).  A
# left factor with Q^T Q = I + E scales the bond's singular values by
# (I + E)^(1/2): relative change at most ||E||_2 / 2 <= ||E||_F / 2, and the
# factors of a sweep add up, so a sweep that keeps sum_k ||E_k||_F <= 1e-10
# moves every singular value by at most 5e-11 relative, half the 1e-10
# agreement asked of the plain form.  CholeskyQR's first pass on the
# 8192 x 4096 unfoldings of the chain N=32 chi=256 leaves 0.8-1.6e-11 each.
ROUND_REORTH_BUDGET = 1e-10


# fp64: every product in FP64.  mixed: the products on the FP32 matrix
# cores.  bf16: the contractions (absorb S R, C V, U^T C, P M) on the BF16
# matrix cores and CholeskyQR's Q = A L^-T on the FP32 ones.  In every form
# the Gram matrices, the Cholesky, the Jacobi rotations and the truncation
# stay FP64.
PRECISIONS = ("fp64", "mixed", "bf16")


def _kinds(precision):
    """(compute kind of the QR products, of the contractions)."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}, got {precision!r}")
    return {"fp64": (False, False), "mixed": (True, True), "bf16": (True, "bf16")}[precision]


def left_orthonormalize(cores, own=False, mixed=False, reorth_budget=None, contract=None):
    """Left QR sweep; returns new cores (left-orthonormal except the last).
    With own=False the input cores are not modified; with own=True the
    caller's list is consumed (its entries are replaced as the sweep goes,
    and an entry that is a callable is formed only when the sweep reaches
    it).  mixed: the products
    (CholeskyQR's Q and the absorb) on the FP32 matrix cores, the Gram
    matrices and the Cholesky in FP64.  reorth_budget (FP64 only): total
    loss of orthogonality sum_k ||Q_k^T Q_k - I||_F the sweep may keep;
    CholeskyQR skips its second pass at a site while the running sum of the
    first passes' estimates stays within it (ops.cholqr2_adaptive).  None
    runs both passes everywhere.  contract: the compute kind of the absorb
    product (mixed when None)."""
    if contract is None:
        contract = mixed
    cores = cores if own else list(cores)
    spent = 0.0
    for k in range(len(cores) - 1):
        c = _core(cores, k).contiguous()
        l, mid, r = c.shape[0], c.shape[1:-1], c.shape[-1]
        L = c.reshape(-1, r)
        m = L.shape[0]
        if m <= r:
            # a wide unfolding: Q = I_m and R = L, no factorisation needed
            Q, S, tri = torch.eye(m, dtype=F64, device=L.device), L, 0
        elif r >= CHOLQR_MIN_COLS:
            if reorth_budget is None or mixed:
                Q, S = ops.cholqr2(L, mixed=mixed)
            else:
                Q, S, est = ops.cholqr2_adaptive(L, reorth_budget - spent)
                spent += est
            tri = 4                      # S = R is upper triangular
        else:
            Q, S = _qr(L)
            tri = 4
        nxt = _core(cores, k + 1)
        cores[k] = Q.reshape(l, *mid, Q.shape[1])
        del c, L
        W = ops.gemm(S, nxt.reshape(nxt.shape[0], -1), tri=tri, mixed=contract)
        cores[k + 1] = W.reshape(S.shape[0], *nxt.shape[1:])
    if cores:
        _core(cores, len(cores) - 1)
    return cores


def _core(cores, k):
    """Core k of a sweep's list; an entry that is a callable (a product core
    not formed yet) is formed now and stored in its place."""
    c = cores[k]
    if callable(c):
        c = cores[k] = c()
    return c


def _qr(L):
    m, n = L.shape
    Q = ops.qr_many([(L, n)])[0]
    S = ops.gemm(Q.t(), L)
    return Q, S


def right_svd_sweep(cores, chi_max=None, cutoff=0.0, report=None, mixed=False):
    """Right-to-left SVD truncation of left-orthonormal cores.  mixed: the
    compute kind of the products with the singular vectors (False: FP64,
    True: FP32, "bf16"); the Gram matrix, the Jacobi rotations and the
    truncation stay FP64."""
    n = len(cores)
    for k in range(n - 1, 0, -1):
        c = cores[k]
        l, mid, r = c.shape[0], c.shape[1:-1], c.shape[-1]
        C = c.reshape(l, -1)
        m, ncol = C.shape
        tall = m >= ncol
        if tall:
            G = ops.gemm(C.t(), C)       # (ncol x ncol) = C^T C
        else:
            G = ops.gemm(C, C.t())       # (m x m) = C C^T
        if G.shape[0] <= LDS_JACOBI_MAX:
            rt, s = ops.jacobi_gram_many([G])[0]
        else:
            rt, s, _ = ops.jacobi_gram_blocked(G)
        keep, disc = ops.truncate_many([(s, chi_max, cutoff)])
        cc = int(keep.item())
        prev = cores[k - 1]
        Pl = prev.reshape(-1, l)                       # (l' * mid' x m)
        if tall:
            Vt = rt[:cc]                               # (cc x ncol) = V^T
            M = ops.gemm(C, Vt.t(), mixed=mixed)       # C V = U S   (m x cc)
            newc = Vt
            if not mixed:
                ops.vecnorms_many([(M, 0, s[:cc])])    # s_i = ||(U S)[:, i]||
        else:
            Ut = rt[:cc].contiguous()                  # (cc x m) = U^T
            newc = ops.gemm(Ut, C, mixed=mixed)        # U^T C = S V^T (cc x ncol)
            if not mixed:
                ops.vecnorms_many([(newc, 1, s[:cc])])  # s_i = ||(S V^T)[i, :]||
            ops.rowscale_many([(newc, s, 1)])          # -> V^T
            ops.rowscale_many([(Ut, s, 0)])            # -> (U S)^T
            M = Ut.t()
        cores[k] = reshape(newc, (cc, *mid, r))
        cores[k - 1] = reshape(ops.gemm(Pl, M, mixed=mixed), (*prev.shape[:-1], cc))
        if report is not None:
            report.svals.insert(0, s)
            report.kept.insert(0, cc)
            report.discarded.insert(0, disc)
    return cores


def tt_round(cores, chi_max=None, cutoff=0.0, report=None, own=False, precision="fp64"):
    """TT-rounding of an MPS given as device cores (l, d, r).  The input
    cores are left untouched unless own=True.  precision "fp64" (plain),
    "mixed" (products on the FP32 matrix cores) or "bf16" (contractions on
    the BF16 matrix cores, QR products on the FP32 ones); Gram matrices,
    Cholesky, Jacobi and truncation in FP64 in every form."""
    qr_kind, con = _kinds(precision)
    left = left_orthonormalize(cores, own=own, mixed=qr_kind, reorth_budget=ROUND_REORTH_BUDGET, contract=con)
    return right_svd_sweep(left, chi_max, cutoff, report, mixed=con)


def apply_and_round(mpo_sites, mps_sites, chi_max=None, cutoff=0.0, report=None, precision="fp64"):
    """One apply+round step: MPO x MPS product cores, then tt_round.  Each
    product core is formed when the left sweep reaches it and dropped once
    its QR factor replaces it, so at most two product cores (P_k, P_k+1)
    exist at a time instead of the whole product (133 GB at N=64, d=4,
    chi=1024, D_w=8, against 2.1 GB per core)."""
    prod = [(lambda o=o, x=x: ops.apply_many([(o, x)])[0]) for o, x in zip(mpo_sites, mps_sites)]
    return tt_round(prod, chi_max, cutoff, report, own=True, precision=precision)
