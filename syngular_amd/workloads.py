"""Synthetic workloads and the algorithmic work model used by the benchmark.

Inputs are seeded NumPy draws (the reference has no datasets on this path):
MPS cores N(0, 1/chi) and MPO cores N(0, 1)/sqrt(D_w * d), literal bond
dimensions (edge bonds larger than the rank are left as they are and the
rounding clips them).

The work model is the textbook materialised apply + TT-rounding count
(Golub & Van Loan flop counts): apply sum_k 2 d_in d_out P_k P_k+1; left
Householder QR sweep F_QR(m, n) plus the absorb GEMM; right SVD sweep
F_SVD(m, n) = 6 M K^2 + 20 K^3 plus the absorb GEMM.  A faster algorithm
is credited with the same count.
"""
from __future__ import annotations

import numpy as np


def chain_shapes(n, chi, bond_mpo):
    b = [1] + [chi] * (n - 1) + [1]
    w = [1] + [bond_mpo] * (n - 1) + [1]
    return b, w


def chain_inputs(n, d, chi, bond_mpo, seed):
    """MPS cores (b_k, d, b_k+1) ~ N(0, 1/chi) and MPO cores
    (w_k, d, d, w_k+1) ~ N(0, 1)/sqrt(D_w d), as float64 NumPy arrays."""
    rng = np.random.default_rng(seed)
    b, w = chain_shapes(n, chi, bond_mpo)
    xs = [rng.standard_normal((b[k], d, b[k + 1])) / np.sqrt(chi) for k in range(n)]
    os_ = [rng.standard_normal((w[k], d, d, w[k + 1])) / np.sqrt(bond_mpo * d) for k in range(n)]
    return xs, os_


def f_qr(m, n):
    return 2.0 * n * n * (m - n / 3.0) if m >= n else 2.0 * m * m * (n - m / 3.0)


def f_svd(m, n):
    big, small = max(m, n), min(m, n)
    return 6.0 * big * small * small + 20.0 * small ** 3


def apply_round_flops(mps_bonds, mpo_bonds, d, kept):
    """Textbook flops of one materialised apply + TT-rounding.

    mps_bonds / mpo_bonds: bond lists including the edge 1s (length N+1);
    kept: realised kept ranks c_1..c_{N-1} of the right sweep."""
    n = len(mps_bonds) - 1
    P = [mps_bonds[k] * mpo_bonds[k] for k in range(n + 1)]
    apply = sum(2.0 * d * d * P[k] * P[k + 1] for k in range(n))
    left = 0.0
    r = [1] * (n + 1)
    for k in range(n - 1):
        m, nn = r[k] * d, P[k + 1]
        r[k + 1] = min(m, nn)
        left += f_qr(m, nn) + 2.0 * r[k + 1] * nn * d * P[k + 2]
    r[n - 1] = min(r[n - 2] * d, P[n - 1]) if n > 1 else 1
    right = 0.0
    c = [1] * (n + 1)
    for k in range(n - 1, 0, -1):
        c[k] = kept[k - 1]
        right += f_svd(r[k], d * c[k + 1]) + 2.0 * (r[k - 1] * d) * r[k] * c[k]
    return apply + left + right
