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


# --------------------------------------------------------------------------
# many-chain workloads: one gate layer per chain (grouped form) and a served
# stream of brickwork circuits
# --------------------------------------------------------------------------
def haar_orthogonal(rng, count, n=4):
    """count Haar-distributed real orthogonal n x n matrices (QR of a Gaussian
    matrix with the signs of diag(R) folded into Q, Mezzadri 2007)."""
    z = rng.standard_normal((count, n, n))
    q, r = np.linalg.qr(z)
    return q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]


def npairs(n, offset):
    return max((n - offset) // 2, 0)


def layer_gates(seed, job, layer, n):
    """Gates of brickwork layer `layer` of chain `job` (pairs start at site
    layer % 2): drawn from a generator keyed by (seed, job, layer), so a
    chain's circuit does not depend on the schedule that runs it."""
    rng = np.random.default_rng([seed, job, layer])
    return haar_orthogonal(rng, npairs(n, layer % 2))


def gate_mpo_cores(n, offset, gates, d=2):
    """The same layer as dense-free MPO cores (list of (D, d, d, D') arrays),
    for the per-chain path and the oracle."""
    from .batch import gate_pair_cores
    A, B = gate_pair_cores(gates, d)
    cores = [np.eye(d).reshape(1, d, d, 1) for _ in range(n)]
    for p in range(npairs(n, offset)):
        cores[offset + 2 * p] = A[p]
        cores[offset + 2 * p + 1] = B[p]
    return cores


def batching_jobs(count, seed, n_range=(8, 48), chis=(8, 16, 32, 64, 128), d=2):
    """count independent chains: N ~ U[n_range], chi ~ U{chis}, bonds
    (1, chi, ..., chi, 1), cores N(0, 1/chi); each gets one layer of Haar
    2-site gates on the pairs (0,1), (2,3), ... and is rounded back to chi.
    Returns a list of dicts with keys n, chi, cores, gates."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    cs = rng.choice(np.asarray(chis), size=count)
    jobs = []
    for j in range(count):
        n, chi = int(ns[j]), int(cs[j])
        b = [1] + [chi] * (n - 1) + [1]
        sizes = [b[k] * d * b[k + 1] for k in range(n)]
        flat = rng.standard_normal(sum(sizes)) / np.sqrt(chi)
        off = np.concatenate([[0], np.cumsum(sizes)])
        cores = [flat[off[k]:off[k + 1]].reshape(b[k], d, b[k + 1]) for k in range(n)]
        gates = haar_orthogonal(rng, npairs(n, 0))
        jobs.append({"n": n, "chi": chi, "cores": cores, "gates": gates})
    return jobs


def batching_layout(count, seed, n_range=(8, 48), chis=(8, 16, 32, 64, 128)):
    """The shapes and gates of the batching workload without its cores (the
    cores are drawn on the device, batch.ChainBatch.gaussian): N ~ U[n_range],
    chi ~ U{chis}, and one layer of Haar 2-site gates on the pairs
    (0,1), (2,3), ... per chain.  Returns (n, chi, gates) with gates stacked
    in chain order."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    cs = rng.choice(np.asarray(chis), size=count)
    gates = haar_orthogonal(rng, int(sum(npairs(int(v), 0) for v in ns)))
    return ns.astype(np.int64), cs.astype(np.int64), gates


def serving_jobs(count, rate, seed, n_range=(16, 128), layer_range=(4, 32), chis=(64, 128, 256, 512),
                 cutoff=1e-10):
    """A Poisson stream of brickwork-circuit jobs: arrival times (s) with
    exponential gaps of mean 1/rate, N ~ U[n_range], layers ~ U[layer_range],
    chi_max ~ U{chis}; every job starts from |0...0>."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=count)
    arrival = np.cumsum(gaps) - gaps[0]
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    ks = rng.integers(layer_range[0], layer_range[1] + 1, size=count)
    cs = rng.choice(np.asarray(chis), size=count)
    return [{"id": j, "arrival": float(arrival[j]), "n": int(ns[j]), "layers": int(ks[j]),
             "chi_max": int(cs[j]), "cutoff": float(cutoff), "seed": int(seed)} for j in range(count)]
