"""TT-rounding oracle (NumPy float64) -- the SVD form of ``>>``.

Test infrastructure only (see oracle/__init__.py).  The reference has no SVD
rounding on its live path (its ``>>`` is QR truncation,
tensor/matrix_product_state.py:432-468; SVD truncation exists only in dead code
such as trash/mpo.py:59-135), so this is the textbook algorithm of
I. Oseledets, "Tensor-Train Decomposition", SIAM J. Sci. Comput. 33 (2011),
Algorithm 2 (TT-rounding):

1. left-to-right sweep of reduced QR factorisations of the left unfoldings,
   absorbing R into the next core (left-orthonormal form);
2. right-to-left sweep of SVDs of the right unfoldings; at bond k keep
   ``c_k`` singular triplets, store V^T as core k and absorb U*S into core k-1.

Truncation rule (stated because the reference has none): keep the smallest
``c`` with  sum_{i>=c} s_i^2 <= cutoff * sum_i s_i^2, then ``c = min(c, chi_max)``
and ``c >= 1``.  The discarded weight of the bond is w_k = sum_{i>=c} s_i^2 of
the *current* (already partially truncated) state, so for unitary-free
sequential projections ||psi||^2 - ||psi'||^2 = sum_k w_k exactly.
"""
from __future__ import annotations

import numpy as np

from .reference import apply_site, overlap


def keep_count(s, chi_max=None, cutoff=0.0):
    """Number of singular values to keep (rule in the module docstring)."""
    s = np.asarray(s, dtype=np.float64)
    if s.size == 0:
        return 0
    w = s * s
    total = w.sum()
    # tail[c] = sum_{i>=c} w_i
    tail = np.concatenate([np.cumsum(w[::-1])[::-1], [0.0]])
    c = s.size
    if cutoff > 0.0:
        ok = np.nonzero(tail <= cutoff * total)[0]
        c = int(ok[0]) if ok.size else s.size
    if chi_max is not None:
        c = min(c, int(chi_max))
    return max(c, 1)


def left_orthonormalize(sites):
    """Left QR sweep (reduced QR); returns new cores, the last one carries the norm."""
    out = [np.array(s, dtype=np.float64) for s in sites]
    for k in range(len(out) - 1):
        l, d, r = out[k].shape
        Q, R = np.linalg.qr(out[k].reshape(l * d, r))
        out[k] = Q.reshape(l, d, Q.shape[1])
        nxt = out[k + 1]
        out[k + 1] = (R @ nxt.reshape(nxt.shape[0], -1)).reshape(R.shape[0], nxt.shape[1], nxt.shape[2])
    return out


def tt_round(sites, chi_max=None, cutoff=0.0):
    """TT-rounding of an MPS given by float64 cores (l, d, r).

    Returns ``(cores, svals, kept, discarded)`` where for bond k (between core
    k-1 and core k, k = 1..N-1) ``svals[k]`` holds all singular values of the
    right unfolding at that step, ``kept[k]`` the retained count and
    ``discarded[k]`` the discarded weight.  Index 0 is unused (None).
    """
    cores = left_orthonormalize(sites)
    n = len(cores)
    svals = [None] * n
    kept = [None] * n
    disc = [None] * n
    for k in range(n - 1, 0, -1):
        l, d, r = cores[k].shape
        U, S, Vt = np.linalg.svd(cores[k].reshape(l, d * r), full_matrices=False)
        c = keep_count(S, chi_max, cutoff)
        svals[k] = S
        kept[k] = c
        disc[k] = float(np.sum(S[c:] ** 2))
        cores[k] = Vt[:c].reshape(c, d, r)
        prev = cores[k - 1]
        M = U[:, :c] * S[:c]
        cores[k - 1] = (prev.reshape(-1, l) @ M).reshape(prev.shape[0], prev.shape[1], c)
    return cores, svals, kept, disc


def apply_mpo(mpo_sites, mps_sites):
    """Materialised MPO x MPS product cores (no truncation)."""
    return [apply_site(O, X) for O, X in zip(mpo_sites, mps_sites)]


def apply_and_round(mpo_sites, mps_sites, chi_max=None, cutoff=0.0):
    """One apply+round step: tt_round(O @ X)."""
    return tt_round(apply_mpo(mpo_sites, mps_sites), chi_max, cutoff)


def norm2(sites):
    return overlap(sites, sites)
