"""Inputs of the parity cases at the benchmark configurations' own shapes,
and the gauge-invariant quantities their fixtures hold.

Shared by ``make_config_fixtures.py`` (which runs the NumPy FP64 oracle on
them once, here, and writes ``*.npz``) and the GPU tests (which rebuild the
same seeded inputs, run the HIP path and compare with the fixture).  The
oracle needs minutes at these shapes on a host CPU, so its outputs are kept
as vectors instead of being recomputed in every test run.

Open-boundary sub-chains reach the interior shapes of the long chains
without the long chains: a chain's left rank grows as d^k from its left
edge, so the 8192 x 4096 (cfg1) and 32768 x 8192 (cfg2) unfoldings of the
left sweep and the 4096 x 512 / 8192 x 4096 right-sweep unfoldings only
appear deep inside.  A three-site piece whose left bond is the interior
bond chi (product bond chi * D_w) and whose MPO closes on the right with
bond 1 (right bond chi) has exactly those unfoldings at both interior
bonds.  Its rounding is compared through quantities that do not depend on
the gauge: the singular values per bond, the kept ranks, the discarded
weights, the squared norm and the tensor contracted with fixed boundary
vectors on both open legs.
"""
from __future__ import annotations

import numpy as np


def interior_chain(seed, chi, bond_mpo, d, sites=3):
    """MPS cores (chi, d, chi) ~ N(0, 1/chi) and MPO cores (D, d, d, D)
    ~ N(0, 1)/sqrt(D d), the last MPO core closing with bond 1."""
    rng = np.random.default_rng(seed)
    xs = [rng.standard_normal((chi, d, chi)) / np.sqrt(chi) for _ in range(sites)]
    os_ = [rng.standard_normal((bond_mpo, d, d, bond_mpo if k < sites - 1 else 1)) / np.sqrt(bond_mpo * d)
           for k in range(sites)]
    return xs, os_


# cfg1: N=32 d=2 chi=256 D_w=16 (product bond 4096), round to 256, no cutoff
CFG1 = dict(seed=2101, chi=256, bond_mpo=16, d=2, chi_max=256, cutoff=0.0)
# cfg2: N=64 d=4 chi=1024 D_w=8 (product bond 8192), round to 1024, cutoff 1e-8
CFG2 = dict(seed=2102, chi=1024, bond_mpo=8, d=4, chi_max=1024, cutoff=1e-8)


def case_inputs(case):
    return interior_chain(case["seed"], case["chi"], case["bond_mpo"], case["d"])


def boundary_vectors(l0, r_end, count=4, seed=7):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((count, l0)), rng.standard_normal((count, r_end))


def boundary_probes(cores, us, ws):
    """For each pair (u, w): the tensor with its left leg contracted with u
    and its right leg with w, flattened (d^N values)."""
    out = []
    for u, w in zip(us, ws):
        t = u.reshape(1, -1)
        for c in cores:
            l, d, r = c.shape
            t = (t @ c.reshape(l, d * r)).reshape(-1, r)
        out.append(t @ w)
    return np.stack(out)


def right_orthonormality_error(cores):
    """max over cores 1.. of max |R R^T - I| of the right unfoldings."""
    err = 0.0
    for c in cores[1:]:
        R = c.reshape(c.shape[0], -1)
        err = max(err, float(np.abs(R @ R.T - np.eye(R.shape[0])).max()))
    return err


# cfg4: brickwork circuits of Haar 2-site gates from |0...0>, cutoff 1e-10,
# one job per chi_max class; enough sites and layers for the middle bonds to
# reach chi_max and stay capped for a few layers (the largest bond of these
# circuits grows about 1.5x per layer under the 1e-10 cutoff).
CFG4_JOBS = [dict(id=0, n=16, layers=10, chi_max=64), dict(id=1, n=18, layers=12, chi_max=128),
             dict(id=2, n=20, layers=13, chi_max=256), dict(id=3, n=24, layers=14, chi_max=512)]
CFG4_SEED = 4104
CFG4_CUTOFF = 1e-10


def cfg4_jobs():
    return [dict(j, arrival=0.0, cutoff=CFG4_CUTOFF, seed=CFG4_SEED) for j in CFG4_JOBS]


def product_probes(n, count=8, seed=9):
    """count random product states: (count, n, 2) local vectors."""
    return np.random.default_rng(seed).standard_normal((count, n, 2))


def product_overlaps(cores, vecs):
    """<v_1 x ... x v_n | psi> for each product state in vecs (count, n, d)."""
    out = []
    for v in vecs:
        e = np.ones((1,))
        for c, vk in zip(cores, v):
            e = e @ np.einsum("lsr,s->lr", c, vk)
        out.append(float(e.reshape(-1)[0]))
    return np.array(out)


def schmidt_values(cores, bond):
    """Schmidt values of a closed chain across the cut before core `bond`:
    left-orthonormalise cores 0..bond-1 and right-orthonormalise cores
    bond..n-1 by QR, then the singular values of the bond matrix."""
    cores = [np.array(c) for c in cores]
    carry = np.ones((1, 1))
    for k in range(bond):
        c = np.einsum("ab,bsr->asr", carry, cores[k])
        l, d, r = c.shape
        q, carry = np.linalg.qr(c.reshape(l * d, r))
    carry_r = np.ones((1, 1))
    for k in range(len(cores) - 1, bond - 1, -1):
        c = np.einsum("lsr,rb->lsb", cores[k], carry_r)
        l, d, r = c.shape
        q, rr = np.linalg.qr(c.reshape(l, d * r).T)
        carry_r = rr.T
    return np.linalg.svd(carry @ carry_r, compute_uv=False)
