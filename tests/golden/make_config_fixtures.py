"""Writes the config-shape parity fixtures from the NumPy FP64 oracle.

    python tests/golden/make_config_fixtures.py [cfg1] [cfg2] [cfg4]

Inputs come from config_cases.py (seeded); the outputs are the oracle's
gauge-invariant results (singular values per bond, kept ranks, discarded
weights, squared norms, boundary and product-state contractions, Schmidt
values).  cfg1 takes about two minutes on an 8-core host, cfg2 about half
an hour (two 32768 x 8192 QRs), cfg4 a few minutes.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import config_cases as cc  # noqa: E402
from oracle import rounding as ornd  # noqa: E402
from syngular_amd import workloads as W  # noqa: E402


def interior(name, case):
    t0 = time.time()
    xs, os_ = cc.case_inputs(case)
    ref, sv, kept, disc = ornd.apply_and_round(os_, xs, chi_max=case["chi_max"], cutoff=case["cutoff"])
    us, ws = cc.boundary_vectors(ref[0].shape[0], ref[-1].shape[2])
    np.savez_compressed(os.path.join(HERE, f"{name}_interior.npz"),
                        svals_1=sv[1], svals_2=sv[2], kept=np.array(kept[1:]), disc=np.array(disc[1:]),
                        norm2=float(np.sum(ref[0] ** 2)), probes=cc.boundary_probes(ref, us, ws),
                        right_orth_err=cc.right_orthonormality_error(ref))
    print(f"{name}: kept {kept[1:]} in {time.time() - t0:.0f} s", flush=True)


def circuits():
    t0 = time.time()
    out = {}
    for j in cc.cfg4_jobs():
        cores = [np.eye(2)[0].reshape(1, 2, 1) for _ in range(j["n"])]
        disc = 0.0
        for layer in range(j["layers"]):
            mpo = W.gate_mpo_cores(j["n"], layer % 2, W.layer_gates(j["seed"], j["id"], layer, j["n"]))
            cores, sv, kept, dk = ornd.apply_and_round(mpo, cores, chi_max=j["chi_max"], cutoff=j["cutoff"])
            disc += float(sum(dk[1:]))
        q = j["id"]
        out[f"norm2_{q}"] = ornd.norm2(cores)
        out[f"disc_{q}"] = disc
        out[f"bonds_{q}"] = np.array([c.shape[2] for c in cores[:-1]])
        out[f"schmidt_{q}"] = cc.schmidt_values(cores, j["n"] // 2)
        out[f"overlaps_{q}"] = cc.product_overlaps(cores, cc.product_probes(j["n"]))
        print(f"cfg4 job {q}: n={j['n']} layers={j['layers']} chi_max={j['chi_max']} "
              f"max bond {max(c.shape[2] for c in cores)} ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(os.path.join(HERE, "cfg4_circuits.npz"), **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["cfg1", "cfg2", "cfg4"]
    if "cfg1" in which:
        interior("cfg1", cc.CFG1)
    if "cfg4" in which:
        circuits()
    if "cfg2" in which:
        interior("cfg2", cc.CFG2)
