#!/usr/bin/env python3
"""apply+round benchmark of the matrix-product core on one MI355X.

Workload (one "step"): the MPO x MPS product of an N=32-site chain
(d = 2, MPS bond chi = 256, MPO bond D_w = 16, product bond 4096) followed by
TT-rounding back to chi = 256 (left CholeskyQR2 sweep, right Jacobi-SVD
sweep), plain FP64, inputs resident in HBM.  Value = sites/s = 32 * steps /
elapsed.

Besides the JSON contract line the script reports:
  roofline      the kernel with the largest share of GPU time in the timed
                region, its algorithmic rate (HIP events recorded by the
                library on the launch stream) against the MI355X peak;
  cpu_baseline  the NumPy FP64 oracle on a bounded sample of the same chain
                (its last three sites), timed on this host;
  parity        errors of the GPU result against the oracle on a reduced
                chain, and the Pythagoras identity ||OX||^2 - ||Z||^2 =
                sum of discarded weights at full size.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"fp64_mfma_tflops": 78.6, "hbm_gbs": 8000.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sites", type=int, default=32)
    ap.add_argument("--chi", type=int, default=256)
    ap.add_argument("--bond-mpo", type=int, default=16)
    ap.add_argument("--phys", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE={world}).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from syngular_amd import linalg, ops, rounding
    from syngular_amd import workloads as W

    dev = ops.device()
    n, d, chi, dw = args.sites, args.phys, args.chi, args.bond_mpo
    xs, os_ = W.chain_inputs(n, d, chi, dw, seed=2024)
    X = [torch.from_numpy(x).to(dev) for x in xs]
    O = [torch.from_numpy(o).to(dev) for o in os_]
    log(f"workload: N={n} d={d} chi={chi} D_w={dw} (product bond {chi * dw}), fp64")

    def step(report=None):
        return rounding.apply_and_round(O, X, chi_max=chi, cutoff=0.0, report=report)

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        log(f"warmup {i}: {time.perf_counter() - t0:.3f} s")

    torch.cuda.synchronize()
    ops.profile_begin()
    t0 = time.perf_counter()
    for i in range(args.steps):
        rep = rounding.RoundingReport()
        Z = step(rep)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ops.profile_end()
    ms_per_step = 1e3 * elapsed / args.steps
    sites_per_s = n * args.steps / elapsed
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {ms_per_step:.1f} ms/step, {sites_per_s:.2f} sites/s")

    # ---- roofline of the dominant kernel ---------------------------------
    total_ms = sum(v[1] for v in prof.values())
    dom, (dl, dms, dfl) = max(prof.items(), key=lambda kv: kv[1][1])
    if dom == "apply_kernel":
        achieved = dfl / (dms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAKS["hbm_gbs"], "unit": "GB/s"}
    else:
        achieved = dfl / (dms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAKS["fp64_mfma_tflops"],
                "unit": "TFLOP/s"}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["traffic"] = None
    roof["kernel"] = dom
    roof["launches_per_step"] = dl / args.steps
    roof["avg_launch_us"] = round(1e3 * dms / dl, 2)
    roof["share_of_gpu_time"] = round(dms / total_ms, 3) if total_ms else None
    kept = list(rep.kept)
    textbook = W.apply_round_flops(*W.chain_shapes(n, chi, dw), d, kept)
    pipeline = {"textbook_flops_per_step": textbook,
                "textbook_tflops": round(textbook / (ms_per_step * 1e-3) / 1e12, 3),
                "kernels": {k: {"launches": v[0], "ms": round(v[1], 2),
                                "rate": round(v[2] / (v[1] * 1e-3) / (1e9 if k == "apply_kernel" else 1e12), 3)}
                            for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}}
    log(f"dominant kernel {dom}: {roof['achieved']} {roof['unit']} ({roof['frac']:.1%} of peak)")

    # ---- parity ----------------------------------------------------------
    parity = None
    if not args.no_parity:
        parity = parity_checks(np, torch, ops, linalg, rounding, W, O, X, Z, rep)
        log(f"parity: {parity}")

    # ---- CPU baseline (oracle on the chain's last three sites) -----------
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(np, xs, os_, chi)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "apply+round sites/s (mixed vs plain FP64) + p50/p99 latency at stated rel err",
        "value": round(sites_per_s, 3),
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian cores)",
        "config": {"workload": f"MPO x MPS apply + SVD rounding, chain N={n} d={d} chi={chi} D_w={dw} "
                               f"(product bond {chi * dw}), plain FP64",
                   "form": "plain-fp64", "sites": n, "phys_dim": d, "chi": chi, "mpo_bond": dw,
                   "round_to": chi, "cutoff": 0.0},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "pipeline": pipeline,
    }
    print(json.dumps(line), flush=True)


def parity_checks(np, torch, ops, linalg, rounding, W, O, X, Z, rep):
    """(1) Reduced chain (N=12, d=2, chi=64, D_w=8) against the NumPy oracle:
    retained singular values per bond (max |ds| / s_max), final norm and the
    overlap <Z|X> (relative to ||Z|| ||X||).  (2) Full size, GPU only:
    Pythagoras ||OX||^2 - ||Z||^2 = sum_k w_k (relative to ||OX||^2)."""
    from oracle import reference as oref
    from oracle import rounding as ornd

    out = {}
    xs, os_ = W.chain_inputs(12, 2, 64, 8, seed=11)
    dev = ops.device()
    Xs = [torch.from_numpy(x).to(dev) for x in xs]
    Os = [torch.from_numpy(o).to(dev) for o in os_]
    r = rounding.RoundingReport()
    cores = rounding.apply_and_round(Os, Xs, chi_max=64, cutoff=0.0, report=r)
    ref, sv, kept, disc = ornd.apply_and_round(os_, xs, chi_max=64)
    err_s = 0.0
    for k in range(1, 12):
        g = r.svals[k - 1].cpu().numpy()[: kept[k]]
        err_s = max(err_s, float(np.max(np.abs(g - sv[k][: kept[k]])) / sv[k][0]))
    n2 = float(linalg.overlap(cores, cores).item())
    n2r = ornd.norm2(ref)
    ov = float(linalg.overlap(cores, Xs).item())
    ovr = oref.overlap(ref, xs)
    out["reduced_chain"] = "N=12 d=2 chi=64 D_w=8"
    out["svals_max_rel_err"] = err_s
    out["norm2_rel_err"] = abs(n2 - n2r) / n2r
    out["overlap_rel_err"] = abs(ov - ovr) / np.sqrt(n2r * ornd.norm2(xs))
    out["kept_match"] = r.kept == kept[1:]
    # full size, GPU only
    P = ops.apply_many(list(zip(O, X)))
    full2 = float(linalg.overlap(P, P).item())
    z2 = float(linalg.overlap(Z, Z).item())
    wsum = float(sum(float(t.item()) for t in rep.discarded))
    out["full_pythagoras_rel_err"] = abs(full2 - z2 - wsum) / full2
    out["full_discarded_weight_rel"] = wsum / full2
    return out


def cpu_baseline(np, xs, os_, chi):
    """NumPy FP64 oracle on the last three sites of the same chain (their
    product cores and the rounding of that 3-site tail), on this host."""
    from oracle import rounding as ornd
    from oracle.reference import apply_site

    threads = os.environ.get("OMP_NUM_THREADS") or str(os.cpu_count())
    n = len(xs)
    tail = slice(n - 3, n)
    t0 = time.perf_counter()
    prod = [apply_site(o, x) for o, x in zip(os_[tail], xs[tail])]
    ornd.tt_round(prod, chi_max=chi)
    dt = time.perf_counter() - t0
    return {"value": round(3 / dt, 4), "unit": "sites/s", "cores": int(threads), "kind": "port",
            "sample": f"NumPy FP64 oracle, apply + TT-rounding of the last 3 sites of the same chain "
                      f"(two 8192x4096 QRs and the right-edge SVDs), {dt:.1f} s"}


if __name__ == "__main__":
    main()
