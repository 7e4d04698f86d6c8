#!/usr/bin/env python3
"""apply+round benchmark of the matrix-product core on one MI355X.

Headline workload (one "step", the JSON `value`): the MPO x MPS product of
an N=32-site chain (d = 2, MPS bond chi = 256, MPO bond D_w = 16, product bond
4096) followed by TT-rounding back to chi = 256 (left CholeskyQR2 sweep,
right Gram + Jacobi SVD sweep), plain FP64, inputs resident in HBM.
Value = sites/s = 32 * steps / elapsed.

Beside the JSON contract fields the line carries:
  roofline      the work kernel with the largest share of GPU time (from an
                untimed census step in which every kernel launch is timed by
                HIP events), its algorithmic rate in the timed steps (events
                recorded by the library on the launch stream around its
                launches only) against the MI355X peak;
  pipeline      the census step's per-kernel launches, milliseconds and
                rates;
  cpu_baseline  the NumPy FP64 oracle on a bounded sample of the same chain
                (its last three sites), timed on this host;
  parity        errors against the oracle on a reduced chain and the
                Pythagoras identity at full size;
  forms         chain_mixed: the same chain in the mixed form (products on
                       the FP32 matrix cores; Gram matrices, Cholesky, Jacobi
                       and truncation in FP64) with its error vs the oracle;
                the many-chain workloads run by the grouped engine:
                batch: 4096 independent chains (N ~ U[8,48], chi ~ U{8..128})
                       with one layer of Haar 2-site gates each, rounded back
                       to chi, all in lock-step grouped launches (FP64 and
                       mixed);
                serve: a Poisson stream of brickwork-circuit jobs (N ~
                       U[16,128], layers ~ U[4,32], chi_max ~ U{64..512},
                       cutoff 1e-10) under the continuous-batching server
                       (FP64 by default, --serve-precision mixed);
                each beside the per-chain FP64 form (one chain at a time
                through the same kernels) on the same jobs, with p50/p99
                latency and the error against the oracle.
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
    ap.add_argument("--batch-jobs", type=int, default=4096)
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-precision", choices=("fp64", "mixed"), default="fp64",
                    help="arithmetic of the served form (the per-chain comparison is always plain FP64); "
                         "the mixed form gains no speed on this Jacobi-bound workload and its FP32 "
                         "accumulation grows the overlap error over long circuits")
    ap.add_argument("--plain-seconds", type=float, default=20.0,
                    help="time budget of the per-chain comparison of the batch form")
    ap.add_argument("--serve-plain-seconds", type=float, default=60.0,
                    help="time budget of the per-chain comparison of the served form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-api", action="store_true", help="skip the timings of the drop-in API paths")
    ap.add_argument("--precision-form", action="store_true",
                    help="only the precision workload (N=64, d=4, chi=1024, D_w=8, cutoff 1e-8): one plain FP64 "
                         "step and one mixed step, their times and the mixed form's error (a few minutes)")
    ap.add_argument("--chain-only", action="store_true", help="only the headline chain (for profiling)")
    ap.add_argument("--cholqr-probe", action="store_true",
                    help="only the copy probe and one CholeskyQR pass on an 8192 x 4096 unfolding (the chain's "
                         "left-sweep shape, whose GEMMs are the dominant kernel), for short counter runs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy (syn_copy_f64) three times: the known-size "
                         "reference that scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: syngular_amd has no multi-GPU path "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE={world}).", file=sys.stderr)
        sys.exit(2)
    if args.chain_only:
        args.no_batch = args.no_serve = args.no_cpu_baseline = args.no_parity = args.no_api = True

    import numpy as np
    import torch

    from syngular_amd import linalg, ops, rounding
    from syngular_amd import workloads as W

    dev = ops.device()
    if args.copy_probe or args.cholqr_probe:
        copy_probe(np, torch)
    if args.precision_form:
        res = precision_form(np, torch)
        print(json.dumps({"precision_form": res}), flush=True)
        return
    if args.cholqr_probe:
        A = torch.randn(8192, 4096, dtype=torch.float64, device=dev) / 64.0
        ops.cholqr(A)
        torch.cuda.synchronize()
        log("cholqr probe: one CholeskyQR pass on 8192 x 4096")
        return
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

    # census step (untimed): every instrumented kernel bracketed by events,
    # for the per-kernel breakdown and the choice of the dominant work kernel
    torch.cuda.synchronize()
    ops.profile_begin()
    t0 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    census_ms = 1e3 * (time.perf_counter() - t0)
    census = ops.profile_end()
    dom = dominant_work_kernel(census)
    log(f"census step (all kernels timed by events): {census_ms:.1f} ms; dominant work kernel {dom}")

    # timed steps: events around the dominant kernel's launches only (two
    # event packets per launch would slow the thousands of short Jacobi and
    # Cholesky launches of a step by a measurable amount)
    torch.cuda.synchronize()
    ops.profile_begin([dom])
    t0 = time.perf_counter()
    for i in range(args.steps):
        rep = rounding.RoundingReport()
        Z = step(rep)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timed = ops.profile_end()
    ms_per_step = 1e3 * elapsed / args.steps
    sites_per_s = n * args.steps / elapsed
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {ms_per_step:.1f} ms/step, {sites_per_s:.2f} sites/s")

    roof, pipeline = roofline(census, timed, dom, args.steps, ms_per_step, census_ms, W, n, chi, dw, d, rep)
    log(f"dominant kernel {roof['kernel']}: {roof['achieved']} {roof['unit']} ({roof['frac']:.1%} of peak)")

    parity = None
    if not args.no_parity:
        parity = parity_checks(np, torch, ops, linalg, rounding, W, O, X, Z, rep)
        log(f"parity: {parity}")
    del Z, rep

    forms = {}
    if not args.chain_only:
        forms["chain_mixed"] = chain_mixed_form(np, torch, args, O, X, sites_per_s)
        log(f"chain mixed form: {forms['chain_mixed']}")
    if not args.no_batch:
        forms["batch"] = batch_form(np, torch, args)
        log(f"batch form: {forms['batch']}")
    if not args.no_serve:
        forms["serve"] = serve_form(np, torch, args)
        log(f"serve form: {forms['serve']}")
    if not args.no_api:
        forms["api_paths"] = api_form(np, torch)
        log(f"api paths: {forms['api_paths']}")

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
        "data": "synthetic (seeded Gaussian cores, Haar-random orthogonal 2-site gates)",
        "config": {"workload": f"MPO x MPS apply + SVD rounding, chain N={n} d={d} chi={chi} D_w={dw} "
                               f"(product bond {chi * dw}), plain FP64",
                   "form": "plain-fp64", "sites": n, "phys_dim": d, "chi": chi, "mpo_bond": dw,
                   "round_to": chi, "cutoff": 0.0},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "pipeline": pipeline,
        "forms": forms or None,
    }
    print(json.dumps(line), flush=True)


# Kernels bound by latency (serial rotation rounds in LDS, 64-column
# Cholesky leaves, one-work-group factorisations): no flop or byte count
# describes them, so they get no roofline; bench.py names the largest of
# them beside the roofline when it takes more GPU time than the dominant
# work kernel.
LATENCY_KERNELS = ("gram_solve_kernel", "chol_diag_kernel", "jacobi_kernel", "qr_kernel", "bjacobi_step_kernel")


def dominant_work_kernel(prof):
    """The kernel with an algorithmic work count (not in LATENCY_KERNELS)
    that took the most event time."""
    work = {k: v for k, v in prof.items() if k not in LATENCY_KERNELS}
    return max(work.items(), key=lambda kv: kv[1][1])[0]


def roofline(census, timed, dom, steps, ms_per_step, census_ms, W, n, chi, dw, d, rep):
    """Dominant work kernel of the chain (the largest share of GPU time among
    kernels with an algorithmic work count, from the census step where every
    kernel is timed): algorithmic work (flops for the GEMMs, bytes for the
    memory-bound kernels) over its event time in the timed steps, where
    only its launches carry events."""
    total_ms = sum(v[1] for v in census.values())
    dl, dms, dwork = timed[dom]
    if dom in BYTE_KERNELS:
        achieved = dwork / (dms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAKS["hbm_gbs"], "unit": "GB/s"}
    else:
        achieved = dwork / (dms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAKS["fp64_mfma_tflops"],
                "unit": "TFLOP/s"}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["traffic"] = measured_traffic(dom)
    roof["kernel"] = dom
    roof["launches_per_step"] = dl / steps
    roof["avg_launch_us"] = round(1e3 * dms / dl, 2)
    roof["share_of_gpu_time"] = round(census[dom][1] / total_ms, 3) if total_ms else None
    lat = {k: v for k, v in census.items() if k in LATENCY_KERNELS}
    if lat:
        lk, (ll, lms, _) = max(lat.items(), key=lambda kv: kv[1][1])
        if lms > census[dom][1]:
            roof["largest_kernel_by_time"] = {"kernel": lk, "share_of_gpu_time": round(lms / total_ms, 3),
                                              "launches_per_step": ll,
                                              "avg_launch_us": round(1e3 * lms / ll, 2),
                                              "bound": "latency (serial rounds in LDS; no flop/byte roofline)"}
    kept = list(rep.kept)
    textbook = W.apply_round_flops(*W.chain_shapes(n, chi, dw), d, kept)
    pipeline = {"textbook_flops_per_step": textbook,
                "textbook_tflops": round(textbook / (ms_per_step * 1e-3) / 1e12, 3),
                "census_step_ms": round(census_ms, 1),
                "census_kernels": {k: {"launches": v[0], "ms": round(v[1], 2),
                                       "rate": round(v[2] / (v[1] * 1e-3) / (1e9 if k in BYTE_KERNELS else 1e12), 3),
                                       "rate_unit": "GB/s" if k in BYTE_KERNELS else "TFLOP/s"}
                                   for k, v in sorted(census.items(), key=lambda kv: -kv[1][1])}}
    return roof, pipeline


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cholqr_8192x4096_gemm_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by a known-size copy), or None
    when no counter run covers it.  The counters cannot be read from inside
    this process; profiles/cholqr_8192x4096_gemm_traffic.md has the commands."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    return rec.get("bytes_per_launch", {}).get(kernel)


# kernels whose profiler work figure is bytes moved (the rest: flops)
BYTE_KERNELS = ("apply_kernel", "gram_update_kernel")


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


def chain_mixed_form(np, torch, args, O, X, plain_sites_per_s):
    """The headline chain in the mixed form (products on the FP32 matrix
    cores; Gram matrices, Cholesky, Jacobi and truncation in FP64), with its
    error against the FP64 oracle on the reduced chain."""
    from oracle import reference as oref
    from oracle import rounding as ornd
    from syngular_amd import linalg, ops, rounding
    from syngular_amd import workloads as W

    n = len(X)
    rounding.apply_and_round(O, X, chi_max=args.chi, precision="mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rounding.apply_and_round(O, X, chi_max=args.chi, precision="mixed")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    xs, os_ = W.chain_inputs(12, 2, 64, 8, seed=11)
    dev = ops.device()
    Xs = [torch.from_numpy(x).to(dev) for x in xs]
    Os = [torch.from_numpy(o).to(dev) for o in os_]
    r = rounding.RoundingReport()
    cores = rounding.apply_and_round(Os, Xs, chi_max=64, report=r, precision="mixed")
    ref, sv, kept, disc = ornd.apply_and_round(os_, xs, chi_max=64)
    n2, n2r = float(linalg.overlap(cores, cores).item()), ornd.norm2(ref)
    ov, ovr = float(linalg.overlap(cores, Xs).item()), oref.overlap(ref, xs)
    err_s = max(float(np.max(np.abs(r.svals[k - 1].cpu().numpy()[: kept[k]] - sv[k][: kept[k]])) / sv[k][0])
                for k in range(1, 12))
    value = n * args.steps / dt
    return {"sites_per_s": round(value, 3), "ms_per_step": round(1e3 * dt / args.steps, 2),
            "speedup_vs_plain": round(value / plain_sites_per_s, 3),
            "error_vs_fp64_oracle": {"reduced_chain": "N=12 d=2 chi=64 D_w=8", "norm2_rel_err": abs(n2 - n2r) / n2r,
                                     "overlap_rel_err": abs(ov - ovr) / float(np.sqrt(n2r * ornd.norm2(xs))),
                                     "svals_max_rel_err": err_s, "kept_match": r.kept == kept[1:]}}


def precision_form(np, torch):
    """The precision workload: N=64 sites, d=4, chi=1024, MPO bond 8
    (product bond 8192), Gaussian cores as in the headline chain, rounded to
    chi=1024 with relative cutoff 1e-8.  One step in plain FP64 and one in the
    mixed form (products on the FP32 matrix cores; Gram matrices, Cholesky,
    Jacobi and truncation in FP64), inputs resident.  The mixed result is
    compared with the plain FP64 result of the same inputs (norm, overlap,
    retained singular values per bond); the plain form itself is compared
    with the NumPy FP64 oracle on a reduced chain of the same d and MPO bond
    (N=8, chi=64), since the oracle cannot hold the full size."""
    from oracle import reference as oref
    from oracle import rounding as ornd
    from syngular_amd import linalg, ops, rounding
    from syngular_amd import workloads as W

    n, d, chi, dw, cut = 64, 4, 1024, 8, 1e-8
    xs, os_ = W.chain_inputs(n, d, chi, dw, seed=64)
    dev = ops.device()
    X = [torch.from_numpy(x).to(dev) for x in xs]
    O = [torch.from_numpy(o).to(dev) for o in os_]
    del xs, os_
    out = {"workload": f"chain N={n} d={d} chi={chi} D_w={dw} (product bond {chi * dw}), round to chi={chi}, "
                       f"relative cutoff {cut:g}", "steps_each": 1}
    reps = {}
    res = {}
    for prec in ("fp64", "mixed"):
        log(f"precision form {prec}: one step of {out['workload']}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = rounding.RoundingReport()
        Z = rounding.apply_and_round(O, X, chi_max=chi, cutoff=cut, report=rep, precision=prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        reps[prec], res[prec] = rep, Z
        out[prec] = {"seconds_per_step": round(dt, 3), "sites_per_s": round(n / dt, 3)}
        log(f"precision form {prec}: {dt:.1f} s per step")
    zp, zm = res["fp64"], res["mixed"]
    n2p = float(linalg.overlap(zp, zp).item())
    n2m = float(linalg.overlap(zm, zm).item())
    ov = float(linalg.overlap(zm, zp).item())
    err_s = 0.0
    for sp, sm, kp in zip(reps["fp64"].svals, reps["mixed"].svals, reps["fp64"].kept):
        a, b = sp.cpu().numpy()[:kp], sm.cpu().numpy()[:kp]
        err_s = max(err_s, float(np.max(np.abs(a - b)) / a[0]))
    out["mixed"]["speedup_vs_fp64"] = round(out["fp64"]["seconds_per_step"] / out["mixed"]["seconds_per_step"], 3)
    out["mixed"]["error_vs_fp64_gpu"] = {"norm2_rel_err": abs(n2m - n2p) / n2p,
                                         "overlap_rel_err": abs(ov - n2p) / n2p,
                                         "svals_max_rel_err": err_s,
                                         "kept_match": reps["fp64"].kept == reps["mixed"].kept}
    out["kept_bonds_fp64"] = [min(reps["fp64"].kept), max(reps["fp64"].kept)]
    del res, zp, zm
    # the plain form against the oracle, reduced chain of the same d and D_w
    xs, os_ = W.chain_inputs(8, d, 64, dw, seed=65)
    Xs = [torch.from_numpy(x).to(dev) for x in xs]
    Os = [torch.from_numpy(o).to(dev) for o in os_]
    r = rounding.RoundingReport()
    cores = rounding.apply_and_round(Os, Xs, chi_max=64, cutoff=cut, report=r)
    ref, sv, kept, _ = ornd.apply_and_round(os_, xs, chi_max=64, cutoff=cut)
    n2, n2r = float(linalg.overlap(cores, cores).item()), ornd.norm2(ref)
    ovg, ovr = float(linalg.overlap(cores, Xs).item()), oref.overlap(ref, xs)
    es = max(float(np.max(np.abs(r.svals[k - 1].cpu().numpy()[: kept[k]] - sv[k][: kept[k]])) / sv[k][0])
             for k in range(1, 8))
    out["fp64_vs_oracle"] = {"reduced_chain": f"N=8 d={d} chi=64 D_w={dw}", "norm2_rel_err": abs(n2 - n2r) / n2r,
                             "overlap_rel_err": abs(ovg - ovr) / float(np.sqrt(n2r * ornd.norm2(xs))),
                             "svals_max_rel_err": es, "kept_match": r.kept == kept[1:]}
    return out


def batch_form(np, torch, args):
    """4096 chains, one gate layer each, rounded back to their chi: the
    grouped engine (all chains in lock-step launches) against the per-chain
    form on a sample of the same jobs; oracle error on a few jobs."""
    from syngular_amd import batch as B
    from syngular_amd import ops, rounding
    from syngular_amd import workloads as W

    t0 = time.perf_counter()
    jobs = W.batching_jobs(args.batch_jobs, seed=3)
    nsites = sum(j["n"] for j in jobs)
    log(f"batch: {len(jobs)} jobs, {nsites} sites, generated in {time.perf_counter() - t0:.1f} s")
    X = B.ChainBatch.from_numpy([j["cores"] for j in jobs])
    Op = B.gate_layer([j["n"] for j in jobs], np.zeros(len(jobs), np.int64),
                      np.concatenate([j["gates"] for j in jobs]))
    chi = np.array([j["chi"] for j in jobs])
    torch.cuda.synchronize()
    B.apply_and_round(X, Op, chi)          # warm-up
    torch.cuda.synchronize()
    res = {}
    for prec in ("fp64", "mixed"):
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            Zp = B.apply_and_round(X, Op, chi, precision=prec)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        res[prec] = (nsites / min(times), min(times), Zp)
        log(f"batch: {prec} {nsites / min(times):.1f} sites/s")
    batched, Z = res["fp64"][0], res["fp64"][2]
    # per-chain form on a strided sample of the same jobs
    dev = ops.device()
    sample = jobs[:: max(1, len(jobs) // 256)]
    done_sites, plain_time, used = 0, 0.0, 0
    for j in sample:
        # inputs go to the device before the clock starts, as for the batch
        Xs = [torch.from_numpy(c).to(dev) for c in j["cores"]]
        Os = [torch.from_numpy(c).to(dev) for c in W.gate_mpo_cores(j["n"], 0, j["gates"])]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rounding.apply_and_round(Os, Xs, chi_max=j["chi"])
        torch.cuda.synchronize()
        plain_time += time.perf_counter() - t1
        done_sites += j["n"]
        used += 1
        if plain_time > args.plain_seconds:
            break
    plain = done_sites / plain_time
    log(f"batch: per-chain form {plain:.1f} sites/s; oracle check of every 512th job")
    # oracle error on a few jobs (every 512th)
    from oracle import reference as oref
    from oracle import rounding as ornd
    errs, errm = [], []
    check = list(range(0, len(jobs), 512))[:8]
    outs = Z.to_numpy(check)
    outm = res["mixed"][2].to_numpy(check)
    for q, c, cm in zip(check, outs, outm):
        j = jobs[q]
        ref = ornd.apply_and_round(W.gate_mpo_cores(j["n"], 0, j["gates"]), j["cores"], chi_max=j["chi"])[0]
        n2r, ovr = ornd.norm2(ref), oref.overlap(ref, j["cores"])
        scale = np.sqrt(n2r * ornd.norm2(j["cores"]))
        errs.append((abs(ornd.norm2(c) - n2r) / n2r, abs(oref.overlap(c, j["cores"]) - ovr) / scale))
        errm.append((abs(ornd.norm2(cm) - n2r) / n2r, abs(oref.overlap(cm, j["cores"]) - ovr) / scale))
    return {"workload": f"{len(jobs)} chains, N~U[8,48], d=2, chi~U{{8,16,32,64,128}}, one layer of Haar "
                        f"O(4) 2-site gates (MPO bond 4) on pairs (0,1),(2,3),..., rounded to chi; plain FP64",
            "sites": nsites, "sites_per_s": round(batched, 1), "seconds": round(min(times), 3),
            "plain_fp64_per_chain_sites_per_s": round(plain, 1),
            "plain_sample": f"{done_sites} sites of {used} strided jobs, one chain at a time",
            "speedup_vs_plain": round(batched / plain, 2),
            "oracle_jobs": len(errs), "norm2_rel_err_max": max(e[0] for e in errs),
            "overlap_rel_err_max": max(e[1] for e in errs),
            "mixed": {"sites_per_s": round(res["mixed"][0], 1), "seconds": round(res["mixed"][1], 3),
                      "speedup_vs_plain": round(res["mixed"][0] / plain, 2),
                      "norm2_rel_err_max": max(e[0] for e in errm),
                      "overlap_rel_err_max": max(e[1] for e in errm)}}


def serve_form(np, torch, args):
    """A Poisson stream through the continuous-batching server against the
    per-chain form on the same jobs (time-boxed), with latency percentiles
    and the error against the oracle on the cheapest jobs."""
    from syngular_amd import ops, rounding, serving
    from syngular_amd import workloads as W

    jobs = W.serving_jobs(args.serve_jobs, args.serve_rate, seed=4)
    prec = args.serve_precision
    srv = serving.Server(serving.DeviceExecutor.lanes(precision=prec), hbm_budget=200 * 2 ** 30,
                         keep_states=False, lanes=serving.chi_lanes)
    # warm-up on a tiny stream (kernel loads, allocator)
    serving.Server(serving.DeviceExecutor(), hbm_budget=2 ** 34).run(
        W.serving_jobs(2, 1e6, seed=99, n_range=(16, 20), layer_range=(2, 2), chis=(64,)))
    torch.cuda.synchronize()
    log(f"serve: {len(jobs)} jobs at {args.serve_rate}/s into the server")
    rep = srv.run(jobs)
    torch.cuda.synchronize()
    log(f"serve: {rep.site_layers} site-layers served in {rep.makespan:.1f} s; per-chain sample next")
    lat = rep.latency_percentiles()
    pyth = max(abs(r.norm2 + r.discarded - 1.0) for r in rep.results)
    # per-chain form: complete circuits of every third job of the same stream,
    # one chain and one layer at a time, time-boxed (a box that closes early
    # leaves the later, heavier layers out, which favours this form)
    dev = ops.device()
    done, t0, boxed = 0, time.perf_counter(), False
    subset = jobs[::3]
    last_log = time.perf_counter()
    for j in subset:
        cores = [torch.from_numpy(np.eye(2)[0].reshape(1, 2, 1).copy()).to(dev) for _ in range(j["n"])]
        for layer in range(j["layers"]):
            mpo = [torch.from_numpy(c).to(dev)
                   for c in W.gate_mpo_cores(j["n"], layer % 2, W.layer_gates(j["seed"], j["id"], layer, j["n"]))]
            cores = rounding.apply_and_round(mpo, cores, chi_max=j["chi_max"], cutoff=j["cutoff"])
            torch.cuda.synchronize()
            done += j["n"]
            if time.perf_counter() - last_log > 30.0:
                last_log = time.perf_counter()
                log(f"serve: per-chain sample at {done} site-layers")
            if time.perf_counter() - t0 > args.serve_plain_seconds:
                boxed = True
                break
        if boxed:
            break
    plain = done / (time.perf_counter() - t0)
    # oracle on the two cheapest jobs (n * layers), full circuits
    from oracle import rounding as ornd
    log(f"serve: per-chain form {plain:.1f} site-layers/s; oracle check of the two cheapest jobs")
    cheap = sorted(jobs, key=lambda j: j["n"] * j["layers"] * j["chi_max"])[:2]
    srv2 = serving.Server(serving.DeviceExecutor(precision=prec), hbm_budget=2 ** 36, keep_states=True)
    got = {r.id: r for r in srv2.run([dict(j, arrival=0.0) for j in cheap]).results}
    errs = []
    for j in cheap:
        cores = [np.eye(2)[0].reshape(1, 2, 1) for _ in range(j["n"])]
        for layer in range(j["layers"]):
            mpo = W.gate_mpo_cores(j["n"], layer % 2, W.layer_gates(j["seed"], j["id"], layer, j["n"]))
            cores = ornd.apply_and_round(mpo, cores, chi_max=j["chi_max"], cutoff=j["cutoff"])[0]
        from oracle import reference as oref
        g = got[j["id"]].cores
        n2r = ornd.norm2(cores)
        errs.append((abs(ornd.norm2(g) - n2r) / n2r, abs(oref.overlap(g, cores) - n2r) / n2r))
    return {"workload": f"Poisson stream of {len(jobs)} jobs at {args.serve_rate}/s: N~U[16,128], layers~U[4,32] "
                        f"brickwork Haar O(4) gates (MPO bond 4), chi_max~U{{64,128,256,512}}, cutoff 1e-10, "
                        f"from |0...0>; {prec} (served, grouped launches)",
            "precision": prec,
            "site_layers": rep.site_layers, "makespan_s": round(rep.makespan, 2),
            "sites_per_s": round(rep.sites_per_s, 1), "p50_latency_s": round(lat["p50"], 2),
            "p99_latency_s": round(lat["p99"], 2), "layer_steps": rep.steps, "peak_resident": rep.peak_resident,
            "admitted_while_running": rep.admitted_during_run,
            "plain_fp64_per_chain_sites_per_s": round(plain, 1),
            "plain_sample": f"{done} site-layers of the complete circuits of jobs {[j['id'] for j in subset]}, "
                            f"one chain and layer at a time" + (f", stopped at the {args.serve_plain_seconds:.0f} s "
                                                                 f"box" if boxed else ""),
            "lanes": "3 lanes by chi_max ({64,128}, 256, 512), one host thread and HIP stream each",
            "speedup_vs_plain": round(rep.sites_per_s / plain, 2),
            "norm_pythagoras_err_max": pyth,
            "oracle_jobs": [f"N={j['n']} layers={j['layers']} chi_max={j['chi_max']}" for j in cheap],
            "norm2_rel_err_max": max(e[0] for e in errs),
            "overlap_rel_err_max": max(e[1] for e in errs)}


def api_form(np, torch, small=False):
    """The drop-in paths beside the headline, each timed on the device
    (mean of a few calls after one warm-up, inputs resident) and as the NumPy
    oracle on this host (one call, the reference's own algorithm), with the
    relative error between the two on the same inputs:
    dense two-site gates (MPS.apply), DMRG environment blocks, the
    tensor-train Linear forward on a batch, MPO.apply on two sites and a
    24-qubit GHZ circuit on the qubit register.  small: reduced sizes, for
    a quick check of this function itself."""
    from oracle import quantum as oq
    from oracle import reference as oref
    from oracle import variational as ovar

    from syngular_amd.core import Linear
    from syngular_amd.quantum import Qbit, gate
    from syngular_amd.tensor import MatrixProductOperator as MPO, MatrixProductState as MPS
    from syngular_amd.variational import left_blocks, right_blocks

    rng = np.random.default_rng(77)

    def sites(dims, bonds, four=False):
        b = [1] + list(bonds) + [1]
        if four:
            return [rng.standard_normal((b[k], d, d, b[k + 1])) / np.sqrt(b[k] * d) for k, d in enumerate(dims)]
        return [rng.standard_normal((b[k], d, b[k + 1])) / np.sqrt(b[k] * d) for k, d in enumerate(dims)]

    def dev_time(f, reps=3):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = f()
        torch.cuda.synchronize()
        return out, 1e3 * (time.perf_counter() - t0) / reps

    def cpu_time(f):
        t0 = time.perf_counter()
        out = f()
        return out, 1e3 * (time.perf_counter() - t0)

    def rel(a, b):
        return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))

    res = {"reduced_sizes": True} if small else {}
    # dense two-site gate in the middle of a chi = 128 chain of 16 sites
    n, chi = (8, 8) if small else (16, 128)
    bonds = [min(chi, 2 ** (k + 1), 2 ** (n - k - 1)) for k in range(n - 1)]
    xs = sites((2,) * n, bonds)
    G = np.linalg.qr(rng.standard_normal((4, 4)))[0].reshape(2, 2, 2, 2)
    X = MPS.from_sites(xs)
    Y, td = dev_time(lambda: X.apply(G, n // 2 - 1))
    Yo, tc = cpu_time(lambda: oref.MPS.from_sites(xs).apply(G, n // 2 - 1))
    res["mps_apply_2site"] = {"workload": "chain N=16 d=2 chi=128, gate on sites 7-8", "device_ms": round(td, 3),
                              "oracle_cpu_ms": round(tc, 3), "rel_err_dense": rel(Y.to_tensor(), Yo.to_tensor())}
    log(f"api mps_apply_2site: {res['mps_apply_2site']}")
    # DMRG environment blocks of a chi = 128 chain with a bond-8 operator
    n, chi = (8, 8) if small else (24, 128)
    bonds = [min(chi, 2 ** (k + 1), 2 ** (n - k - 1)) for k in range(n - 1)]
    xs, ws = sites((2,) * n, bonds), sites((2,) * n, [8] * (n - 1), four=True)
    X, Wm = MPS.from_sites(xs), MPO.from_sites(list(ws))
    (Lg, Rg), td = dev_time(lambda: (left_blocks(Wm, X), right_blocks(Wm, X)))
    (Lo, Ro), tc = cpu_time(lambda: (ovar.left_blocks(ws, xs), ovar.right_blocks(ws, xs)))
    err = max(rel(g.cpu().numpy(), w) for g, w in zip(Lg + Rg, Lo + Ro) if w is not None)
    res["dmrg_blocks"] = {"workload": "N=24 d=2 chi=128 D_w=8, all left and right blocks", "device_ms": round(td, 3),
                          "oracle_cpu_ms": round(tc, 3), "max_rel_err": err}
    log(f"api dmrg_blocks: {res['dmrg_blocks']}")
    # tensor-train Linear forward on a batch of 64 states
    np.random.seed(3)
    layer = Linear(256, 256, core=4, bond=4) if small else Linear(4096, 4096, core=6, bond=8)
    layer.build(None)
    layer.is_built = True
    Wl = layer.trainable_tensor_weights[0]["weight"]
    xs_b = [sites((4,) * 4, (4, 8, 4)) if small else sites((4,) * 6, (4, 16, 16, 16, 4))
            for _ in range(4 if small else 64)]
    Xb = [MPS.from_sites(x) for x in xs_b]
    outs, td = dev_time(lambda: layer(Xb))
    _, td1 = dev_time(lambda: [Wl @ x for x in Xb], reps=1)
    Wo = oref.MPO.from_sites([s.cpu().numpy() for s in Wl.sites])
    wants, tc = cpu_time(lambda: [Wo @ oref.MPS.from_sites(x) for x in xs_b])
    err = max(rel(g.to_tensor(), w.to_tensor()) for g, w in zip(outs[:4], wants[:4]))
    res["linear_forward_batch64"] = {"workload": "Linear 4096->4096, core 6, bond 8, 64 states of bonds (4,16,16,16,4)",
                                     "device_batched_ms": round(td, 3), "device_one_by_one_ms": round(td1, 3),
                                     "oracle_cpu_ms": round(tc, 3), "max_rel_err_first4": err}
    log(f"api linear_forward_batch64: {res['linear_forward_batch64']}")
    # MPO.apply of a two-site operator on an operator chain of bond 16
    ms_ = sites((2,) * 8, [4, 4, 4, 4, 4, 4, 4] if small else [4, 16, 16, 16, 16, 16, 4], four=True)
    op2 = sites((2, 2), [4], four=True)

    def dev_mpo():
        W = MPO.from_sites(list(ms_))
        W.apply(MPO.from_sites(list(op2)), [3, 4])
        return W
    Wd, td = dev_time(dev_mpo)

    def cpu_mpo():
        W = oref.MPO.from_sites(list(ms_))
        W.apply(oref.MPO.from_sites(list(op2)), [3, 4])
        return W
    Wc, tc = cpu_time(cpu_mpo)
    res["mpo_apply_2site"] = {"workload": "operator chain 8 sites d=2 bond 16, 2-site operator on sites 3-4",
                              "device_ms": round(td, 3), "oracle_cpu_ms": round(tc, 3),
                              "max_rel_err": max(rel(g.cpu().numpy(), w) for g, w in zip(Wd.sites, Wc.sites))}
    log(f"api mpo_apply_2site: {res['mpo_apply_2site']}")
    # 24-qubit GHZ on the register
    nq = 6 if small else 24
    ghz = [(gate.H, 0)] + [(gate.CX, i, i + 1) for i in range(nq - 1)]

    def dev_ghz():
        q = Qbit(nq)
        for op in ghz:
            q = q @ op
        return q
    q, td = dev_time(dev_ghz, reps=1)
    qo, tc = cpu_time(lambda: oq.run_circuit(nq, [(oq.H, 0)] + [(oq.CX, i, i + 1) for i in range(nq - 1)]))
    idx = [(0,) * nq, (1,) * nq]
    res["qbit_ghz24"] = {"workload": "24-qubit GHZ (H + 23 CX) on the bond-2 register", "device_ms": round(td, 3),
                         "oracle_cpu_ms": round(tc, 3),
                         "max_abs_err_amplitudes": max(abs(q.state[i] - qo.state[i]) for i in idx)}
    log(f"api qbit_ghz24: {res['qbit_ghz24']}")
    return res


def copy_probe(np, torch):
    """Three launches of the library's copy kernel over 256 MiB (reads 256 MiB,
    writes 256 MiB each): the reference of known traffic for counter runs."""
    from syngular_amd import batch as B
    n = 32 * 2 ** 20                       # doubles = 256 MiB
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        B.copy_table(np.array([src.data_ptr()], np.uint64), np.array([dst.data_ptr()], np.uint64),
                     np.array([n], np.int64))
    torch.cuda.synchronize()
    log(f"copy probe: 3 x syn_copy_f64 of {n * 8} bytes")


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
